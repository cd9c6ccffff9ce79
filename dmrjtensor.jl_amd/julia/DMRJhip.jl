#  DMRJhip.jl: Julia front-end of libdmrjhip (MI355X two-site DMRG hot path).
#
#  Keeps the reference's API for the two-site path (bakerte/DMRJtensor.jl):
#    dmrg(psi, mpo; m, minm, sweeps, cutoff, silent, goal, method)  algs/DMRG.jl:549
#    twostep / optmps / dmrgcvg / SvNcheck!                         algs/DMRG.jl:362,610,172
#    step3S / alpha_update (dmrg(...; method="3S"))                  algs/DMRG.jl:322,288
#                                                                   algs/optimizeMPS.jl:759
#    Lupdate / Rupdate / makeEnv / expect                           lib/MPO_MPS/Lupdate.jl:17,
#                                                                   Rupdate.jl:17, lib/Env/makeEnv.jl:196,
#                                                                   lib/MPO_MPS/expect.jl:19
#  The tensors live in HBM behind opaque handles; every method below is a
#  `ccall` into include/dmrj.h.  Host arrays are column-major, the same memory
#  order as Julia arrays, so no transposition happens at the boundary.
#
#  With DMRJtensor loaded first (`using DMRJtensor; include("DMRJhip.jl")`):
#    * gpuMPS <: DMRJtensor.regMPS and gpuMPO <: DMRJtensor.regMPO (lib/types.jl:
#      84,99), the environments are Vectors (DMRJtensor's Env includes Array);
#    * dmrg, twostep, Lupdate, Rupdate, makeEnv and expect are methods of
#      DMRJtensor's own generic functions for the GPU types, so user code
#      calling dmrg(psi, mpo; ...) dispatches here when psi and mpo are GPU
#      handles, and nothing is ambiguous;
#    * twostep has the stepfct signature of the reference's optmps
#      (algs/optimizeMPS.jl:805), and dmrg runs the reference's own optmps
#      with it (Envfct = gpu_setEnv, measfct = gpu_startval, cvgfct =
#      DMRJtensor.dmrgcvg).
#  Without DMRJtensor the same names are this module's own functions and dmrg
#  runs the sweep loop of optmps restated below.
#
#  This file needs no package beyond Base.  Point DMRJHIP_LIB at the built
#  shared library (dmrjtensor.jl_amd/lib/libdmrjhip.so by default).
module DMRJhip

const HAS_DMRJ = isdefined(Main, :DMRJtensor)
const _MPSsuper = HAS_DMRJ ? Main.DMRJtensor.regMPS : Any
const _MPOsuper = HAS_DMRJ ? Main.DMRJtensor.regMPO : Any
if HAS_DMRJ
  # new methods below extend the reference's generic functions
  @eval import Main.DMRJtensor: dmrg, twostep, Lupdate, Rupdate, makeEnv, expect
end

export gpuMPS, gpuMPO, gpuEnv, gpuSite, dmrg, makeEnv, Lupdate, Rupdate, expect, twostep, step3S, tdvp, free!,
       applyOps!

const libdmrj = get(ENV, "DMRJHIP_LIB", joinpath(@__DIR__, "..", "lib", "libdmrjhip.so"))

const PREC = Dict("f64" => Cint(0), "bf16x3" => Cint(1), "f16x3" => Cint(2))
const CORRECT = Cint(16)

mutable struct Context
  h::Ptr{Cvoid}
end

function check(ctx::Context, rc::Cint, what::String)
  if rc != 0
    msg = unsafe_string(ccall((:dmrj_last_error, libdmrj), Cstring, (Ptr{Cvoid},), ctx.h))
    error("$what failed ($rc): $msg")
  end
  nothing
end

const _ctx = Ref{Union{Nothing,Context}}(nothing)

function context(device::Integer=0)
  if _ctx[] === nothing
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:dmrj_ctx_create, libdmrj), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
    rc == 0 || error("dmrj_ctx_create failed ($rc): no usable GPU")
    _ctx[] = Context(h[])
  end
  return _ctx[]
end

"device tensor handle (MPS tensor, two-site tensor or environment)"
mutable struct gpuTens
  h::Ptr{Cvoid}
end

function free!(t::gpuTens)
  if t.h != C_NULL
    ctx = context()
    ccall((:dmrj_tensor_free, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, t.h)
    t.h = C_NULL
  end
  nothing
end

"charges as an Int32 [nq, n] matrix (column j = charge of basis state j)"
charges(q::AbstractMatrix{<:Integer}) = Matrix{Int32}(q)

function upload(A::Array{Float64,3}; qL=nothing, qP=nothing, qR=nothing)
  ctx = context()
  nq = qL === nothing ? 0 : size(qL, 1)
  out = Ref{Ptr{Cvoid}}(C_NULL)
  pl = nq == 0 ? C_NULL : pointer(charges(qL))
  pp = nq == 0 ? C_NULL : pointer(charges(qP))
  pr = nq == 0 ? C_NULL : pointer(charges(qR))
  GC.@preserve qL qP qR A begin
    rc = ccall((:dmrj_site_from_dense, libdmrj), Cint,
               (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                Ref{Ptr{Cvoid}}),
               ctx.h, nq, size(A, 1), size(A, 2), size(A, 3), A, pl, pp, pr, out)
  end
  check(ctx, rc, "dmrj_site_from_dense")
  return gpuTens(out[])
end

"ComplexF64 upload: Julia's complex arrays are interleaved (re, im), the layout dmrj_site_from_dense_z takes"
function upload(A::Array{ComplexF64,3}; qL=nothing, qP=nothing, qR=nothing)
  ctx = context()
  nq = qL === nothing ? 0 : size(qL, 1)
  out = Ref{Ptr{Cvoid}}(C_NULL)
  pl = nq == 0 ? C_NULL : pointer(charges(qL))
  pp = nq == 0 ? C_NULL : pointer(charges(qP))
  pr = nq == 0 ? C_NULL : pointer(charges(qR))
  GC.@preserve qL qP qR A begin
    rc = ccall((:dmrj_site_from_dense_z, libdmrj), Cint,
               (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{ComplexF64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                Ref{Ptr{Cvoid}}),
               ctx.h, nq, size(A, 1), size(A, 2), size(A, 3), A, pl, pp, pr, out)
  end
  check(ctx, rc, "dmrj_site_from_dense_z")
  return gpuTens(out[])
end

iscomplex(t::gpuTens) = ccall((:dmrj_tensor_is_complex, libdmrj), Cint, (Ptr{Cvoid},), t.h) == 1

function info(t::gpuTens)
  d = zeros(Int64, 9)
  ccall((:dmrj_tensor_info, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Int64}), t.h, d)
  return d
end

function download(t::gpuTens)
  ctx = context()
  d = info(t)
  dims = d[1] == 1 ? (d[2], d[3], d[4], d[5]) : (d[2], d[3], d[4])
  if iscomplex(t)
    z = zeros(ComplexF64, dims...)
    rc = ccall((:dmrj_tensor_to_dense_z, libdmrj), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Int32}, Ptr{Int32}), ctx.h, t.h, z, C_NULL, C_NULL)
    check(ctx, rc, "dmrj_tensor_to_dense_z")
    return z
  end
  out = zeros(Float64, dims...)
  rc = ccall((:dmrj_tensor_to_dense, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}),
             ctx.h, t.h, out, C_NULL, C_NULL)
  check(ctx, rc, "dmrj_tensor_to_dense")
  return out
end

"matrixproductstate on the GPU (lib/types.jl:84); a regMPS when DMRJtensor is loaded"
mutable struct gpuMPS <: _MPSsuper
  A::Vector{gpuTens}
  oc::Int
  nq::Int
  qtot::Vector{Int32}   # total charge (right boundary of the environments); empty when nq == 0
end
gpuMPS(A::Vector{gpuTens}, oc::Integer, nq::Integer) = gpuMPS(A, oc, nq, zeros(Int32, nq))
Base.getindex(psi::gpuMPS, i::Integer) = psi.A[i]
Base.setindex!(psi::gpuMPS, v::gpuTens, i::Integer) = (psi.A[i] = v)
Base.length(psi::gpuMPS) = length(psi.A)

"Upload a dense (canonical, oc-centred) MPS; optional charges per bond (sorted)."
function gpuMPS(A::Vector{Array{Float64,3}}; oc::Integer=1, qbond=nothing, qphys=nothing)
  if qbond === nothing
    return gpuMPS([upload(a) for a in A], oc, 0)
  end
  nq = size(qphys, 1)
  return gpuMPS([upload(A[i]; qL=qbond[i], qP=qphys, qR=qbond[i+1]) for i = 1:length(A)], oc, nq,
                Vector{Int32}(qbond[end][:, 1]))
end

"one MPO site on the GPU (dmrj_mpo_create handle)"
struct gpuSite
  h::Ptr{Cvoid}
end

"matrixproductoperator on the GPU (lib/types.jl:99); a regMPO when DMRJtensor is loaded"
struct gpuMPO <: _MPOsuper
  W::Vector{gpuSite}
end
Base.getindex(mpo::gpuMPO, i::Integer) = mpo.W[i]
Base.length(mpo::gpuMPO) = length(mpo.W)

function gpuMPO(H::Vector{Array{Float64,4}}; qphys=nothing, deltas=nothing)
  ctx = context()
  nq = qphys === nothing ? 0 : size(qphys, 1)
  W = gpuSite[]
  for i = 1:length(H)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    w = H[i]
    qp = nq == 0 ? C_NULL : pointer(charges(qphys))
    dl = nq == 0 ? C_NULL : pointer(charges(deltas[i]))
    dr = nq == 0 ? C_NULL : pointer(charges(deltas[i+1]))
    GC.@preserve w qphys deltas begin
      rc = ccall((:dmrj_mpo_create, libdmrj), Cint,
                 (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                  Ref{Ptr{Cvoid}}),
                 ctx.h, nq, size(w, 1), size(w, 2), size(w, 4), w, qp, dl, dr, out)
    end
    check(ctx, rc, "dmrj_mpo_create")
    push!(W, gpuSite(out[]))
  end
  return gpuMPO(W)
end

const gpuEnv = Vector{Union{Nothing,gpuTens}}

function boundary(left::Bool, nq::Integer, q=nothing)
  ctx = context()
  out = Ref{Ptr{Cvoid}}(C_NULL)
  qq = nq == 0 ? C_NULL : pointer(Vector{Int32}(q))
  rc = ccall((:dmrj_env_boundary, libdmrj), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Int32}, Ref{Ptr{Cvoid}}),
             ctx.h, left ? 1 : 0, nq, qq, out)
  check(ctx, rc, "dmrj_env_boundary")
  return gpuTens(out[])
end

"""Lupdate(Lenv, dualpsi, psi, mpo) (lib/MPO_MPS/Lupdate.jl:17-26), the
argument order of the reference's call in optmps (algs/optimizeMPS.jl:810).
The GPU environment update takes one tensor for bra and ket: dualpsi must be
psi (the two-site DMRG path passes the same MPS twice)."""
function Lupdate(Lenv::gpuTens, dualpsi::gpuTens, psi::gpuTens, W::gpuSite)
  dualpsi.h == psi.h || error("Lupdate on the GPU: dualpsi must be psi")
  ctx = context()
  out = Ref{Ptr{Cvoid}}(C_NULL)
  rc = ccall((:dmrj_env_left, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
             ctx.h, Lenv.h, psi.h, W.h, out)
  check(ctx, rc, "dmrj_env_left")
  return gpuTens(out[])
end
Lupdate(Lenv::gpuTens, psi::gpuTens, W::gpuSite) = Lupdate(Lenv, psi, psi, W)

"""Rupdate(Renv, dualpsi, psi, mpo) (lib/MPO_MPS/Rupdate.jl:17-26), the
argument order of optmps (algs/optimizeMPS.jl:815); dualpsi must be psi."""
function Rupdate(Renv::gpuTens, dualpsi::gpuTens, psi::gpuTens, W::gpuSite)
  dualpsi.h == psi.h || error("Rupdate on the GPU: dualpsi must be psi")
  ctx = context()
  out = Ref{Ptr{Cvoid}}(C_NULL)
  rc = ccall((:dmrj_env_right, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
             ctx.h, Renv.h, psi.h, W.h, out)
  check(ctx, rc, "dmrj_env_right")
  return gpuTens(out[])
end
Rupdate(Renv::gpuTens, psi::gpuTens, W::gpuSite) = Rupdate(Renv, psi, psi, W)

"makeEnv(dualpsi, psi, mpo) (lib/Env/makeEnv.jl:196-222) with dualpsi = psi"
function makeEnv(dualpsi::gpuMPS, psi::gpuMPS, mpo::gpuMPO; qtot=nothing)
  dualpsi === psi || error("makeEnv on the GPU: dualpsi must be psi")
  return makeEnv(psi, mpo; qtot=qtot)
end

"makeEnv (lib/Env/makeEnv.jl:196-222,244); qtot = total charge of the state"
function makeEnv(psi::gpuMPS, mpo::gpuMPO; qtot=nothing)
  Ns = length(psi)
  Lenv = gpuEnv(nothing, Ns)
  Renv = gpuEnv(nothing, Ns)
  Lenv[1] = boundary(true, psi.nq, psi.nq == 0 ? nothing : zeros(Int32, psi.nq))
  Renv[Ns] = boundary(false, psi.nq, qtot === nothing ? (psi.nq == 0 ? nothing : psi.qtot) : qtot)
  for i = 1:psi.oc-1
    Lenv[i+1] = Lupdate(Lenv[i], psi[i], mpo[i])
  end
  for i = Ns:-1:psi.oc+1
    Renv[i-1] = Rupdate(Renv[i], psi[i], mpo[i])
  end
  return Lenv, Renv
end

"""expect(psi, mpo) (lib/MPO_MPS/expect.jl:19-57): <psi|H|psi> by a
right-to-left chain of GPU environment updates, closed at the left boundary."""
function expect(psi::gpuMPS, mpo::gpuMPO; qtot=nothing)
  ctx = context()
  Ns = length(psi)
  R = boundary(false, psi.nq, qtot === nothing ? (psi.nq == 0 ? nothing : psi.qtot) : qtot)
  for i = Ns:-1:1
    Rn = Rupdate(R, psi[i], mpo[i])
    free!(R)
    R = Rn
  end
  L = boundary(true, psi.nq, psi.nq == 0 ? nothing : zeros(Int32, psi.nq))
  e = Ref{Float64}(0.0)
  check(ctx, ccall((:dmrj_dot, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}),
                   ctx.h, L.h, R.h, e), "dmrj_dot")
  free!(L); free!(R)
  return e[]
end
expect(dualpsi::gpuMPS, psi::gpuMPS, mpo::gpuMPO; qtot=nothing) =
  (dualpsi === psi ? expect(psi, mpo; qtot=qtot) : error("expect on the GPU: dualpsi must be psi"))

"""applyOps!(psi, sites, Op; trail) (lib/MPO_MPS/applyMPO.jl:21-33) on the GPU:
psi[p] = Op psi[p] for every p in sites, the trail on the sites left of p.
dq = charge moved by Op (q(k) - q(j) of its non-zero Op[k, j]): on a charged
MPS the bonds right of p move by dq; the trail must be neutral."""
function applyOps!(psi::gpuMPS, sites::Vector{<:Integer}, Op::AbstractMatrix; trail=nothing, dq=nothing)
  ctx = context()
  nq = psi.nq
  zero = zeros(Int32, max(nq, 1))
  shift = nq == 0 ? zero : Vector{Int32}(dq)
  function site!(i, O, l, r)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    buf = Matrix{Float64}(O)
    GC.@preserve buf l r begin
      rc = ccall((:dmrj_site_apply_op, libdmrj), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ref{Ptr{Cvoid}}),
                 ctx.h, psi[i].h, buf, nq == 0 ? C_NULL : pointer(l), nq == 0 ? C_NULL : pointer(r), out)
    end
    check(ctx, rc, "dmrj_site_apply_op")
    free!(psi[i])
    psi[i] = gpuTens(out[])
  end
  for p in sites
    site!(p, Op, zero, shift)
    if nq > 0 && any(shift .!= 0)
      for i = p+1:length(psi)
        d = info(psi[i])[3]
        site!(i, Matrix{Float64}(I_d(d)), shift, shift)
      end
    end
    if trail !== nothing
      for i = 1:p-1
        site!(i, trail, zero, zero)
      end
    end
  end
  return psi
end
I_d(d) = [Float64(a == b) for a = 1:d, b = 1:d]

mutable struct algvars
  maxm::Int
  minm::Int
  sweeps::Int
  cutoff::Float64
  goal::Float64
  silent::Bool
  SvNbond::Int
  precision::String
  energy::Float64
  lastenergy::Float64
  truncerr::Float64
  maxtrunc::Float64
  biggestm::Int
  entropy::Float64
  storeD::Vector{Float64}
  # 3S: Krylov vectors per site and the subspace-expansion noise control
  r::Int
  noise::Float64
  noise_goal::Float64
  noise_incr::Float64
end

_precision(params) = hasproperty(params, :precision) ? params.precision : "f64"

"SvNcheck! (algs/DMRG.jl:172-210, allSvNbond = false) on the kept singular values"
function svn_check!(params, i, j, Ns, sv::Vector{Float64})
  params.maxtrunc = max(params.maxtrunc, params.truncerr)
  params.biggestm = max(params.biggestm, length(sv))
  if i == min(max(2, params.SvNbond + 1), Ns) && j < 0
    D = sort(sv, rev=true)
    params.entropy = -sum(x -> x > 0 ? x^2 * log(x^2) : 0.0, D)
    params.storeD = D[1:min(8, length(D))]
  end
  nothing
end

"""twostep (algs/DMRG.jl:362-405) with the stepfct signature of the
reference's optmps (algs/optimizeMPS.jl:805): one dmrj_twostep call
(normalise the centre, simplelanczos, truncated SVD, split U | D*V or U*D | V);
the reference's unused twosite_update product is not computed.  params is
the reference's algvars (through DMRJtensor.optmps) or this module's.  With a
split precision the FP64 Rayleigh-quotient energy is taken at the last bond
of a sweep (j < 0, iL == 1), the energy a sweep reports."""
function twostep(n::Integer, j::Integer, i::Integer, iL::Integer, iR::Integer, dualpsi::gpuMPS, psi::gpuMPS,
                 mpo::gpuMPO, Lenv, Renv, psiLenv, psiRenv, beta::Array{Y,1}, prevpsi...;
                 params=nothing) where Y <: Number
  dualpsi === psi || error("twostep on the GPU: dualpsi must be psi")
  isempty(prevpsi) || error("twostep on the GPU: previous-state penalties are not on the GPU path")
  ctx = context()
  cap = max(8, 4 * (params.maxm > 0 ? params.maxm : 4096))
  sv = zeros(Float64, cap)
  nl = Ref{Ptr{Cvoid}}(C_NULL); nr = Ref{Ptr{Cvoid}}(C_NULL)
  e = Ref{Float64}(0.0); te = Ref{Float64}(0.0); kept = Ref{Int64}(0)
  pname = _precision(params)
  correct = pname != "f64" && j < 0 && iL == 1
  prec = PREC[pname] | (correct ? CORRECT : Cint(0))
  rc = ccall((:dmrj_twostep, libdmrj), Cint,
             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64,
              Int64, Float64, Cint, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}, Ref{Float64}, Ref{Float64}, Ref{Int64},
              Ptr{Float64}, Int64),
             ctx.h, Lenv[iL].h, Renv[iR].h, psi[iL].h, psi[iR].h, mpo[iL].h, mpo[iR].h, j > 0 ? 1 : -1,
             params.maxm, params.minm, params.cutoff, prec, nl, nr, e, te, kept, sv, cap)
  check(ctx, rc, "dmrj_twostep")
  free!(psi[iL]); free!(psi[iR])
  psi[iL] = gpuTens(nl[]); psi[iR] = gpuTens(nr[])
  params.energy = e[]
  params.truncerr = te[]
  svn_check!(params, i, j, length(psi), sv[1:min(cap, kept[])])
  nothing
end

"noise control of 3S (algs/DMRG.jl:288-303); x/0 follows IEEE as in the reference"
function alpha_update(alpha, truncerr, cutoff, lastE, currE, goal, dalpha)
  small_trunc = truncerr <= max(cutoff, 1e-12)
  small_step = abs(currE - lastE) <= max(abs(currE * cutoff), 1e-12)
  site = abs((truncerr * currE) / (lastE / currE - 1)) > goal
  alpha *= (currE <= lastE && (site || small_step || small_trunc)) ? 1 - dalpha : 1 + dalpha
  return max(min(alpha, 1.0), 1e-42)
end

"""step3S (algs/DMRG.jl:322-360) with the stepfct signature of optmps:
lanczos1, expansion, truncated SVD, restricted absorb (FP64)."""
function step3S(n::Integer, j::Integer, i::Integer, iL::Integer, iR::Integer, dualpsi::gpuMPS, psi::gpuMPS,
                mpo::gpuMPO, Lenv, Renv, psiLenv, psiRenv, beta::Array{Y,1}, prevpsi...;
                params=nothing) where Y <: Number
  dualpsi === psi || error("step3S on the GPU: dualpsi must be psi")
  ctx = context()
  A = Ref{Ptr{Cvoid}}(C_NULL); e = Ref{Float64}(0.0); k = Ref{Cint}(0)
  rc = ccall((:dmrj_lanczos1, libdmrj), Cint,
             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ref{Ptr{Cvoid}}, Ref{Float64},
              Ref{Cint}), ctx.h, Lenv[i].h, Renv[i].h, psi[i].h, mpo[i].h, params.r, 0, A, e, k)
  check(ctx, rc, "dmrj_lanczos1")
  noise = params.noise[1]   # a scalar or, in the reference's algvars, possibly a vector
  alpha = noise > params.noise_incr ? noise : 0.0
  dir = j > 0 ? Cint(1) : Cint(-1)
  T = Ref{Ptr{Cvoid}}(C_NULL)
  rc = ccall((:dmrj_expand3s, libdmrj), Cint,
             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Cint, Ref{Ptr{Cvoid}}),
             ctx.h, j > 0 ? Lenv[iL].h : Renv[iR].h, A[], mpo[i].h, alpha, dir, T)
  check(ctx, rc, "dmrj_expand3s")
  ccall((:dmrj_tensor_free, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, A[])
  check(ctx, ccall((:dmrj_normalize, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, T[]), "dmrj_normalize")
  cap = max(8, 4 * (params.maxm > 0 ? params.maxm : 4096))
  sv = zeros(Float64, cap)
  nl = Ref{Ptr{Cvoid}}(C_NULL); nr = Ref{Ptr{Cvoid}}(C_NULL)
  te = Ref{Float64}(0.0); kept = Ref{Int64}(0)
  rc = ccall((:dmrj_svd_trunc2, libdmrj), Cint,
             (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Float64, Cint, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}, Ref{Float64},
              Ref{Int64}, Ptr{Float64}, Int64),
             ctx.h, T[], params.maxm, params.minm, params.cutoff, dir, nl, nr, te, kept, sv, cap)
  check(ctx, rc, "dmrj_svd_trunc2")
  ccall((:dmrj_tensor_free, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, T[])
  out = Ref{Ptr{Cvoid}}(C_NULL)
  F, N = j > 0 ? (nr[], psi[iR].h) : (nl[], psi[iL].h)
  rc = ccall((:dmrj_absorb3s, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}),
             ctx.h, F, N, dir, out)
  check(ctx, rc, "dmrj_absorb3s")
  ccall((:dmrj_tensor_free, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, F)
  free!(psi[iL]); free!(psi[iR])
  if j > 0
    psi[iL] = gpuTens(nl[]); psi[iR] = gpuTens(out[])
  else
    psi[iL] = gpuTens(out[]); psi[iR] = gpuTens(nr[])
  end
  params.energy = e[]
  params.noise = alpha_update(noise, te[], params.cutoff, params.lastenergy[1], params.energy[1], params.noise_goal,
                              params.noise_incr)
  params.truncerr = te[]
  svn_check!(params, i, j, length(psi), sv[1:min(cap, kept[])])
  nothing
end

"""startval (algs/optimizeMPS.jl:255-258): [<dualpsi|H_1site psi>], the
measfct of optmps"""
function gpu_startval(Lenv::gpuTens, Renv::gpuTens, dualpsi::gpuTens, psi::gpuTens, W::gpuSite)
  ctx = context()
  H1 = Ref{Ptr{Cvoid}}(C_NULL); e0 = Ref{Float64}(0.0)
  rc = ccall((:dmrj_heff1_apply, libdmrj), Cint,
             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}),
             ctx.h, Lenv.h, W.h, Renv.h, psi.h, 0, H1)
  check(ctx, rc, "dmrj_heff1_apply")
  rc = ccall((:dmrj_dot, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}), ctx.h, dualpsi.h, H1[], e0)
  ccall((:dmrj_tensor_free, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.h, H1[])
  check(ctx, rc, "dmrj_dot")
  return [e0[]]
end

"""setEnv (algs/optimizeMPS.jl:260-410) for GPU states, the Envfct of optmps:
environments up to the centre, the opening energy from measfct, the number
of bond steps per sweep and the first direction.  Returns (Lenv, Renv,
psiLenv, psiRenv, Nsteps, j) as setEnv does; previous-state environments are
not on the GPU path."""
function gpu_setEnv(dualpsi::gpuMPS, psi::gpuMPS, mpo::gpuMPO, params, prevpsi...; measfct=gpu_startval,
                    mover=nothing)
  dualpsi === psi || error("setEnv on the GPU: dualpsi must be psi")
  isempty(prevpsi) || error("setEnv on the GPU: previous-state penalties are not on the GPU path")
  Ns = length(psi)
  Lenv, Renv = makeEnv(psi, mpo)
  if isapprox(sum(params.energy), 0.0)
    params.energy = real(measfct(Lenv[psi.oc], Renv[psi.oc], psi[psi.oc], psi[psi.oc], mpo[psi.oc])[1])
  end
  params.lastenergy = copy(params.energy)
  j = psi.oc == Ns ? -1 : 1
  return Lenv, Renv, gpuEnv[], gpuEnv[], 2 * (Ns - 1), j
end

"""dmrg(psi, mpo; m, minm, sweeps, cutoff, silent, goal, method) (algs/DMRG.jl:
549-589), method = "twosite" or "3S", on GPU states.  With DMRJtensor loaded
and FP64 arithmetic it runs the reference's own optmps with this module's
stepfct (twostep / step3S), Envfct and measfct and the reference's dmrgcvg;
otherwise (or for precision = "bf16x3" / "f16x3") the same sweep loop
restated here."""
function dmrg(psi::gpuMPS, mpo::gpuMPO; m::Integer=0, minm::Integer=2, sweeps::Integer=1, cutoff::Float64=0.,
              silent::Bool=false, goal::Float64=0., method::String="twosite", precision::String="f64",
              qtot=nothing, r::Integer=2, noise::Float64=1.0, noise_goal::Float64=0.3, noise_incr::Float64=0.01)
  method in ("twosite", "3S") || error("DMRG method ($method) not defined on the GPU path")
  method == "3S" && precision != "f64" && error("3S runs in FP64 on the GPU path")
  qtot === nothing || (psi.qtot = Vector{Int32}(qtot))
  stepfct = method == "3S" ? step3S : twostep
  Ns = length(psi)
  if HAS_DMRJ && precision == "f64"
    D = Main.DMRJtensor
    params = D.algvars(D.U1, D.default_boundary, gpuEnv(), gpuEnv())
    params.method = "DMRG-" * method
    params.maxm = m; params.minm = minm; params.sweeps = sweeps; params.cutoff = cutoff
    params.silent = silent; params.goal = goal; params.SvNbond = cld(Ns, 2)
    params.r = r; params.startnoise = noise; params.noise = noise
    params.noise_goal = noise_goal; params.noise_incr = noise_incr
    params.energy = 0.0
    return D.optmps(psi, psi, mpo, [1.]; params=params, Envfct=gpu_setEnv, measfct=gpu_startval,
                    stepfct=stepfct, cvgfct=D.dmrgcvg, displayfct=p -> nothing)
  end
  params = algvars(m, minm, sweeps, cutoff, goal, silent, cld(Ns, 2), precision, 0., 0., 0., 0., 0, 0., zeros(8),
                   r, noise, noise_goal, noise_incr)
  Lenv, Renv, psiLenv, psiRenv, Nsteps, j = gpu_setEnv(psi, psi, mpo, params)
  ccall((:dmrj_ctx_set_env_precision, libdmrj), Cint, (Ptr{Cvoid}, Cint), context().h, PREC[precision])
  try
    for n = 1:sweeps
      timer = -time()
      for ns = 1:Nsteps
        i = psi.oc
        iL, iR = j > 0 ? (i, i + 1) : (i - 1, i)
        stepfct(n, j, i, iL, iR, psi, psi, mpo, Lenv, Renv, psiLenv, psiRenv, [1.0]; params=params)
        if j > 0
          Lenv[i+1] = Lupdate(Lenv[i], psi[i], psi[i], mpo[i])
        else
          Renv[i-1] = Rupdate(Renv[i], psi[i], psi[i], mpo[i])
        end
        psi.oc += j
        if (psi.oc == Ns && j > 0) || (psi.oc == 1 && j < 0)
          j *= -1
        end
      end
      timer += time()
      if !silent
        println("Sweep $n (back and forth): $(round(timer, digits=5)) sec")
        println("Largest truncation = $(params.maxtrunc), m = $(params.biggestm)")
        println("Energy at sweep $n is $(params.energy)")
      end
      params.truncerr = params.maxtrunc
      params.maxtrunc = 0.; params.biggestm = 0
      if abs(params.energy - params.lastenergy) < goal
        break
      end
      params.lastenergy = params.energy
    end
  finally
    ccall((:dmrj_ctx_set_env_precision, libdmrj), Cint, (Ptr{Cvoid}, Cint), context().h, 0)
  end
  return params.energy
end

"krylovTimeEvol (algs/time.jl:182-202) of a two-site (two=true) or one-site tensor; complex prefactor"
function krylovTimeEvol(Lenv::gpuTens, Renv::gpuTens, A::gpuTens, mpo::gpuMPO, i::Integer, r::Integer,
                        prefactor::Number; two::Bool=true)
  ctx = context()
  out = Ref{Ptr{Cvoid}}(C_NULL); e = Ref{Float64}(0.0); k = Ref{Cint}(0)
  pf = ComplexF64(prefactor)
  if two
    rc = ccall((:dmrj_krylov_expm2, libdmrj), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Float64, Float64, Cint,
                Ref{Ptr{Cvoid}}, Ref{Float64}, Ref{Cint}),
               ctx.h, Lenv.h, Renv.h, A.h, mpo[i].h, mpo[i+1].h, r, real(pf), imag(pf), 0, out, e, k)
  else
    rc = ccall((:dmrj_krylov_expm1, libdmrj), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Float64, Float64, Cint,
                Ref{Ptr{Cvoid}}, Ref{Float64}, Ref{Cint}),
               ctx.h, Lenv.h, Renv.h, A.h, mpo[i].h, r, real(pf), imag(pf), 0, out, e, k)
  end
  check(ctx, rc, "dmrj_krylov_expm")
  return gpuTens(out[]), e[]
end

normalize!(t::gpuTens) = check(context(), ccall((:dmrj_normalize, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                                                context().h, t.h), "dmrj_normalize")

"tdvp_twosite (algs/time.jl:288-362): one half-sweep, step prefactor/(Ns-1)/2 per bond"
function tdvp(psi::gpuMPS, mpo::gpuMPO; m::Integer=0, minm::Integer=2, cutoff::Float64=0., r::Integer=2,
              prefactor::Number=1., Lenv=nothing, Renv=nothing, qtot=nothing)
  ctx = context()
  Ns = length(psi)
  timestep = prefactor / (Ns - 1) / 2
  if Lenv === nothing || Renv === nothing
    Lenv, Renv = makeEnv(psi, mpo; qtot=qtot)
  end
  j = psi.oc == Ns ? -1 : 1
  En = 0.
  for _ = 1:Ns-1
    i = psi.oc
    iL, iR = j > 0 ? (i, i + 1) : (i - 1, i)
    th = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:dmrj_theta2, libdmrj), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                     ctx.h, psi[iL].h, psi[iR].h, th), "dmrj_theta2")
    AA = gpuTens(th[])
    newpsi, En = krylovTimeEvol(Lenv[iL], Renv[iR], AA, mpo, iL, r, timestep)
    free!(AA)
    normalize!(newpsi)
    nl = Ref{Ptr{Cvoid}}(C_NULL); nr = Ref{Ptr{Cvoid}}(C_NULL); te = Ref{Float64}(0.); kept = Ref{Int64}(0)
    sv = zeros(8)
    check(ctx, ccall((:dmrj_svd_trunc2, libdmrj), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Float64, Cint, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}},
                      Ref{Float64}, Ref{Int64}, Ptr{Float64}, Int64),
                     ctx.h, newpsi.h, m, minm, cutoff, j > 0 ? 1 : -1, nl, nr, te, kept, sv, 8), "dmrj_svd_trunc2")
    free!(newpsi); free!(psi[iL]); free!(psi[iR])
    left = gpuTens(nl[]); right = gpuTens(nr[])
    if j > 0
      psi[iL] = left
      Lenv[iR] = Lupdate(Lenv[iL], psi[iL], mpo[iL])
      normalize!(right)
      psi[iR], En = krylovTimeEvol(Lenv[iR], Renv[iR], right, mpo, iR, r, timestep; two=false)
      free!(right)
    else
      psi[iR] = right
      Renv[iL] = Rupdate(Renv[iR], psi[iR], mpo[iR])
      normalize!(left)
      psi[iL], En = krylovTimeEvol(Lenv[iL], Renv[iL], left, mpo, iL, r, timestep; two=false)
      free!(left)
    end
    psi.oc += j
    normalize!(psi[psi.oc])
    if psi.oc == Ns && j > 0
      j = -1
    elseif psi.oc == 1 && j < 0
      j = 1
    end
  end
  return En, Lenv, Renv
end

end # module
