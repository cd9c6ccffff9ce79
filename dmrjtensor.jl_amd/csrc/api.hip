// libdmrjhip: runtime, block-sparse metadata, contraction plans, Lanczos and
// Krylov drivers, the 3S expansion and the C ABI declared in include/dmrj.h
// (the truncated two-site SVD is in svd.hip).
//
// Every contraction of the two-site DMRG step is expressed as
//   GEMM (all symmetry blocks of the step in one grouped launch)
//   -> MIX (MPO coefficients over the small labels, HBM-bound)
//   -> GEMM (all blocks in one grouped launch)
// planned on the host from the block tables; tensors stay resident in HBM
// between calls and only scalars (energy, truncation error, singular values)
// come back to the host.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

#include "../../include/dmrj.h"
#include "internal.h"

#define HIP_OK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw dmrj::Error(-3, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

namespace dmrj {

// ====================================================================== spaces
int Space::find(const Charge& c) const {
  auto it = std::lower_bound(q.begin(), q.end(), c);
  if (it == q.end() || !(*it == c)) return -1;
  return (int)(it - q.begin());
}

void Space::finish() {
  off.resize(q.size());
  total = 0;
  for (size_t i = 0; i < q.size(); ++i) {
    off[i] = total;
    total += dim[i];
  }
}

bool Space::same(const Space& o) const {
  if (q.size() != o.q.size()) return false;
  for (size_t i = 0; i < q.size(); ++i)
    if (!(q[i] == o.q[i]) || dim[i] != o.dim[i]) return false;
  return true;
}

void SiteMeta::build() {
  blocks.clear();
  bidx.assign((size_t)left.nsec() * P, -1);
  panel_off.assign(left.nsec(), 0);
  panel_cols.assign(left.nsec(), 0);
  int64_t cur = 0;
  for (int a = 0; a < left.nsec(); ++a) {
    panel_off[a] = cur;
    int64_t cols = 0;
    for (int s = 0; s < P; ++s) {
      const int b = right.find(left.q[a] + qp[s]);
      if (b < 0) continue;
      bidx[(size_t)a * P + s] = (int)blocks.size();
      blocks.push_back({a, s, b, cur + cols * left.dim[a], cols});
      cols += right.dim[b];
    }
    panel_cols[a] = cols;
    cur += cols * left.dim[a];
  }
  size = cur;
}

void EnvMeta::build() {
  blocks.clear();
  bidx.assign((size_t)bra.nsec() * D, -1);
  panel_off.assign(bra.nsec(), 0);
  panel_cols.assign(bra.nsec(), 0);
  int64_t cur = 0;
  for (int ap = 0; ap < bra.nsec(); ++ap) {
    panel_off[ap] = cur;
    int64_t cols = 0;
    for (int w = 0; w < D; ++w) {
      const int a = ket.find(bra.q[ap] - delta[w]);
      if (a < 0) continue;
      bidx[(size_t)ap * D + w] = (int)blocks.size();
      blocks.push_back({ap, w, a, cur + cols * bra.dim[ap], cols});
      cols += ket.dim[a];
    }
    panel_cols[ap] = cols;
    cur += cols * bra.dim[ap];
  }
  size = cur;
}

// =================================================================== allocator
static size_t round_bytes(size_t b) {
  if (b < 256) return 256;
  // size classes: 8 per power of two
  size_t p = 1;
  while (p < b) p <<= 1;
  const size_t step = std::max<size_t>(p / 8, 256);
  return (b + step - 1) / step * step;
}

void* Allocator::alloc(size_t bytes) {
  const size_t r = round_bytes(bytes);
  auto it = free_.find(r);
  void* p = nullptr;
  if (it != free_.end()) {
    p = it->second;
    free_.erase(it);
    bytes_cached -= r;
  } else {
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      trim();
      HIP_OK(hipMalloc(&p, r));
    }
  }
  live_[p] = r;
  bytes_live += r;
  peak = std::max(peak, bytes_live);
  return p;
}

void Allocator::release(void* p) {
  if (!p) return;
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.insert({it->second, p});
  bytes_cached += it->second;
  bytes_live -= it->second;
  live_.erase(it);
}

void Allocator::trim() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  bytes_cached = 0;
}

Allocator::~Allocator() {
  trim();
  for (auto& kv : live_) (void)hipFree(kv.first);
}

// ====================================================================== staging
void* Context::upload(const void* host, size_t bytes) {
  const size_t r = (bytes + 255) / 256 * 256;
  if (stage.used + r > stage.cap) {
    // all earlier staged copies must have landed before the buffer is reused
    HIP_OK(hipStreamSynchronize(stream));
    stage.used = 0;
    if (r > stage.cap) {
      if (stage.base) HIP_OK(hipHostFree(stage.base));
      stage.cap = std::max(r * 2, stage.cap * 2);
      HIP_OK(hipHostMalloc((void**)&stage.base, stage.cap, hipHostMallocDefault));
    }
  }
  char* h = stage.base + stage.used;
  if (host && bytes)
    std::memcpy(h, host, bytes);
  else
    std::memset(h, 0, r);
  stage.used += r;
  void* d = alloc.alloc(r);
  HIP_OK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
  // the device copy is returned to the pool right away: later users of the
  // same memory are ordered after the kernels that read it (single stream)
  transient.push_back(d);
  return d;
}

void Context::stage_reset() { stage.used = 0; }

void Context::release_transients() {
  for (void* p : transient) alloc.release(p);
  transient.clear();
  prof_resolve();
}

int Context::prof_begin() {
  if (!prof) return -1;
  while (prof_pool.size() < 2) {
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing only: no cache writeback
    prof_pool.push_back(e);
  }
  ProfRec r;
  r.a = prof_pool.back();
  prof_pool.pop_back();
  r.b = prof_pool.back();
  prof_pool.pop_back();
  r.kind = -1;
  r.flops = r.bytes = 0.0;
  HIP_OK(hipEventRecord(r.a, stream));
  prof_pending.push_back(r);
  return (int)prof_pending.size() - 1;
}

int Context::prof_begin_ext() {
  if (!prof) return -1;
  while (prof_pool.size() < 2) {
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing only: no cache writeback
    prof_pool.push_back(e);
  }
  ProfRec r;
  r.a = prof_pool.back();
  prof_pool.pop_back();
  r.b = prof_pool.back();
  prof_pool.pop_back();
  r.kind = -1;
  r.flops = r.bytes = 0.0;
  prof_pending.push_back(r);
  ext_a = r.a;
  ext_b = r.b;
  return (int)prof_pending.size() - 1;
}

void Context::prof_end_ext(int slot, int kind, double flops, double bytes) {
  ext_a = ext_b = nullptr;
  if (slot < 0) return;
  ProfRec& r = prof_pending[slot];
  r.kind = kind;
  r.flops = flops;
  r.bytes = bytes;
}

void Context::prof_end(int slot, int kind, double flops, double bytes) {
  if (slot < 0) return;
  ProfRec& r = prof_pending[slot];
  r.kind = kind;
  r.flops = flops;
  r.bytes = bytes;
  HIP_OK(hipEventRecord(r.b, stream));
}

void Context::prof_resolve() {
  for (auto& r : prof_pending) {
    if (r.kind >= 0 && r.kind < 8) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        prof_ms[r.kind] += ms;
        prof_flops[r.kind] += r.flops;
        prof_bytes[r.kind] += r.bytes;
        prof_count[r.kind] += 1;
      }
    }
    prof_pool.push_back(r.a);
    prof_pool.push_back(r.b);
  }
  prof_pending.clear();
}

}  // namespace dmrj

using namespace dmrj;

namespace dmrj {

dmrj_tensor* new_tensor(Context& c, int kind) {
  auto* t = new dmrj_tensor();
  t->kind = kind;
  (void)c;
  return t;
}

void alloc_data(Context& c, Tensor& t) {
  const int64_t n = t.nstore();
  t.data = (double*)c.alloc.alloc(std::max<int64_t>(n, 1) * sizeof(double));
}

int nq_of(const Tensor& t) { return t.nq; }

}  // namespace dmrj

namespace {

// ===================================================================== helpers
Charge charge_of(const int32_t* q, int nq, int64_t i) {
  Charge c;
  for (int k = 0; k < nq; ++k) c.v[k] = q ? q[i * nq + k] : 0;
  return c;
}

Space space_from_list(const int32_t* q, int nq, int64_t n) {
  Space s;
  if (n <= 0) throw Error(DMRJ_ERR_ARG, "empty bond index");
  for (int64_t i = 0; i < n; ++i) {
    Charge c = nq ? charge_of(q, nq, i) : Charge();
    if (!s.q.empty()) {
      if (c < s.q.back()) throw Error(DMRJ_ERR_ARG, "bond charges must be sorted");
      if (c == s.q.back()) {
        s.dim.back()++;
        continue;
      }
    }
    s.q.push_back(c);
    s.dim.push_back(1);
  }
  s.finish();
  return s;
}

void space_to_list(const Space& s, int nq, int32_t* out) {
  if (!out) return;
  int64_t i = 0;
  for (int a = 0; a < s.nsec(); ++a)
    for (int k = 0; k < s.dim[a]; ++k, ++i)
      for (int c = 0; c < nq; ++c) out[i * nq + c] = s.q[a].v[c];
}

void check_finite_host(const double* p, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) throw Error(DMRJ_ERR_NUMERIC, "non-finite input");
}

// Upload a dense column-major site tensor [mL, P, mR] into block layout.
dmrj_tensor* site_from_dense(Context& c, int nq, int64_t mL, int64_t P, int64_t mR,
                             const double* A, const int32_t* qL,
                             const std::vector<Charge>& qp, const int32_t* qR, int d1, int d2,
                             const std::vector<Charge>& qp1, const std::vector<Charge>& qp2,
                             const double* Aim = nullptr) {
  if (nq < 0 || nq > kMaxQ) throw Error(DMRJ_ERR_ARG, "nq must be 0, 1 or 2");
  if (mL <= 0 || P <= 0 || mR <= 0) throw Error(DMRJ_ERR_ARG, "bad dimensions");
  check_finite_host(A, mL * P * mR);
  if (Aim) check_finite_host(Aim, mL * P * mR);
  auto* t = new_tensor(c, 0);
  t->nq = nq;
  SiteMeta& m = t->site;
  m.left = space_from_list(qL, nq, mL);
  m.right = space_from_list(qR, nq, mR);
  m.P = (int)P;
  m.qp = qp;
  m.d1 = d1;
  m.d2 = d2;
  m.qp1 = qp1;
  m.qp2 = qp2;
  m.build();
  t->cplx = Aim ? 1 : 0;
  const int64_t plane = std::max<int64_t>(m.size, 1);
  std::vector<double> h(plane * (Aim ? 2 : 1), 0.0);
  std::vector<char> used((size_t)(mL * P * mR), 0);
  for (const auto& b : m.blocks) {
    const int na = m.left.dim[b.a], nb = m.right.dim[b.b];
    const int64_t oa = m.left.off[b.a], ob = m.right.off[b.b];
    for (int cb = 0; cb < nb; ++cb)
      for (int ra = 0; ra < na; ++ra) {
        const int64_t src = (oa + ra) + mL * (b.s + P * (ob + cb));
        h[b.off + ra + (int64_t)cb * na] = A[src];
        if (Aim) h[plane + b.off + ra + (int64_t)cb * na] = Aim[src];
        used[src] = 1;
      }
  }
  double drop2 = 0.0, tot2 = 0.0;
  for (int64_t i = 0; i < mL * P * mR; ++i) {
    const double a2 = A[i] * A[i] + (Aim ? Aim[i] * Aim[i] : 0.0);
    tot2 += a2;
    if (!used[i]) drop2 += a2;
  }
  if (drop2 > 1e-24 * std::max(1.0, tot2)) {
    delete t;
    throw Error(DMRJ_ERR_NUMERIC, "tensor has weight outside the charge-allowed blocks");
  }
  alloc_data(c, *t);
  HIP_OK(hipMemcpyAsync(t->data, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  return t;
}

// ================================================================ H_eff plan
// Fused MPO coefficient provider: coef(w, J, K, u) for the combined physical
// index of one or two sites.
struct FusedMpo {
  int DL, DR, P;
  std::vector<double> c;  // [w + DL*(J + P*(K + P*u))]
  std::vector<Charge> dL, dR;
  double at(int w, int J, int K, int u) const {
    return c[(size_t)w + (size_t)DL * ((size_t)J + (size_t)P * ((size_t)K + (size_t)P * u))];
  }
};

FusedMpo fuse_one(const MpoSite& W) {
  FusedMpo f;
  f.DL = W.DL;
  f.DR = W.DR;
  f.P = W.d;
  f.c = W.W;
  f.dL = W.dL;
  f.dR = W.dR;
  return f;
}

// W12[w, (j1,j2), (k1,k2), u] = sum_v W1[w,j1,k1,v] W2[v,j2,k2,u], J = j1 + d1*j2
FusedMpo fuse_two(const MpoSite& A, const MpoSite& B) {
  if (A.DR != B.DL) throw Error(DMRJ_ERR_ARG, "MPO bond mismatch");
  FusedMpo f;
  f.DL = A.DL;
  f.DR = B.DR;
  f.P = A.d * B.d;
  f.dL = A.dL;
  f.dR = B.dR;
  f.c.assign((size_t)f.DL * f.P * f.P * f.DR, 0.0);
  for (int w = 0; w < A.DL; ++w)
    for (int j1 = 0; j1 < A.d; ++j1)
      for (int k1 = 0; k1 < A.d; ++k1)
        for (int v = 0; v < A.DR; ++v) {
          const double a = A.at(w, j1, k1, v);
          if (a == 0.0) continue;
          for (int j2 = 0; j2 < B.d; ++j2)
            for (int k2 = 0; k2 < B.d; ++k2)
              for (int u = 0; u < B.DR; ++u) {
                const double b = B.at(v, j2, k2, u);
                if (b == 0.0) continue;
                const int J = j1 + A.d * j2, K = k1 + A.d * k2;
                f.c[(size_t)w + (size_t)f.DL * ((size_t)J + (size_t)f.P * ((size_t)K + (size_t)f.P * u))] +=
                    a * b;
              }
        }
  return f;
}

struct HeffPlan {
  std::vector<GemmTask> tA, tC;
  std::vector<GemmSeg> sA, sC;
  std::vector<MixDst> md;
  std::vector<MixTerm> mt;
  int64_t t1 = 0, t3 = 0;
  double flops_gemm = 0.0, flops_alg = 0.0;
};

// Plan y = H_eff x for a site tensor x (P states) with environments L (on
// x.left) and R (on x.right), MPO coefficients f.  Buffers T1/T3 are
// allocated here (returned in the plan's pointers).
HeffPlan plan_heff(Context& c, const SiteMeta& x, const double* xd, double* yd,
                   const EnvMeta& L, const double* Ld, const EnvMeta& R, const double* Rd,
                   const FusedMpo& f, double** T1out, double** T3out, int zmask = 0) {
  // zmask: bit 0 = L complex, bit 1 = x complex, bit 2 = R complex
  const int zf = zmask ? 2 : 1;
  if (!L.ket.same(x.left) || !L.bra.same(x.left))
    throw Error(DMRJ_ERR_ARG, "left environment does not match the tensor's left bond");
  if (!R.ket.same(x.right) || !R.bra.same(x.right))
    throw Error(DMRJ_ERR_ARG, "right environment does not match the tensor's right bond");
  if (L.D != f.DL || R.D != f.DR || x.P != f.P)
    throw Error(DMRJ_ERR_ARG, "MPO dimensions do not match the environments / tensor");
  HeffPlan p;
  const Space& S = x.left;
  const int nL = S.nsec();
  // T1 panels (w, a): rows n_ap, cols C_a
  std::vector<int64_t> t1off((size_t)f.DL * nL, -1);
  for (int w = 0; w < f.DL; ++w)
    for (int a = 0; a < nL; ++a) {
      if (x.panel_cols[a] == 0) continue;
      const int ap = S.find(S.q[a] + f.dL[w]);
      if (ap < 0 || L.block(ap, w) < 0) continue;
      t1off[(size_t)w * nL + a] = p.t1;
      p.t1 += (int64_t)S.dim[ap] * x.panel_cols[a];
    }
  // T3 panels (ap, K): blocks over u aligned with R's bra panel of bp
  std::vector<int64_t> t3pan((size_t)nL * f.P, -1);
  for (int ap = 0; ap < nL; ++ap)
    for (int K = 0; K < f.P; ++K) {
      const int yb = x.block(ap, K);
      if (yb < 0) continue;
      const int bp = x.blocks[yb].b;
      t3pan[(size_t)ap * f.P + K] = p.t3;
      p.t3 += (int64_t)S.dim[ap] * R.panel_cols[bp];
    }
  double* T1 = (double*)c.alloc.alloc(std::max<int64_t>(p.t1, 1) * zf * sizeof(double));
  double* T3 = (double*)c.alloc.alloc(std::max<int64_t>(p.t3, 1) * zf * sizeof(double));
  *T1out = T1;
  *T3out = T3;
  // step A: T1(w,a) = L(ap,w) * Xpanel(a)
  for (int w = 0; w < f.DL; ++w)
    for (int a = 0; a < nL; ++a) {
      const int64_t o = t1off[(size_t)w * nL + a];
      if (o < 0) continue;
      const int ap = S.find(S.q[a] + f.dL[w]);
      const EnvBlock& lb = L.blocks[L.block(ap, w)];
      GemmTask t{T1 + o, S.dim[ap], S.dim[ap], (int)x.panel_cols[a], (int)p.sA.size(), 1, 0.0};
      p.sA.push_back({Ld + lb.off, xd + x.panel_off[a], S.dim[ap], S.dim[a], S.dim[a], 0, 1.0});
      p.tA.push_back(t);
    }
  // step B: T3(ap,K,u,b) = sum_{w,J} f(w,J,K,u) T1(w, a=ap-dL[w], J)
  for (int ap = 0; ap < nL; ++ap)
    for (int K = 0; K < f.P; ++K) {
      const int64_t pan = t3pan[(size_t)ap * f.P + K];
      if (pan < 0) continue;
      const int bp = x.blocks[x.block(ap, K)].b;
      const int nap = S.dim[ap];
      for (int u = 0; u < f.DR; ++u) {
        const int rbi = R.block(bp, u);
        if (rbi < 0) continue;
        const EnvBlock& rb = R.blocks[rbi];
        const int b = rb.a;  // ket sector of the right environment block
        MixDst d{T3 + pan + rb.colpos * nap, (int64_t)nap * x.right.dim[b], (int)p.mt.size(), 0};
        for (int w = 0; w < f.DL; ++w) {
          const int a = S.find(S.q[ap] - f.dL[w]);
          if (a < 0) continue;
          const int64_t o = t1off[(size_t)w * nL + a];
          if (o < 0) continue;
          for (int J = 0; J < f.P; ++J) {
            const int xb = x.block(a, J);
            if (xb < 0 || x.blocks[xb].b != b) continue;
            const double cf = f.at(w, J, K, u);
            if (cf == 0.0) continue;
            p.mt.push_back({T1 + o + x.blocks[xb].colpos * nap, cf});
            d.nt++;
          }
        }
        p.flops_alg += 2.0 * (double)d.n * f.DL * f.P;  // dense-W count of the reference
        p.md.push_back(d);
      }
    }
  // step C: Y(ap,K) = T3panel(ap,K) * Rpanel(bp)^T
  for (int ap = 0; ap < nL; ++ap)
    for (int K = 0; K < f.P; ++K) {
      const int yb = x.block(ap, K);
      if (yb < 0) continue;
      const SiteBlock& blk = x.blocks[yb];
      const int bp = blk.b;
      const int64_t pan = t3pan[(size_t)ap * f.P + K];
      GemmTask t{yd + blk.off, S.dim[ap], S.dim[ap], x.right.dim[bp], (int)p.sC.size(), 0, 0.0};
      const int64_t Kt = R.panel_cols[bp];
      if (Kt > 0) {
        p.sC.push_back({T3 + pan, Rd + R.panel_off[bp], S.dim[ap], x.right.dim[bp], (int)Kt, 2, 1.0});
        t.nseg = 1;
      }
      p.tC.push_back(t);
    }
  p.flops_gemm = gemm_flops(p.tA, p.sA) + gemm_flops(p.tC, p.sC);
  // the reference's algorithmic count uses the dense W contraction for the
  // middle step (4 D^2 d^3 mL mR for two sites); here: the mix terms counted
  // as dense over (w, J) per output element.
  p.flops_alg += p.flops_gemm;
  if (zmask) {
    BufMap bm;
    bm.add(Ld, L.size, zmask & 1);
    bm.add(xd, x.size, (zmask & 2) != 0);
    bm.add(Rd, R.size, (zmask & 4) != 0);
    bm.add(T1, p.t1, true);
    bm.add(T3, p.t3, true);
    bm.add(yd, x.size, true);
    complexify_gemm(p.tA, p.sA, bm);
    complexify_mix(p.md, p.mt, bm);
    complexify_gemm(p.tC, p.sC, bm);
    p.flops_gemm = gemm_flops(p.tA, p.sA) + gemm_flops(p.tC, p.sC);
    p.flops_alg *= 4.0;  // one complex multiply-add = 8 real flops
  }
  return p;
}

void rebase_plan(HeffPlan& p, const double* oldx, const double* newx, int64_t nx,
                 const double* oldy, double* newy, int64_t ny) {
  for (auto& s : p.sA)
    if (s.B >= oldx && s.B < oldx + nx) s.B = newx + (s.B - oldx);
  for (auto& t : p.tC)
    if (t.C >= oldy && t.C < oldy + ny) t.C = newy + (t.C - oldy);
}

// One H_eff application with its descriptor tables on the device.
struct HeffLaunch {
  GemmLaunch a, c;
  MixLaunch m;
  int arith = 0;
  double fa = 0, fc = 0, ba = 0, bc = 0, flops_alg = 0;
};

HeffLaunch prepare_heff(Context& c, const HeffPlan& p, int prec) {
  HeffLaunch h;
  h.arith = prec;  // DMRJ_PREC_F64 / _BF16X3 / _F16X3 = GEMM arithmetic 0 / 1 / 2
  h.a = prepare_gemm(c, p.tA, p.sA, p.tA.empty() ? 0 : pick_gemm_cfg(c, p.tA, p.sA, h.arith), h.arith);
  h.m = prepare_mix(c, p.md, p.mt);
  h.c = prepare_gemm(c, p.tC, p.sC, p.tC.empty() ? 0 : pick_gemm_cfg(c, p.tC, p.sC, h.arith), h.arith);
  h.fa = gemm_flops(p.tA, p.sA);
  h.fc = gemm_flops(p.tC, p.sC);
  h.ba = gemm_bytes(p.tA, p.sA);
  h.bc = gemm_bytes(p.tC, p.sC);
  h.flops_alg = p.flops_alg;
  return h;
}

// Issues the three kernels; with per-kernel timing unless the stream is
// being captured into a graph.
void issue_heff(Context& c, const HeffLaunch& h, bool timed) {
  const int kind = h.arith == 1 ? KIND_GEMM_LP : (h.arith == 2 ? KIND_GEMM_F16 : KIND_GEMM_F64);
  auto gemm = [&](const GemmLaunch& g, double f, double b) {
    if (g.ntiles <= 0) return;
    if (h.arith == 2 && !g.presplit) {  // on-load split: the whole sequence
      const int ps = timed ? c.prof_begin() : -1;
      issue_gemm(c, g);
      if (ps >= 0) c.prof_end(ps, kind, f, b);
    } else if (h.arith == 2) {  // pre-passes, then the GEMM timed by its own dispatch
      const int pp = timed ? c.prof_begin() : -1;
      issue_gemm_f16_prep(c, g);
      if (pp >= 0) c.prof_end(pp, KIND_F16_PREP, 0.0, 0.0);
      const int ps = timed ? c.prof_begin_ext() : -1;
      issue_gemm_f16_main(c, g);
      c.prof_end_ext(ps, kind, f, b);
    } else {  // one kernel: timed by its own dispatch
      const int ps = timed ? c.prof_begin_ext() : -1;
      issue_gemm(c, g);
      c.prof_end_ext(ps, kind, f, b);
    }
  };
  gemm(h.a, h.fa, h.ba);
  if (h.m.nwork > 0) {
    const int ps = timed ? c.prof_begin() : -1;
    issue_mix(c, h.m);
    if (ps >= 0) c.prof_end(ps, KIND_MIX, h.m.flops, h.m.bytes);
  }
  gemm(h.c, h.fc, h.bc);
  c.flops_heff += h.flops_alg;
}

void run_heff(Context& c, const HeffPlan& p, int prec = DMRJ_PREC_F64) {
  issue_heff(c, prepare_heff(c, p, prec), true);
}

// Captures body() on the context stream into a graph and launches it; the
// executable graph of the previous call is updated in place when the
// topology matches (same kernels, new arguments), else re-instantiated.
template <class F>
void run_captured(Context& c, F&& body, double flops, int slot = 0) {
  HIP_OK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(c.stream, &g) == hipSuccess && g) hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamEndCapture(c.stream, &g));
  bool ok = false;
  if (c.graph_exec[slot]) {
    // the previous launch of this executable graph must have finished before
    // its kernel arguments are rewritten
    HIP_OK(hipEventSynchronize(c.graph_done[slot]));
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    ok = hipGraphExecUpdate(c.graph_exec[slot], g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!ok) {
      (void)hipGetLastError();
      hipGraphExecDestroy(c.graph_exec[slot]);
      c.graph_exec[slot] = nullptr;
    } else {
      c.graph_updates += 1;
    }
  }
  if (!ok) {
    hipError_t e = hipGraphInstantiate(&c.graph_exec[slot], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      c.graph_exec[slot] = nullptr;
      throw Error(DMRJ_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c.graph_instantiations += 1;
  }
  const int ps = c.prof_begin();
  hipError_t e = hipGraphLaunch(c.graph_exec[slot], c.stream);
  hipGraphDestroy(g);
  if (e != hipSuccess) throw Error(DMRJ_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
  HIP_OK(hipEventRecord(c.graph_done[slot], c.stream));
  if (ps >= 0) c.prof_end(ps, KIND_GRAPH, flops, 0.0);
}

int base_prec(int precision) {
  const int b = precision & 15;
  if (b != DMRJ_PREC_F64 && b != DMRJ_PREC_BF16X3 && b != DMRJ_PREC_F16X3)
    throw Error(DMRJ_ERR_ARG,
                "precision must be DMRJ_PREC_F64, _BF16X3 or _F16X3 (optionally | DMRJ_PREC_CORRECT)");
  return b;
}

// ======================================================== environment plans
struct EnvPlan {
  std::vector<GemmTask> t1, t3;
  std::vector<GemmSeg> s1, s3;
  std::vector<MixDst> md;
  std::vector<MixTerm> mt;
  double* X = nullptr;
  double* Y = nullptr;
  double flops = 0.0;
};

// Lupdate: Lout on A.right with D = W.DR.  The bra tensor is B (bra bond
// spaces B.left / B.right) and the ket tensor A; B == A for the environments
// of one state, where every step below is the one-state plan.
EnvPlan plan_left(Context& c, const EnvMeta& L, const double* Ld, const SiteMeta& A,
                  const double* Ad, const MpoSite& W, EnvMeta& Lo, double* Lod, int zmask = 0,
                  const SiteMeta* Bp = nullptr, const double* Bd = nullptr) {
  // zmask: bit 0 = L complex, bit 1 = A complex, bit 2 = B complex (then Lod
  // holds two planes)
  const SiteMeta& B = Bp ? *Bp : A;
  if (!Bp) Bd = Ad;
  const int zf = zmask ? 2 : 1;
  if (!L.ket.same(A.left) || !L.bra.same(B.left))
    throw Error(DMRJ_ERR_ARG, "left environment does not match the MPS tensors");
  if (L.D != W.DL || A.P != W.d || B.P != W.d) throw Error(DMRJ_ERR_ARG, "MPO does not match environment/MPS");
  EnvPlan p;
  const Space& S = A.left;   // ket sectors
  const Space& T = A.right;
  const Space& Sb = B.left;  // bra sectors
  const Space& Tb = B.right;
  const int nL = S.nsec(), nLb = Sb.nsec();
  // step 1 X(w,a) = L(ap,w) * Apanel(a)
  std::vector<int64_t> xoff((size_t)W.DL * nL, -1);
  int64_t nx = 0;
  for (int w = 0; w < W.DL; ++w)
    for (int a = 0; a < nL; ++a) {
      if (A.panel_cols[a] == 0) continue;
      const int ap = Sb.find(S.q[a] + W.dL[w]);
      if (ap < 0 || L.block(ap, w) < 0) continue;
      xoff[(size_t)w * nL + a] = nx;
      nx += (int64_t)Sb.dim[ap] * A.panel_cols[a];
    }
  // step 2 Y(ap,k,u,b): blocks n_ap x n_b (ap a bra sector, b a ket sector)
  std::vector<int64_t> yoff((size_t)nLb * B.P * W.DR, -1);
  int64_t ny = 0;
  for (int ap = 0; ap < nLb; ++ap)
    for (int k = 0; k < B.P; ++k) {
      const int ab = B.block(ap, k);
      if (ab < 0) continue;
      const int bp = B.blocks[ab].b;
      for (int u = 0; u < W.DR; ++u) {
        const int b = T.find(Tb.q[bp] - W.dR[u]);
        if (b < 0) continue;
        yoff[((size_t)ap * B.P + k) * W.DR + u] = ny;
        ny += (int64_t)Sb.dim[ap] * T.dim[b];
      }
    }
  p.X = (double*)c.alloc.alloc(std::max<int64_t>(nx, 1) * zf * sizeof(double));
  p.Y = (double*)c.alloc.alloc(std::max<int64_t>(ny, 1) * zf * sizeof(double));
  for (int w = 0; w < W.DL; ++w)
    for (int a = 0; a < nL; ++a) {
      const int64_t o = xoff[(size_t)w * nL + a];
      if (o < 0) continue;
      const int ap = Sb.find(S.q[a] + W.dL[w]);
      const EnvBlock& lb = L.blocks[L.block(ap, w)];
      p.t1.push_back({p.X + o, Sb.dim[ap], Sb.dim[ap], (int)A.panel_cols[a], (int)p.s1.size(), 1, 0.0});
      p.s1.push_back({Ld + lb.off, Ad + A.panel_off[a], Sb.dim[ap], S.dim[a], S.dim[a], 0, 1.0});
    }
  for (int ap = 0; ap < nLb; ++ap)
    for (int k = 0; k < B.P; ++k) {
      const int ab = B.block(ap, k);
      if (ab < 0) continue;
      const int bp = B.blocks[ab].b;
      for (int u = 0; u < W.DR; ++u) {
        const int64_t yo = yoff[((size_t)ap * B.P + k) * W.DR + u];
        if (yo < 0) continue;
        const int b = T.find(Tb.q[bp] - W.dR[u]);
        MixDst d{p.Y + yo, (int64_t)Sb.dim[ap] * T.dim[b], (int)p.mt.size(), 0};
        for (int w = 0; w < W.DL; ++w) {
          const int a = S.find(Sb.q[ap] - W.dL[w]);
          if (a < 0) continue;
          const int64_t xo = xoff[(size_t)w * nL + a];
          if (xo < 0) continue;
          for (int j = 0; j < A.P; ++j) {
            const int xb = A.block(a, j);
            if (xb < 0 || A.blocks[xb].b != b) continue;
            const double cf = W.at(w, j, k, u);
            if (cf == 0.0) continue;
            p.mt.push_back({p.X + xo + A.blocks[xb].colpos * Sb.dim[ap], cf});
            d.nt++;
          }
        }
        p.flops += 2.0 * (double)d.n * W.DL * A.P;
        p.md.push_back(d);
      }
    }
  // step 3: Lout(bp,u) = sum_k B(ap_k,k)^T * Y(ap_k,k,u,b)
  Lo.bra = Tb;
  Lo.ket = T;
  Lo.D = W.DR;
  Lo.delta = W.dR;
  Lo.build();
  for (const EnvBlock& ob : Lo.blocks) {
    const int bp = ob.ap, u = ob.w, b = ob.a;
    GemmTask t{Lod + ob.off, Tb.dim[bp], Tb.dim[bp], T.dim[b], (int)p.s3.size(), 0, 0.0};
    for (int k = 0; k < B.P; ++k) {
      const int ap = Sb.find(Tb.q[bp] - B.qp[k]);
      if (ap < 0) continue;
      const int ab = B.block(ap, k);
      if (ab < 0 || B.blocks[ab].b != bp) continue;
      const int64_t yo = yoff[((size_t)ap * B.P + k) * W.DR + u];
      if (yo < 0) continue;
      // bra tensor: enters complex-conjugated
      p.s3.push_back({Bd + B.blocks[ab].off, p.Y + yo, Sb.dim[ap], Sb.dim[ap], Sb.dim[ap], 1 | kSegConjA, 1.0});
      t.nseg++;
    }
    p.t3.push_back(t);
  }
  p.flops += gemm_flops(p.t1, p.s1) + gemm_flops(p.t3, p.s3);
  if (zmask) {
    BufMap bm;
    bm.add(Ld, L.size, zmask & 1);
    bm.add(Ad, A.size, (zmask & 2) != 0);
    if (Bd != Ad) bm.add(Bd, B.size, (zmask & 4) != 0);
    bm.add(p.X, nx, true);
    bm.add(p.Y, ny, true);
    bm.add(Lod, Lo.size, true);
    complexify_gemm(p.t1, p.s1, bm);
    complexify_mix(p.md, p.mt, bm);
    complexify_gemm(p.t3, p.s3, bm);
    p.flops *= 4.0;
  }
  return p;
}

// Rupdate: Rout (stored [bra, w, ket]) on A.left with D = W.DL; the bra
// tensor is B (B == A for the environments of one state).
EnvPlan plan_right(Context& c, const EnvMeta& R, const double* Rd, const SiteMeta& A,
                   const double* Ad, const MpoSite& W, EnvMeta& Ro, double* Rod, int zmask = 0,
                   const SiteMeta* Bp = nullptr, const double* Bd = nullptr) {
  // zmask: bit 0 = R complex, bit 1 = A complex, bit 2 = B complex (then Rod
  // holds two planes)
  const SiteMeta& B = Bp ? *Bp : A;
  if (!Bp) Bd = Ad;
  const int zf = zmask ? 2 : 1;
  if (!R.ket.same(A.right) || !R.bra.same(B.right))
    throw Error(DMRJ_ERR_ARG, "right environment does not match the MPS tensors");
  if (R.D != W.DR || A.P != W.d || B.P != W.d) throw Error(DMRJ_ERR_ARG, "MPO does not match environment/MPS");
  EnvPlan p;
  const Space& S = A.left;   // ket sectors
  const Space& T = A.right;
  const Space& Sb = B.left;  // bra sectors
  const Space& Tb = B.right;
  const int nLb = Sb.nsec();
  // step 1 X(ap,k) = B(ap,k) * Rpanel(bp): blocks over u as in R's panel
  std::vector<int64_t> xpan((size_t)nLb * B.P, -1);
  int64_t nx = 0;
  for (int ap = 0; ap < nLb; ++ap)
    for (int k = 0; k < B.P; ++k) {
      const int ab = B.block(ap, k);
      if (ab < 0) continue;
      const int bp = B.blocks[ab].b;
      if (R.panel_cols[bp] == 0) continue;
      xpan[(size_t)ap * B.P + k] = nx;
      nx += (int64_t)Sb.dim[ap] * R.panel_cols[bp];
    }
  // step 2 Y(ap,w): aligned with A's row panel of a = ap - dL[w]
  std::vector<int64_t> ypan((size_t)nLb * W.DL, -1);
  int64_t ny = 0;
  for (int ap = 0; ap < nLb; ++ap)
    for (int w = 0; w < W.DL; ++w) {
      const int a = S.find(Sb.q[ap] - W.dL[w]);
      if (a < 0) continue;
      ypan[(size_t)ap * W.DL + w] = ny;
      ny += (int64_t)Sb.dim[ap] * A.panel_cols[a];
    }
  p.X = (double*)c.alloc.alloc(std::max<int64_t>(nx, 1) * zf * sizeof(double));
  p.Y = (double*)c.alloc.alloc(std::max<int64_t>(ny, 1) * zf * sizeof(double));
  for (int ap = 0; ap < nLb; ++ap)
    for (int k = 0; k < B.P; ++k) {
      const int64_t o = xpan[(size_t)ap * B.P + k];
      if (o < 0) continue;
      const SiteBlock& ab = B.blocks[B.block(ap, k)];
      const int bp = ab.b;
      p.t1.push_back({p.X + o, Sb.dim[ap], Sb.dim[ap], (int)R.panel_cols[bp], (int)p.s1.size(), 1, 0.0});
      // bra tensor: enters complex-conjugated
      p.s1.push_back({Bd + ab.off, Rd + R.panel_off[bp], Sb.dim[ap], Tb.dim[bp], Tb.dim[bp], kSegConjA, 1.0});
    }
  for (int ap = 0; ap < nLb; ++ap)
    for (int w = 0; w < W.DL; ++w) {
      const int64_t yo = ypan[(size_t)ap * W.DL + w];
      if (yo < 0) continue;
      const int a = S.find(Sb.q[ap] - W.dL[w]);
      const int nap = Sb.dim[ap];
      for (int j = 0; j < A.P; ++j) {
        const int jb = A.block(a, j);
        if (jb < 0) continue;
        const SiteBlock& abj = A.blocks[jb];
        const int b = abj.b;
        MixDst d{p.Y + yo + abj.colpos * nap, (int64_t)nap * T.dim[b], (int)p.mt.size(), 0};
        for (int k = 0; k < B.P; ++k) {
          const int64_t xo = xpan[(size_t)ap * B.P + k];
          if (xo < 0) continue;
          const int bp = B.blocks[B.block(ap, k)].b;
          for (int u = 0; u < W.DR; ++u) {
            const int rbi = R.block(bp, u);
            if (rbi < 0 || R.blocks[rbi].a != b) continue;
            const double cf = W.at(w, j, k, u);
            if (cf == 0.0) continue;
            p.mt.push_back({p.X + xo + R.blocks[rbi].colpos * nap, cf});
            d.nt++;
          }
        }
        p.flops += 2.0 * (double)d.n * W.DR * A.P;
        p.md.push_back(d);
      }
    }
  Ro.bra = Sb;
  Ro.ket = S;
  Ro.D = W.DL;
  Ro.delta = W.dL;
  Ro.build();
  for (const EnvBlock& ob : Ro.blocks) {
    const int ap = ob.ap, w = ob.w, a = ob.a;
    GemmTask t{Rod + ob.off, Sb.dim[ap], Sb.dim[ap], S.dim[a], (int)p.s3.size(), 0, 0.0};
    const int64_t yo = ypan[(size_t)ap * W.DL + w];
    if (yo >= 0 && A.panel_cols[a] > 0) {
      p.s3.push_back({p.Y + yo, Ad + A.panel_off[a], Sb.dim[ap], S.dim[a], (int)A.panel_cols[a], 2, 1.0});
      t.nseg = 1;
    }
    p.t3.push_back(t);
  }
  p.flops += gemm_flops(p.t1, p.s1) + gemm_flops(p.t3, p.s3);
  if (zmask) {
    BufMap bm;
    bm.add(Rd, R.size, zmask & 1);
    bm.add(Ad, A.size, (zmask & 2) != 0);
    if (Bd != Ad) bm.add(Bd, B.size, (zmask & 4) != 0);
    bm.add(p.X, nx, true);
    bm.add(p.Y, ny, true);
    bm.add(Rod, Ro.size, true);
    complexify_gemm(p.t1, p.s1, bm);
    complexify_mix(p.md, p.mt, bm);
    complexify_gemm(p.t3, p.s3, bm);
    p.flops *= 4.0;
  }
  return p;
}

void run_env(Context& c, EnvPlan& p) {
  launch_gemm_arith(c, p.t1, p.s1, -1, c.env_prec);
  launch_mix(c, p.md, p.mt);
  launch_gemm_arith(c, p.t3, p.s3, -1, c.env_prec);
  c.flops_env += p.flops;
  c.alloc.release(p.X);
  c.alloc.release(p.Y);
}

// ================================================================ theta = AL*AR
dmrj_tensor* make_theta(Context& c, const Tensor& AL, const Tensor& AR) {
  if (AL.kind != 0 || AR.kind != 0 || AL.site.d1 != 0 || AR.site.d1 != 0)
    throw Error(DMRJ_ERR_ARG, "theta2 needs two MPS tensors");
  const SiteMeta& L = AL.site;
  const SiteMeta& R = AR.site;
  if (!L.right.same(R.left)) throw Error(DMRJ_ERR_ARG, "psiL right bond != psiR left bond");
  auto* t = new_tensor(c, 0);
  t->nq = AL.nq;
  SiteMeta& m = t->site;
  m.left = L.left;
  m.right = R.right;
  m.d1 = L.P;
  m.d2 = R.P;
  m.qp1 = L.qp;
  m.qp2 = R.qp;
  m.P = L.P * R.P;
  m.qp.resize(m.P);
  for (int j2 = 0; j2 < R.P; ++j2)
    for (int j1 = 0; j1 < L.P; ++j1) m.qp[j1 + L.P * j2] = L.qp[j1] + R.qp[j2];
  m.build();
  t->cplx = (AL.cplx || AR.cplx) ? 1 : 0;
  alloc_data(c, *t);
  std::vector<GemmTask> tasks;
  std::vector<GemmSeg> segs;
  for (const SiteBlock& b : m.blocks) {
    const int j1 = b.s % L.P, j2 = b.s / L.P;
    GemmTask tk{t->data + b.off, m.left.dim[b.a], m.left.dim[b.a], m.right.dim[b.b], (int)segs.size(), 0, 0.0};
    const int lb = L.block(b.a, j1);
    if (lb >= 0) {
      const int cc = L.blocks[lb].b;
      const int rb = R.block(cc, j2);
      if (rb >= 0 && R.blocks[rb].b == b.b) {
        segs.push_back({AL.data + L.blocks[lb].off, AR.data + R.blocks[rb].off, m.left.dim[b.a],
                        L.right.dim[cc], L.right.dim[cc], 0, 1.0});
        tk.nseg = 1;
      }
    }
    tasks.push_back(tk);
  }
  if (t->cplx) {
    BufMap bm;
    bm.add(AL.data, L.size, AL.cplx != 0);
    bm.add(AR.data, R.size, AR.cplx != 0);
    bm.add(t->data, m.size, true);
    complexify_gemm(tasks, segs, bm);
  }
  launch_gemm(c, tasks, segs, -1);
  return t;
}

// ======================================================================= SVD


// ================================================================== Lanczos
struct LanczosOut {
  dmrj_tensor* AA;
};

dmrj_tensor* lanczos2(Context& c, const Tensor& L, const Tensor& R, const Tensor& AL,
                      const Tensor& AR, const MpoSite& W1, const MpoSite& W2, int precision) {
  const int prec = base_prec(precision);
  const bool correct = (precision & DMRJ_PREC_CORRECT) != 0;
  if (L.kind != 1 || R.kind != 2) throw Error(DMRJ_ERR_ARG, "lanczos2: bad environment handles");
  if (L.cplx || R.cplx || AL.cplx || AR.cplx) throw Error(DMRJ_ERR_ARG, "lanczos2 takes real tensors");
  dmrj_tensor* th = make_theta(c, AL, AR);
  const SiteMeta& x = th->site;
  const int64_t n = x.size;
  FusedMpo f = fuse_two(W1, W2);
  double* y1 = (double*)c.alloc.alloc(std::max<int64_t>(n, 1) * sizeof(double));
  double* y2 = (double*)c.alloc.alloc(std::max<int64_t>(n, 1) * sizeof(double));
  double *T1 = nullptr, *T3 = nullptr;
  HeffPlan p = plan_heff(c, x, th->data, y1, L.env, L.data, R.env, R.data, f, &T1, &T3);
  // descriptor tables of every H_eff application of the step go to the
  // device first, so the step itself is a fixed kernel sequence
  HeffLaunch h1 = prepare_heff(c, p, prec);          // H psi
  rebase_plan(p, th->data, y1, n, y1, y2, n);
  HeffLaunch h2 = prepare_heff(c, p, prec);          // H psi2
  HeffLaunch h3;
  if (correct) {
    rebase_plan(p, y1, y2, n, y2, y1, n);
    h3 = prepare_heff(c, p, DMRJ_PREC_F64);          // H AA' in FP64
  }
  const bool graph = c.graph_mode != 0;
  auto step = [&]() {
    issue_heff(c, h1, !graph);
    {
      const double* xs[2] = {th->data, th->data};
      const double* ys[2] = {th->data, y1};
      launch_dots(c, 2, xs, ys, n, c.scal + S_TT);
    }
    launch_lz_scalar(c, 1);
    launch_axpy_dev(c, y1, th->data, n, c.scal + S_CAXPY);  // y1 := psi2
    {
      const double* xs[1] = {y1};
      launch_dots(c, 1, xs, xs, n, c.scal + S_PP);
    }
    launch_lz_scalar(c, 2);
    launch_scale_dev(c, y1, n, c.scal + S_INVB);
    issue_heff(c, h2, !graph);
    {
      const double* xs[1] = {y1};
      const double* ys[1] = {y2};
      launch_dots(c, 1, xs, ys, n, c.scal + S_PHP);
    }
    launch_lz_scalar(c, 3);
    launch_comb_dev(c, y2, th->data, y1, n, c.scal + S_CA, c.scal + S_CB);
    if (correct) {
      // FP64 residual correction of the energy: Rayleigh quotient of the Ritz
      // vector with an FP64 H_eff application; the residual norm
      // |H AA - E AA| is kept in S_RES
      issue_heff(c, h3, !graph);
      {
        const double* xs[2] = {y2, y2};
        const double* ys[2] = {y2, y1};
        launch_dots(c, 2, xs, ys, n, c.scal + S_RR);
      }
      launch_lz_scalar(c, 4);
      launch_axpy_dev(c, y1, y2, n, c.scal + S_NEGE);
      {
        const double* xs[1] = {y1};
        launch_dots(c, 1, xs, xs, n, c.scal + S_RES);
      }
    }
  };
  if (graph)
    run_captured(c, step, h1.flops_alg + h2.flops_alg + h3.flops_alg);
  else
    step();
  c.alloc.release(T1);
  c.alloc.release(T3);
  c.alloc.release(y1);
  c.alloc.release(th->data);
  th->data = y2;
  return th;
}

// ============================================================ Krylov exponential
// Lanczos with r vectors and full reorthogonalisation (FP64 dots) on x,
// then out = sum_j c_j v_j with c = exp(pf T) e1 (TensorPACK lanczos +
// exp(M, prefactor) as used by krylovTimeEvol, algs/time.jl:182-202); the
// tridiagonal eigensolve and the coefficients run on the device
// (k_krylov_coef).  Returns the lowest Ritz value in *energy and the Krylov
// dimension used when those pointers are given (one stream synchronisation);
// with both NULL the call enqueues work only.
// slot: graph slot of the iteration's kernel sequence (1: two-site, 2: one-site)
// ritz: out is the lowest Ritz vector instead (c = eigenvector of T of the
// lowest eigenvalue, signed so that c_1 = <v_1, out> >= 0; the prefactor is
// ignored), the eigensolver of TensorPACK lanczos as 3S calls it
// (algs/DMRG.jl:328)
dmrj_tensor* krylov_expm(Context& c, const Tensor& L, const Tensor& R, const Tensor& x, const FusedMpo& f, int r,
                         double pf_re, double pf_im, int prec, double* energy, int* kdim, int slot,
                         bool ritz = false) {
  if (ritz) pf_re = pf_im = 0.0;
  if (r < 1 || r > 64) throw Error(DMRJ_ERR_ARG, "Krylov dimension must be 1..64");
  const SiteMeta& m = x.site;
  const int64_t n = std::max<int64_t>(m.size, 1);
  // Krylov vectors are complex when the start vector or an environment is;
  // a complex prefactor alone only makes the final combination complex
  const bool zv = x.cplx || L.cplx || R.cplx;
  const bool zout = zv || pf_im != 0.0;
  const int64_t nv = zv ? 2 * n : n;         // stride of one Krylov vector
  const int64_t ne = zv ? 2 * m.size : m.size;  // stored elements of one vector
  double* V = (double*)c.alloc.alloc((size_t)nv * r * sizeof(double));
  double* w = (double*)c.alloc.alloc((size_t)nv * sizeof(double));
  double* xin = (double*)c.alloc.alloc((size_t)nv * sizeof(double));  // fixed input of the prepared H_eff
  double* sc = (double*)c.alloc.alloc(256 * sizeof(double));  // alpha[0..63], beta[64..127], work[128..]
  double* gs = (double*)c.alloc.alloc((size_t)(2 * 64 * 256 + 128) * sizeof(double));  // CGS partials, coefs
  launch_fill(c, sc, 256, 0.0);
  // v0 = x / |x|
  {
    const double* xs[1] = {x.data};
    launch_dots(c, 1, xs, xs, x.nstore(), c.scal + S_OC);
    launch_lz_scalar(c, 0);
    if (zv && !x.cplx) launch_fill(c, V + n, m.size, 0.0);
    HIP_OK(hipMemcpyAsync(V, x.data, x.nstore() * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    launch_scale_dev(c, V, ne, c.scal + S_INVOC);
  }
  double *T1 = nullptr, *T3 = nullptr;
  const int zmask = zv ? ((L.cplx ? 1 : 0) | 2 | (R.cplx ? 4 : 0)) : 0;
  HIP_OK(hipMemcpyAsync(xin, V, ne * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  HeffPlan p = plan_heff(c, m, xin, w, L.env, L.data, R.env, R.data, f, &T1, &T3, zmask);
  const HeffLaunch hl = prepare_heff(c, p, prec);  // descriptor tables stay valid until the final sync
  // the whole iteration is a fixed kernel sequence on device scalars: with
  // graphs on it is captured and launched as one graph
  const bool graph = c.graph_mode != 0;
  // short vectors: everything after H_eff in one single-work-group launch,
  // which also writes v_{j+1} into the fixed input buffer
  const bool fused = ne <= kKrylovFusedMax;
  auto iterate = [&]() {
  for (int j = 0; j < r; ++j) {
    double* vj = V + (size_t)j * nv;
    if (j > 0 && !fused) HIP_OK(hipMemcpyAsync(xin, vj, ne * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    issue_heff(c, hl, !graph);
    if (fused) {
      launch_krylov_tail(c, V, nv, j, w, m.size, zv ? 1 : 0, sc, xin, j == r - 1 ? 1 : 0);
      continue;
    }
    // alpha_j = Re <v_j, w> (real for a Hermitian H_eff)
    {
      const double* xs[1] = {vj};
      const double* ys[1] = {w};
      launch_dots(c, 1, xs, ys, ne, sc + j);
    }
    if (j == r - 1) break;
    // w -= <v_i, w> v_i for all i <= j: two block passes of classical
    // Gram-Schmidt (this removes alpha_j v_j and beta_{j-1} v_{j-1} as well)
    for (int pass = 0; pass < 2; ++pass)
      launch_cgs_pass(c, w, V, nv, m.size, j + 1, zv ? 1 : 0, gs, gs + 2 * 64 * 256);
    {
      const double* xs[1] = {w};
      launch_dots(c, 1, xs, xs, ne, sc + 130);
    }
    // beta_j = |w|, v_{j+1} = w / beta_j (zero when the Krylov space is exhausted)
    HIP_OK(hipMemcpyAsync(c.scal + S_PP, sc + 130, sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    launch_lz_scalar(c, 2);
    HIP_OK(hipMemcpyAsync(sc + 64 + j, c.scal + S_B1, sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    double* vn = V + (size_t)(j + 1) * nv;
    HIP_OK(hipMemcpyAsync(vn, w, ne * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    launch_scale_dev(c, vn, ne, c.scal + S_INVB);
  }
  };
  if (graph)
    run_captured(c, iterate, r * hl.flops_alg, slot);
  else
    iterate();
  // the tridiagonal eigensolve and the exponential (or Ritz) coefficients on
  // the device: no host round trip unless the caller asks for the energy or
  // the Krylov dimension (sc[128..]: coefficients, sc[200..202]: outputs)
  double* dcoef = (double*)c.alloc.alloc((size_t)4 * r * sizeof(double));
  double* douts = sc + 200;
  launch_krylov_coef(c, sc, r, pf_re, pf_im, ritz ? 1 : 0, zv ? 1 : 0, dcoef, douts);
  auto* out = new dmrj_tensor();
  out->kind = 0;
  out->nq = x.nq;
  out->site = m;
  out->cplx = zout ? 1 : 0;
  alloc_data(c, *out);
  if (!zv) {
    launch_combine(c, out->data, V, n, m.size, dcoef, r);
    if (zout) launch_combine(c, out->data + m.size, V, n, m.size, dcoef + r, r);
  } else {
    // planes of the Krylov basis: plane 2j = Re v_j, plane 2j+1 = Im v_j, stride n
    launch_combine(c, out->data, V, n, m.size, dcoef, 2 * r);
    launch_combine(c, out->data + m.size, V, n, m.size, dcoef + 2 * r, 2 * r);
  }
  double hout[3] = {0.0, (double)r, 0.0};
  const bool need = energy != nullptr || kdim != nullptr;
  if (need) {
    HIP_OK(hipMemcpyAsync(hout, douts, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    c.release_transients();
    c.stage_reset();
  }
  c.alloc.release(T1);
  c.alloc.release(T3);
  c.alloc.release(V);
  c.alloc.release(w);
  c.alloc.release(xin);
  c.alloc.release(gs);
  c.alloc.release(sc);
  c.alloc.release(dcoef);
  if (need && hout[2] != 0.0) {
    dmrj_tensor* o = out;
    c.alloc.release(o->data);
    delete o;
    throw Error(DMRJ_ERR_NUMERIC, "Krylov: non-finite value");
  }
  if (energy) *energy = hout[0];
  if (kdim) *kdim = (int)hout[1];
  return out;
}

// ============================================== 3S subspace expansion and move
// Lexpand / Rexpand (algs/DMRG.jl:273-286) followed by joinindex!: the site
// tensor A joined with alpha times the environment-side half of H_1site A
// along the bond the centre moves across.  The result is laid out as a
// two-site tensor with one trivial physical index, so that the truncated SVD
// of the two-site path splits it as moveR! / moveL! do:
//   dir > 0: [a, s, 1, b'] with b' = b + sum_u (u, b), the new column (u, b)
//            carrying charge q_b + dR[u];  P[x,k,(u,b)] = sum L[x,w,a] A[a,j,b] W[w,j,k,u]
//   dir < 0: [a', 1, s, b] with a' = a + sum_w (a, w), the new row (a, w)
//            carrying charge q_a + dL[w];  P[(a,w),k,B] = sum W[w,j,k,u] A[a,j,b] R[b,u,B]
// Inside every charge sector of the enlarged bond the original states come
// first, so "restricted to the original bond" is a prefix of each sector.
// Groups (u, b) / (a, w) that no MPO coefficient reaches are left out (they
// would be zero columns / rows).  alpha == 0: A alone in the same layout.
dmrj_tensor* expand3s(Context& c, const Tensor& E, const Tensor& A, const MpoSite& W, double alpha, int dir) {
  if (A.kind != 0 || A.site.d1 || A.cplx || E.cplx)
    throw Error(DMRJ_ERR_ARG, "expand3s: needs a real one-site tensor and a real environment");
  const SiteMeta& x = A.site;
  const Space& S = x.left;
  const Space& T = x.right;
  if (x.P != W.d) throw Error(DMRJ_ERR_ARG, "expand3s: MPO does not match the site tensor");
  const bool grow = alpha != 0.0;
  std::vector<GemmTask> t1;
  std::vector<GemmSeg> s1;
  std::vector<MixDst> md;
  std::vector<MixTerm> mt;
  struct Piece {  // one expansion block: Y block -> (tensor block, offset inside the enlarged sector)
    int64_t yo;
    int blk_a, blk_s;   // block (left sector, physical state) of the output tensor
    int group;          // expansion group
    int rows, cols;
  };
  std::vector<Piece> pieces;
  double* X = nullptr;
  double* Y = nullptr;
  auto* t = new_tensor(c, 0);
  t->nq = A.nq;
  SiteMeta& m = t->site;
  std::vector<CopyOp> ops;
  try {
    if (dir > 0) {
      if (E.kind != 1) throw Error(DMRJ_ERR_ARG, "expand3s: moving right needs the left environment");
      const EnvMeta& L = E.env;
      if (!L.ket.same(S) || !L.bra.same(S) || L.D != W.DL)
        throw Error(DMRJ_ERR_ARG, "expand3s: left environment does not match the site tensor / MPO");
      const int nL = S.nsec(), nR = T.nsec();
      std::vector<int64_t> gofs((size_t)W.DR * nR, -1);   // group (u, b): offset inside its enlarged sector
      std::map<Charge, int64_t> dimQ;
      for (int b = 0; b < nR; ++b) dimQ[T.q[b]] += T.dim[b];
      if (grow) {
        // X(w,a) = L(ap,w) * Apanel(a);  Y(ap,k,u,b) = sum_{w,j} W(w,j,k,u) X(w,a)[., (j,b)]
        std::vector<int64_t> xoff((size_t)W.DL * nL, -1);
        int64_t nx = 0;
        for (int w = 0; w < W.DL; ++w)
          for (int a = 0; a < nL; ++a) {
            if (x.panel_cols[a] == 0) continue;
            const int ap = S.find(S.q[a] + W.dL[w]);
            if (ap < 0 || L.block(ap, w) < 0) continue;
            xoff[(size_t)w * nL + a] = nx;
            nx += (int64_t)S.dim[ap] * x.panel_cols[a];
          }
        int64_t ny = 0;
        std::vector<char> used((size_t)W.DR * nR, 0);
        for (int ap = 0; ap < nL; ++ap)
          for (int k = 0; k < x.P; ++k) {
            // the row (x, k) may reach charges the original bond does not
            // hold: new sectors of the enlarged bond
            const Charge Q = S.q[ap] + x.qp[k];
            for (int u = 0; u < W.DR; ++u) {
              const int b = T.find(Q - W.dR[u]);
              if (b < 0) continue;
              MixDst d{nullptr, (int64_t)S.dim[ap] * T.dim[b], (int)mt.size(), 0};
              for (int w = 0; w < W.DL; ++w) {
                const int a = S.find(S.q[ap] - W.dL[w]);
                if (a < 0) continue;
                const int64_t xo = xoff[(size_t)w * nL + a];
                if (xo < 0) continue;
                for (int j = 0; j < x.P; ++j) {
                  const int xb = x.block(a, j);
                  if (xb < 0 || x.blocks[xb].b != b) continue;
                  const double cf = W.at(w, j, k, u);
                  if (cf == 0.0) continue;
                  mt.push_back({(const double*)(intptr_t)(xo + x.blocks[xb].colpos * S.dim[ap]), cf});
                  d.nt++;
                }
              }
              if (d.nt == 0) {
                mt.resize(d.t0);
                continue;
              }
              d.dst = (double*)(intptr_t)ny;
              md.push_back(d);
              pieces.push_back({ny, ap, k, u * nR + b, S.dim[ap], T.dim[b]});
              used[(size_t)u * nR + b] = 1;
              ny += d.n;
            }
          }
        X = (double*)c.alloc.alloc(std::max<int64_t>(nx, 1) * sizeof(double));
        Y = (double*)c.alloc.alloc(std::max<int64_t>(ny, 1) * sizeof(double));
        for (auto& d : md) d.dst = Y + (intptr_t)d.dst;
        for (auto& q : mt) q.src = X + (intptr_t)q.src;
        for (int w = 0; w < W.DL; ++w)
          for (int a = 0; a < nL; ++a) {
            const int64_t o = xoff[(size_t)w * nL + a];
            if (o < 0) continue;
            const int ap = S.find(S.q[a] + W.dL[w]);
            const EnvBlock& lb = L.blocks[L.block(ap, w)];
            t1.push_back({X + o, S.dim[ap], S.dim[ap], (int)x.panel_cols[a], (int)s1.size(), 1, 0.0});
            s1.push_back({E.data + lb.off, A.data + x.panel_off[a], S.dim[ap], S.dim[a], S.dim[a], 0, 1.0});
          }
        // groups in the reference's column order: b outer, u inner
        for (int b = 0; b < nR; ++b)
          for (int u = 0; u < W.DR; ++u) {
            if (!used[(size_t)u * nR + b]) continue;
            int64_t& cur = dimQ[T.q[b] + W.dR[u]];
            gofs[(size_t)u * nR + b] = cur;
            cur += T.dim[b];
          }
      }
      Space Tx;
      for (auto& kv : dimQ) {
        Tx.q.push_back(kv.first);
        Tx.dim.push_back((int)kv.second);
      }
      Tx.finish();
      m.left = S;
      m.right = Tx;
      m.d1 = x.P;
      m.d2 = 1;
      m.qp1 = x.qp;
      m.qp2 = std::vector<Charge>(1);
      m.P = x.P;
      m.qp = x.qp;
      m.build();
      alloc_data(c, *t);
      launch_fill(c, t->data, m.size, 0.0);
      for (const SiteBlock& ab : x.blocks) {
        const int tb = m.block(ab.a, ab.s);
        if (tb < 0 || m.right.q[m.blocks[tb].b] != T.q[ab.b]) throw Error(DMRJ_ERR_STATE, "expand3s: block table");
        ops.push_back({A.data + ab.off, t->data + m.blocks[tb].off, S.dim[ab.a], S.dim[ab.a], S.dim[ab.a],
                       T.dim[ab.b], nullptr, nullptr, 1.0});
      }
      for (const Piece& p : pieces) {
        const int tb = m.block(p.blk_a, p.blk_s);
        if (tb < 0) throw Error(DMRJ_ERR_STATE, "expand3s: expansion block missing");
        const int64_t co = gofs[p.group];
        ops.push_back({Y + p.yo, t->data + m.blocks[tb].off + co * p.rows, p.rows, p.rows, p.rows, p.cols, nullptr,
                       nullptr, alpha});
      }
    } else {
      if (E.kind != 2) throw Error(DMRJ_ERR_ARG, "expand3s: moving left needs the right environment");
      const EnvMeta& R = E.env;
      if (!R.ket.same(T) || !R.bra.same(T) || R.D != W.DR)
        throw Error(DMRJ_ERR_ARG, "expand3s: right environment does not match the site tensor / MPO");
      const int nL = S.nsec();
      std::vector<int64_t> gofs((size_t)W.DL * nL, -1);   // group (a, w): row offset inside its enlarged sector
      std::map<Charge, int64_t> dimQ;
      for (int a = 0; a < nL; ++a) dimQ[S.q[a]] += S.dim[a];
      if (grow) {
        // X(a,j,u) = A(a,j) * Rs(B,u)^T, B = q_b + dR[u];  Y(a,w,k) = sum_{j,u} W(w,j,k,u) X(a,j,u)
        std::vector<int64_t> xoff((size_t)nL * x.P * W.DR, -1);
        std::vector<int> xB((size_t)nL * x.P * W.DR, -1);
        int64_t nx = 0;
        for (int a = 0; a < nL; ++a)
          for (int j = 0; j < x.P; ++j) {
            const int ab = x.block(a, j);
            if (ab < 0) continue;
            const int b = x.blocks[ab].b;
            for (int u = 0; u < W.DR; ++u) {
              const int B = T.find(T.q[b] + W.dR[u]);
              if (B < 0) continue;
              const int rb = R.block(B, u);
              if (rb < 0 || R.blocks[rb].a != b) continue;
              const size_t key = ((size_t)a * x.P + j) * W.DR + u;
              xoff[key] = nx;
              xB[key] = B;
              nx += (int64_t)S.dim[a] * T.dim[B];
            }
          }
        int64_t ny = 0;
        std::vector<char> used((size_t)W.DL * nL, 0);
        for (int a = 0; a < nL; ++a)
          for (int w = 0; w < W.DL; ++w)
            for (int k = 0; k < x.P; ++k) {
              const int B = T.find(S.q[a] + W.dL[w] + x.qp[k]);
              if (B < 0) continue;
              MixDst d{nullptr, (int64_t)S.dim[a] * T.dim[B], (int)mt.size(), 0};
              for (int j = 0; j < x.P; ++j)
                for (int u = 0; u < W.DR; ++u) {
                  const size_t key = ((size_t)a * x.P + j) * W.DR + u;
                  if (xoff[key] < 0 || xB[key] != B) continue;
                  const double cf = W.at(w, j, k, u);
                  if (cf == 0.0) continue;
                  mt.push_back({(const double*)(intptr_t)xoff[key], cf});
                  d.nt++;
                }
              if (d.nt == 0) {
                mt.resize(d.t0);
                continue;
              }
              d.dst = (double*)(intptr_t)ny;
              md.push_back(d);
              pieces.push_back({ny, a, k, w * nL + a, S.dim[a], T.dim[B]});
              used[(size_t)w * nL + a] = 1;
              ny += d.n;
            }
        X = (double*)c.alloc.alloc(std::max<int64_t>(nx, 1) * sizeof(double));
        Y = (double*)c.alloc.alloc(std::max<int64_t>(ny, 1) * sizeof(double));
        for (auto& d : md) d.dst = Y + (intptr_t)d.dst;
        for (auto& q : mt) q.src = X + (intptr_t)q.src;
        for (int a = 0; a < nL; ++a)
          for (int j = 0; j < x.P; ++j) {
            const int ab = x.block(a, j);
            if (ab < 0) continue;
            const int b = x.blocks[ab].b;
            for (int u = 0; u < W.DR; ++u) {
              const size_t key = ((size_t)a * x.P + j) * W.DR + u;
              if (xoff[key] < 0) continue;
              const int B = xB[key];
              const EnvBlock& rb = R.blocks[R.block(B, u)];
              // X[n_a x n_B] = A(a,j)[n_a x n_b] * (Rs(B,u)[n_B x n_b])^T
              t1.push_back({X + xoff[key], S.dim[a], S.dim[a], T.dim[B], (int)s1.size(), 1, 0.0});
              s1.push_back({A.data + x.blocks[ab].off, E.data + rb.off, S.dim[a], T.dim[B], T.dim[b], 2, 1.0});
            }
          }
        // groups in the reference's row order: w outer, a inner
        for (int w = 0; w < W.DL; ++w)
          for (int a = 0; a < nL; ++a) {
            if (!used[(size_t)w * nL + a]) continue;
            int64_t& cur = dimQ[S.q[a] + W.dL[w]];
            gofs[(size_t)w * nL + a] = cur;
            cur += S.dim[a];
          }
      }
      Space Sx;
      for (auto& kv : dimQ) {
        Sx.q.push_back(kv.first);
        Sx.dim.push_back((int)kv.second);
      }
      Sx.finish();
      m.left = Sx;
      m.right = T;
      m.d1 = 1;
      m.d2 = x.P;
      m.qp1 = std::vector<Charge>(1);
      m.qp2 = x.qp;
      m.P = x.P;
      m.qp = x.qp;
      m.build();
      alloc_data(c, *t);
      launch_fill(c, t->data, m.size, 0.0);
      for (const SiteBlock& ab : x.blocks) {
        const int ax = Sx.find(S.q[ab.a]);
        const int tb = ax < 0 ? -1 : m.block(ax, ab.s);
        if (tb < 0 || m.blocks[tb].b != ab.b) throw Error(DMRJ_ERR_STATE, "expand3s: block table");
        ops.push_back({A.data + ab.off, t->data + m.blocks[tb].off, S.dim[ab.a], Sx.dim[ax], S.dim[ab.a],
                       T.dim[ab.b], nullptr, nullptr, 1.0});
      }
      for (const Piece& p : pieces) {
        const int a = p.blk_a, k = p.blk_s, w = p.group / nL;
        const int ax = Sx.find(S.q[a] + W.dL[w]);
        const int tb = ax < 0 ? -1 : m.block(ax, k);
        if (tb < 0) throw Error(DMRJ_ERR_STATE, "expand3s: expansion block missing");
        ops.push_back({Y + p.yo, t->data + m.blocks[tb].off + gofs[p.group], p.rows, Sx.dim[ax], p.rows, p.cols,
                       nullptr, nullptr, alpha});
      }
    }
    if (grow) {
      launch_gemm(c, t1, s1, -1);
      launch_mix(c, md, mt);
    }
    launch_copy(c, ops);
  } catch (...) {
    c.alloc.release(X);
    c.alloc.release(Y);
    if (t->data) c.alloc.release(t->data);
    delete t;
    throw;
  }
  c.flops_env += gemm_flops(t1, s1);
  // the tables above are staged asynchronously; X and Y go back to the pool,
  // which hands them out again only to later work on the same stream
  c.alloc.release(X);
  c.alloc.release(Y);
  return t;
}

// moveR! / moveL! with condition = true (lib/MPS/move.jl:124-139,175-190):
// the new centre is the neighbour contracted with the SVD factor restricted
// to the original bond (a prefix of every sector of the enlarged bond).
//   dir > 0: F = D V [k, 1, b'], N = psi[iR] [b, s, c] -> [k, s, c]
//   dir < 0: F = U D [a', 1, k], N = psi[iL] [aL, s, a] -> [aL, s, k]
dmrj_tensor* absorb3s(Context& c, const Tensor& F, const Tensor& N, int dir) {
  const SiteMeta& f = F.site;
  const SiteMeta& n = N.site;
  if (F.kind || N.kind || f.d1 || n.d1 || F.cplx || N.cplx || f.P != 1)
    throw Error(DMRJ_ERR_ARG, "absorb3s: needs a real SVD factor with one trivial index and a real site tensor");
  auto* t = new_tensor(c, 0);
  t->nq = N.nq;
  SiteMeta& m = t->site;
  std::vector<GemmTask> tasks;
  std::vector<GemmSeg> segs;
  if (dir > 0) {
    m.left = f.left;
    m.right = n.right;
  } else {
    m.left = n.left;
    m.right = f.right;
  }
  m.P = n.P;
  m.qp = n.qp;
  m.build();
  alloc_data(c, *t);
  try {
    if (dir > 0) {
      for (int K = 0; K < m.left.nsec(); ++K) {
        if (m.panel_cols[K] == 0) continue;
        const int dK = m.left.dim[K];
        GemmTask tk{t->data + m.panel_off[K], dK, dK, (int)m.panel_cols[K], (int)segs.size(), 0, 0.0};
        const int fb = f.block(K, 0);
        const int b = n.left.find(m.left.q[K]);
        if (fb >= 0 && b >= 0 && n.panel_cols[b] > 0) {
          const int bx = f.blocks[fb].b;
          if (n.panel_cols[b] != m.panel_cols[K] || f.right.dim[bx] < n.left.dim[b] ||
              f.right.q[bx] != n.left.q[b])
            throw Error(DMRJ_ERR_ARG, "absorb3s: factor and neighbour do not share the original bond");
          segs.push_back({F.data + f.blocks[fb].off, N.data + n.panel_off[b], dK, n.left.dim[b], n.left.dim[b], 0,
                          1.0});
          tk.nseg = 1;
        }
        tasks.push_back(tk);
      }
    } else {
      for (const SiteBlock& ob : m.blocks) {
        const int daL = m.left.dim[ob.a];
        GemmTask tk{t->data + ob.off, daL, daL, m.right.dim[ob.b], (int)segs.size(), 0, 0.0};
        const Charge Q = m.left.q[ob.a] + n.qp[ob.s];
        const int nb = n.block(ob.a, ob.s);
        const int ax = f.left.find(Q);
        const int fb = ax < 0 ? -1 : f.block(ax, 0);
        if (nb >= 0 && fb >= 0) {
          const int a = n.blocks[nb].b;
          if (f.blocks[fb].b != ob.b || f.left.dim[ax] < n.right.dim[a])
            throw Error(DMRJ_ERR_ARG, "absorb3s: factor and neighbour do not share the original bond");
          segs.push_back({N.data + n.blocks[nb].off, F.data + f.blocks[fb].off, daL, f.left.dim[ax], n.right.dim[a],
                          0, 1.0});
          tk.nseg = 1;
        }
        tasks.push_back(tk);
      }
    }
  } catch (...) {
    c.alloc.release(t->data);
    delete t;
    throw;
  }
  launch_gemm(c, tasks, segs, -1);
  return t;
}

double read_scalar(Context& c, int slot) {
  HIP_OK(hipMemcpyAsync(c.hscal, c.scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  c.release_transients();
  c.stage_reset();
  return c.hscal[slot];
}

template <class F>
int guard(dmrj_ctx* ctx, F&& f) {
  if (!ctx) return DMRJ_ERR_ARG;
  try {
    HIP_OK(hipSetDevice(ctx->device));
    f();
    return DMRJ_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return DMRJ_ERR_STATE;
  }
}

std::vector<Charge> charges_vec(const int32_t* q, int nq, int64_t n) {
  std::vector<Charge> v(n);
  for (int64_t i = 0; i < n; ++i) v[i] = nq ? charge_of(q, nq, i) : Charge();
  return v;
}

}  // namespace

// Scratch of dmrj_dots_noalloc: the per-work-group partial sums of up to
// kNoAllocPairs inner products (32 KiB).  Every GPU the library runs on holds
// its own copy until the process ends.  Host code never passes the name
// `scratch` to a kernel: it is a host-side stand-in; the device address comes
// from hipGetSymbolAddress, looked up once per device at context creation.
constexpr int kNoAllocPairs = 16;
static __device__ __align__(256) double scratch[kNoAllocPairs * kDotPartials];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (which must be `device`),
// looked up on the first call for that device.
double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(DMRJ_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(DMRJ_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)p;
  }
  return g_scratch_ptr[device];
}
}  // namespace

// =================================================================== C ABI
extern "C" {

int dmrj_version(void) { return 100; }

int dmrj_ctx_create(int device, dmrj_ctx** out) {
  if (!out) return DMRJ_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DMRJ_ERR_DEVICE;
  if (device < 0 || device >= n) return DMRJ_ERR_ARG;
  auto* c = new dmrj_ctx();
  c->device = device;
  try {
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&c->side_in, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&c->side_out, hipEventDisableTiming));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlock;
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocblas handle");
    rocblas_set_stream(h, c->stream);
    c->rocblas = h;
    HIP_OK(hipMalloc((void**)&c->scal, S_COUNT * sizeof(double)));
    HIP_OK(hipMemset(c->scal, 0, S_COUNT * sizeof(double)));
    HIP_OK(hipHostMalloc((void**)&c->hscal, S_COUNT * sizeof(double), hipHostMallocDefault));
    HIP_OK(hipMalloc((void**)&c->partial, 4 * kDotPartials * sizeof(double)));
    c->noalloc = noalloc_scratch(device);
    c->stage.cap = 8 << 20;
    HIP_OK(hipHostMalloc((void**)&c->stage.base, c->stage.cap, hipHostMallocDefault));
    for (auto& e : c->graph_done) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  } catch (const Error& e) {
    delete c;
    return e.code;
  }
  *out = c;
  return DMRJ_OK;
}

int dmrj_ctx_destroy(dmrj_ctx* c) {
  if (!c) return DMRJ_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  c->release_transients();
  if (c->side_in) (void)hipEventDestroy(c->side_in);
  if (c->side_out) (void)hipEventDestroy(c->side_out);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (auto& g : c->graph_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& e : c->graph_done)
    if (e) (void)hipEventDestroy(e);
  if (c->rocblas) rocblas_destroy_handle((rocblas_handle)c->rocblas);
  if (c->scal) (void)hipFree(c->scal);
  if (c->partial) (void)hipFree(c->partial);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hsig) (void)hipHostFree(c->hsig);
  if (c->stage.base) (void)hipHostFree(c->stage.base);
  c->alloc.trim();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DMRJ_OK;
}

const char* dmrj_last_error(const dmrj_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dmrj_ctx_sync(dmrj_ctx* c) {
  return guard(c, [&] {
    HIP_OK(hipStreamSynchronize(c->stream));
    c->release_transients();
    c->stage_reset();
  });
}

int dmrj_ctx_memstats(dmrj_ctx* c, int64_t* out3) {
  return guard(c, [&] {
    out3[0] = (int64_t)c->alloc.bytes_live;
    out3[1] = (int64_t)c->alloc.bytes_cached;
    out3[2] = (int64_t)c->alloc.peak;
  });
}

int dmrj_ctx_set_env_precision(dmrj_ctx* c, int precision) {
  return guard(c, [&] { c->env_prec = base_prec(precision); });
}

int dmrj_ctx_last_residual(dmrj_ctx* c, double* out) {
  return guard(c, [&] {
    if (!out) throw Error(DMRJ_ERR_ARG, "null pointer");
    *out = std::sqrt(std::max(0.0, read_scalar(*c, S_RES)));
  });
}

int dmrj_selftest_mfma_bf16(dmrj_ctx* c, float* out768) {
  return guard(c, [&] {
    if (!out768) throw Error(DMRJ_ERR_ARG, "null pointer");
    float* d = (float*)c->alloc.alloc(768 * sizeof(float));
    launch_mfma_bf16_probe(*c, d);
    HIP_OK(hipMemcpyAsync(out768, d, 768 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->alloc.release(d);
  });
}

int dmrj_ctx_set_svd_backend(dmrj_ctx* c, int mode) {
  return guard(c, [&] {
    if (mode < 0 || mode > 5) throw Error(DMRJ_ERR_ARG, "svd backend must be 0..5");
    c->svd_backend = mode;
  });
}

int dmrj_ctx_set_eig_batching(dmrj_ctx* c, int mode) {
  return guard(c, [&] {
    if (mode < 0 || mode > 3) throw Error(DMRJ_ERR_ARG, "eigensolve batching mode must be 0..3");
    c->eig_batching = mode;
  });
}

int dmrj_ctx_set_gemm_cfg(dmrj_ctx* c, int cfg) {
  return guard(c, [&] {
    if (cfg < -1 || cfg >= kNumGemmShapes) throw Error(DMRJ_ERR_ARG, "gemm cfg must be -1..4");
    c->gemm_cfg_override = cfg;
  });
}

int dmrj_ctx_svd_stats(dmrj_ctx* c, double* out2, int reset) {
  return guard(c, [&] {
    if (!out2) throw Error(DMRJ_ERR_ARG, "null output");
    out2[0] = c->bj_calls;
    out2[1] = c->bj_sweeps;
    if (reset) c->bj_calls = c->bj_sweeps = 0;
  });
}

int dmrj_ctx_set_tile_order(dmrj_ctx* c, int mode) {
  return guard(c, [&] {
    if (mode < 0 || mode > 1) throw Error(DMRJ_ERR_ARG, "tile order must be 0 or 1");
    c->xcd_order = mode;
  });
}

int dmrj_ctx_set_f16_presplit(dmrj_ctx* c, int on) {
  return guard(c, [&] {
    if (on < 0 || on > 1) throw Error(DMRJ_ERR_ARG, "f16 presplit must be 0 or 1");
    c->f16_presplit = on;
  });
}

int dmrj_ctx_set_autotune(dmrj_ctx* c, int mode) {
  return guard(c, [&] {
    if (mode < 0 || mode > 3) throw Error(DMRJ_ERR_ARG, "autotune mode must be 0, 1, 2 or 3");
    if (mode == 2) {
      c->tune_cache.clear();
      c->tune_hits = c->tune_misses = c->tune_ms = c->tune_skipped = 0;
      return;
    }
    c->autotune = mode;
  });
}

int dmrj_ctx_marker(dmrj_ctx* c, int id) {
  return guard(c, [&] {
    if (id < 0) throw Error(DMRJ_ERR_ARG, "marker id must be >= 0");
    launch_marker(*c, id);
  });
}

int dmrj_ctx_set_graph(dmrj_ctx* c, int mode) {
  return guard(c, [&] {
    if (mode < 0 || mode > 1) throw Error(DMRJ_ERR_ARG, "graph mode must be 0 or 1");
    c->graph_mode = mode;
  });
}

int dmrj_ctx_graph_stats(dmrj_ctx* c, double* out2) {
  return guard(c, [&] {
    if (!out2) throw Error(DMRJ_ERR_ARG, "null output");
    out2[0] = c->graph_instantiations;
    out2[1] = c->graph_updates;
  });
}

int dmrj_ctx_autotune_stats(dmrj_ctx* c, double* out5) {
  return guard(c, [&] {
    if (!out5) throw Error(DMRJ_ERR_ARG, "null output");
    out5[0] = (double)c->tune_cache.size();
    out5[1] = c->tune_hits;
    out5[2] = c->tune_misses;
    out5[3] = c->tune_skipped;
    out5[4] = c->tune_ms;
  });
}

int dmrj_ctx_profile(dmrj_ctx* c, int enable) {
  return guard(c, [&] {
    HIP_OK(hipStreamSynchronize(c->stream));
    c->release_transients();
    c->stage_reset();
    c->prof = enable != 0;
  });
}

int dmrj_ctx_kernel_stats(dmrj_ctx* c, int kind, double* out4, int reset) {
  return guard(c, [&] {
    if (kind < 0 || kind >= 8 || !out4) throw Error(DMRJ_ERR_ARG, "kernel kind 0..7");
    HIP_OK(hipStreamSynchronize(c->stream));
    c->release_transients();
    c->stage_reset();
    out4[0] = c->prof_count[kind];
    out4[1] = c->prof_ms[kind];
    out4[2] = c->prof_flops[kind];
    out4[3] = c->prof_bytes[kind];
    if (reset) c->prof_count[kind] = c->prof_ms[kind] = c->prof_flops[kind] = c->prof_bytes[kind] = 0.0;
  });
}

int dmrj_ctx_flops(dmrj_ctx* c, double* out2, int reset) {
  return guard(c, [&] {
    out2[0] = c->flops_heff;
    out2[1] = c->flops_env;
    if (reset) c->flops_heff = c->flops_env = 0.0;
  });
}

int dmrj_site_from_dense(dmrj_ctx* c, int nq, int64_t mL, int64_t d, int64_t mR, const double* A,
                         const int32_t* qL, const int32_t* qP, const int32_t* qR, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!A || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    *out = site_from_dense(*c, nq, mL, d, mR, A, qL, charges_vec(qP, nq, d), qR, 0, 0, {}, {});
  });
}

int dmrj_site2_from_dense(dmrj_ctx* c, int nq, int64_t mL, int64_t d1, int64_t d2, int64_t mR,
                          const double* A, const int32_t* qL, const int32_t* qP1, const int32_t* qP2,
                          const int32_t* qR, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!A || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    auto q1 = charges_vec(qP1, nq, d1), q2 = charges_vec(qP2, nq, d2);
    std::vector<Charge> qp(d1 * d2);
    for (int64_t j2 = 0; j2 < d2; ++j2)
      for (int64_t j1 = 0; j1 < d1; ++j1) qp[j1 + d1 * j2] = q1[j1] + q2[j2];
    *out = site_from_dense(*c, nq, mL, d1 * d2, mR, A, qL, qp, qR, (int)d1, (int)d2, q1, q2);
  });
}

static void split_planes(const double* z, int64_t n, std::vector<double>& re, std::vector<double>& im) {
  re.resize(n);
  im.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    re[i] = z[2 * i];
    im[i] = z[2 * i + 1];
  }
}

int dmrj_site_from_dense_z(dmrj_ctx* c, int nq, int64_t mL, int64_t d, int64_t mR, const double* Az,
                           const int32_t* qL, const int32_t* qP, const int32_t* qR, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!Az || !out || mL <= 0 || d <= 0 || mR <= 0) throw Error(DMRJ_ERR_ARG, "bad arguments");
    std::vector<double> re, im;
    split_planes(Az, mL * d * mR, re, im);
    *out = site_from_dense(*c, nq, mL, d, mR, re.data(), qL, charges_vec(qP, nq, d), qR, 0, 0, {}, {}, im.data());
  });
}

int dmrj_site2_from_dense_z(dmrj_ctx* c, int nq, int64_t mL, int64_t d1, int64_t d2, int64_t mR,
                            const double* Az, const int32_t* qL, const int32_t* qP1, const int32_t* qP2,
                            const int32_t* qR, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!Az || !out || mL <= 0 || d1 <= 0 || d2 <= 0 || mR <= 0) throw Error(DMRJ_ERR_ARG, "bad arguments");
    auto q1 = charges_vec(qP1, nq, d1), q2 = charges_vec(qP2, nq, d2);
    std::vector<Charge> qp(d1 * d2);
    for (int64_t j2 = 0; j2 < d2; ++j2)
      for (int64_t j1 = 0; j1 < d1; ++j1) qp[j1 + d1 * j2] = q1[j1] + q2[j2];
    std::vector<double> re, im;
    split_planes(Az, mL * d1 * d2 * mR, re, im);
    *out = site_from_dense(*c, nq, mL, d1 * d2, mR, re.data(), qL, qp, qR, (int)d1, (int)d2, q1, q2, im.data());
  });
}

int dmrj_tensor_is_complex(const dmrj_tensor* t) { return t ? (t->cplx ? 1 : 0) : DMRJ_ERR_ARG; }

int dmrj_tensor_to_complex(dmrj_ctx* c, dmrj_tensor* t) {
  return guard(c, [&] {
    if (!t) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (t->cplx) return;
    const int64_t n = std::max<int64_t>(t->nelem(), 1);
    double* d = (double*)c->alloc.alloc(2 * n * sizeof(double));
    HIP_OK(hipMemcpyAsync(d, t->data, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    launch_fill(*c, d + n, n, 0.0);
    c->alloc.release(t->data);
    t->data = d;
    t->cplx = 1;
  });
}

int dmrj_tensor_info(const dmrj_tensor* t, int64_t* d) {
  if (!t || !d) return DMRJ_ERR_ARG;
  for (int i = 0; i < 9; ++i) d[i] = 0;
  if (t->kind == 0) {
    const SiteMeta& m = t->site;
    d[0] = m.d1 ? 1 : 0;
    d[1] = m.left.total;
    if (m.d1) {
      d[2] = m.d1;
      d[3] = m.d2;
      d[4] = m.right.total;
    } else {
      d[2] = m.P;
      d[3] = m.right.total;
      d[4] = 1;
    }
    d[5] = m.size;
    d[6] = (int64_t)m.blocks.size();
    d[7] = m.left.nsec();
    d[8] = m.right.nsec();
  } else {
    const EnvMeta& m = t->env;
    d[0] = t->kind == 1 ? 2 : 3;
    d[1] = m.bra.total;
    d[2] = m.D;
    d[3] = m.ket.total;
    d[4] = 1;
    d[5] = m.size;
    d[6] = (int64_t)m.blocks.size();
    d[7] = m.bra.nsec();
    d[8] = m.ket.nsec();
  }
  return DMRJ_OK;
}

// Dense export; S = 1: real output, S = 2: interleaved complex output.
static void to_dense_impl(dmrj_ctx* c, const dmrj_tensor* t, double* out, int32_t* qfirst, int32_t* qlast, int S) {
  const int64_t n = t->nelem();
  const int planes = t->cplx ? 2 : 1;
  std::vector<double> h(std::max<int64_t>(n * planes, 1));
  if (n > 0) HIP_OK(hipMemcpyAsync(h.data(), t->data, n * planes * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->release_transients();
  c->stage_reset();
  const int np = std::min(planes, S);
  if (t->kind == 0) {
    const SiteMeta& m = t->site;
    const int64_t mL = m.left.total, mR = m.right.total, P = m.P;
    std::fill(out, out + mL * P * mR * S, 0.0);
    for (int p = 0; p < np; ++p)
      for (const auto& b : m.blocks) {
        const int na = m.left.dim[b.a], nb = m.right.dim[b.b];
        const int64_t oa = m.left.off[b.a], ob = m.right.off[b.b];
        for (int cb = 0; cb < nb; ++cb)
          for (int ra = 0; ra < na; ++ra)
            out[((oa + ra) + mL * (b.s + P * (ob + cb))) * S + p] = h[p * n + b.off + ra + (int64_t)cb * na];
      }
    space_to_list(m.left, t->nq, qfirst);
    space_to_list(m.right, t->nq, qlast);
  } else {
    const EnvMeta& m = t->env;
    const int64_t nb = m.bra.total, nk = m.ket.total, D = m.D;
    std::fill(out, out + nb * D * nk * S, 0.0);
    for (int p = 0; p < np; ++p)
      for (const auto& b : m.blocks) {
        const int nap = m.bra.dim[b.ap], na = m.ket.dim[b.a];
        const int64_t oap = m.bra.off[b.ap], oa = m.ket.off[b.a];
        for (int cc = 0; cc < na; ++cc)
          for (int r = 0; r < nap; ++r) {
            const double v = h[p * n + b.off + r + (int64_t)cc * nap];
            if (t->kind == 1)
              out[((oap + r) + nb * (b.w + D * (oa + cc))) * S + p] = v;   // Lenv[bra, w, ket]
            else
              out[((oa + cc) + nk * (b.w + D * (oap + r))) * S + p] = v;   // Renv[ket, u, bra]
          }
      }
    if (t->kind == 1) {
      space_to_list(m.bra, t->nq, qfirst);
      space_to_list(m.ket, t->nq, qlast);
    } else {
      space_to_list(m.ket, t->nq, qfirst);
      space_to_list(m.bra, t->nq, qlast);
    }
  }
}

int dmrj_tensor_to_dense(dmrj_ctx* c, const dmrj_tensor* t, double* out, int32_t* qfirst, int32_t* qlast) {
  return guard(c, [&] {
    if (!t || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (t->cplx) throw Error(DMRJ_ERR_ARG, "tensor is complex: export it with dmrj_tensor_to_dense_z");
    to_dense_impl(c, t, out, qfirst, qlast, 1);
  });
}

int dmrj_tensor_to_dense_z(dmrj_ctx* c, const dmrj_tensor* t, double* out, int32_t* qfirst, int32_t* qlast) {
  return guard(c, [&] {
    if (!t || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    to_dense_impl(c, t, out, qfirst, qlast, 2);
  });
}

int dmrj_tensor_free(dmrj_ctx* c, dmrj_tensor* t) {
  return guard(c, [&] {
    if (!t) return;
    c->alloc.release(t->data);
    delete t;
  });
}

int dmrj_tensor_copy(dmrj_ctx* c, const dmrj_tensor* t, dmrj_tensor** out) {
  return guard(c, [&] {
    auto* n = new dmrj_tensor();
    n->kind = t->kind;
    n->nq = t->nq;
    n->site = t->site;
    n->env = t->env;
    n->cplx = t->cplx;
    alloc_data(*c, *n);
    HIP_OK(hipMemcpyAsync(n->data, t->data, std::max<int64_t>(t->nstore(), 1) * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    *out = n;
  });
}

int dmrj_site_export(dmrj_ctx* c, const dmrj_tensor* t, int64_t* info8, int32_t* qL, int32_t* qP, int32_t* qR,
                     double* data) {
  return guard(c, [&] {
    if (!t || !info8) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (t->kind != 0 || t->site.d1) throw Error(DMRJ_ERR_ARG, "site_export: single-site MPS tensors only");
    const SiteMeta& m = t->site;
    const int nq = t->nq;
    info8[0] = nq;
    info8[1] = m.left.total;
    info8[2] = m.P;
    info8[3] = m.right.total;
    info8[4] = t->nstore();
    info8[5] = t->cplx;
    info8[6] = info8[7] = 0;
    space_to_list(m.left, nq, qL);
    space_to_list(m.right, nq, qR);
    if (qP)
      for (int s = 0; s < m.P; ++s)
        for (int k = 0; k < nq; ++k) qP[s * nq + k] = m.qp[s].v[k];
    if (data && t->nstore() > 0) {
      HIP_OK(hipMemcpyAsync(data, t->data, t->nstore() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_OK(hipStreamSynchronize(c->stream));
    }
  });
}

int dmrj_site_import(dmrj_ctx* c, int nq, int64_t mL, int64_t P, int64_t mR, const int32_t* qL,
                     const int32_t* qP, const int32_t* qR, const double* data, int64_t n, int cplx,
                     dmrj_tensor** out) {
  return guard(c, [&] {
    if (!out || !data) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (nq < 0 || nq > kMaxQ) throw Error(DMRJ_ERR_ARG, "nq must be 0, 1 or 2");
    if (mL <= 0 || P <= 0 || mR <= 0 || (nq && (!qL || !qP || !qR))) throw Error(DMRJ_ERR_ARG, "bad dimensions");
    auto* t = new_tensor(*c, 0);
    try {
      t->nq = nq;
      SiteMeta& m = t->site;
      m.left = space_from_list(qL, nq, mL);
      m.right = space_from_list(qR, nq, mR);
      m.P = (int)P;
      m.qp = nq ? charges_vec(qP, nq, P) : std::vector<Charge>(P);
      m.build();
      t->cplx = cplx ? 1 : 0;
      if (n != t->nstore()) throw Error(DMRJ_ERR_ARG, "site_import: stored size does not match the charges");
      check_finite_host(data, n);
      alloc_data(*c, *t);
      if (n > 0) {
        HIP_OK(hipMemcpyAsync(t->data, data, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
      }
    } catch (...) {
      if (t->data) c->alloc.release(t->data);
      delete t;
      throw;
    }
    *out = t;
  });
}

int dmrj_mpo_create(dmrj_ctx* c, int nq, int64_t DL, int64_t d, int64_t DR, const double* W,
                    const int32_t* qP, const int32_t* deltaL, const int32_t* deltaR, dmrj_mpo** out) {
  return guard(c, [&] {
    if (!W || !out || DL <= 0 || d <= 0 || DR <= 0) throw Error(DMRJ_ERR_ARG, "bad MPO arguments");
    if (nq < 0 || nq > kMaxQ) throw Error(DMRJ_ERR_ARG, "nq must be 0, 1 or 2");
    auto* m = new dmrj_mpo();
    m->DL = (int)DL;
    m->d = (int)d;
    m->DR = (int)DR;
    m->W.assign(W, W + DL * d * d * DR);
    m->qp = charges_vec(qP, nq, d);
    m->dL = charges_vec(deltaL, nq, DL);
    m->dR = charges_vec(deltaR, nq, DR);
    check_finite_host(W, DL * d * d * DR);
    for (int u = 0; u < DR; ++u)
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < d; ++j)
          for (int w = 0; w < DL; ++w)
            if (m->at(w, j, k, u) != 0.0 && !(m->dR[u] == m->dL[w] + m->qp[k] - m->qp[j])) {
              delete m;
              throw Error(DMRJ_ERR_ARG, "MPO entry violates delta(u) = delta(w) + q(k) - q(j)");
            }
    *out = m;
  });
}

int dmrj_mpo_free(dmrj_ctx* c, dmrj_mpo* W) {
  return guard(c, [&] { delete W; });
}

int dmrj_env_boundary(dmrj_ctx* c, int left, int nq, const int32_t* q, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!out) throw Error(DMRJ_ERR_ARG, "null pointer");
    auto* t = new_tensor(*c, left ? 1 : 2);
    t->nq = nq;
    Space s;
    s.q.push_back(nq ? charge_of(q, nq, 0) : Charge());
    s.dim.push_back(1);
    s.finish();
    t->env.bra = s;
    t->env.ket = s;
    t->env.D = 1;
    t->env.delta = {Charge()};
    t->env.build();
    alloc_data(*c, *t);
    launch_fill(*c, t->data, 1, 1.0);
    *out = t;
  });
}

int dmrj_env_left(dmrj_ctx* c, const dmrj_tensor* Lin, const dmrj_tensor* A, const dmrj_mpo* W,
                  dmrj_tensor** Lout) {
  return guard(c, [&] {
    if (!Lin || !A || !W || !Lout || Lin->kind != 1 || A->kind != 0 || A->site.d1)
      throw Error(DMRJ_ERR_ARG, "env_left: bad handles");
    auto* t = new_tensor(*c, 1);
    t->nq = A->nq;
    EnvMeta lo;
    // metadata first (to size the output), then the plan fills the tasks
    lo.bra = A->site.right;
    lo.ket = A->site.right;
    lo.D = W->DR;
    lo.delta = W->dR;
    lo.build();
    t->env = lo;
    const int zm = (Lin->cplx ? 1 : 0) | (A->cplx ? 2 : 0);
    t->cplx = zm ? 1 : 0;
    alloc_data(*c, *t);
    EnvPlan p = plan_left(*c, Lin->env, Lin->data, A->site, A->data, *W, t->env, t->data, zm);
    run_env(*c, p);
    *Lout = t;
  });
}

int dmrj_env_right(dmrj_ctx* c, const dmrj_tensor* Rin, const dmrj_tensor* A, const dmrj_mpo* W,
                   dmrj_tensor** Rout) {
  return guard(c, [&] {
    if (!Rin || !A || !W || !Rout || Rin->kind != 2 || A->kind != 0 || A->site.d1)
      throw Error(DMRJ_ERR_ARG, "env_right: bad handles");
    auto* t = new_tensor(*c, 2);
    t->nq = A->nq;
    EnvMeta ro;
    ro.bra = A->site.left;
    ro.ket = A->site.left;
    ro.D = W->DL;
    ro.delta = W->dL;
    ro.build();
    t->env = ro;
    const int zm = (Rin->cplx ? 1 : 0) | (A->cplx ? 2 : 0);
    t->cplx = zm ? 1 : 0;
    alloc_data(*c, *t);
    EnvPlan p = plan_right(*c, Rin->env, Rin->data, A->site, A->data, *W, t->env, t->data, zm);
    run_env(*c, p);
    *Rout = t;
  });
}

int dmrj_env_left2(dmrj_ctx* c, const dmrj_tensor* Lin, const dmrj_tensor* B, const dmrj_tensor* A,
                   const dmrj_mpo* W, dmrj_tensor** Lout) {
  return guard(c, [&] {
    if (!Lin || !A || !B || !W || !Lout || Lin->kind != 1 || A->kind != 0 || A->site.d1 || B->kind != 0 ||
        B->site.d1)
      throw Error(DMRJ_ERR_ARG, "env_left2: bad handles");
    if (A->nq != B->nq) throw Error(DMRJ_ERR_ARG, "env_left2: bra and ket carry different charge types");
    auto* t = new_tensor(*c, 1);
    t->nq = A->nq;
    EnvMeta lo;
    lo.bra = B->site.right;
    lo.ket = A->site.right;
    lo.D = W->DR;
    lo.delta = W->dR;
    lo.build();
    t->env = lo;
    const int zm = (Lin->cplx ? 1 : 0) | (A->cplx ? 2 : 0) | (B->cplx ? 4 : 0);
    t->cplx = zm ? 1 : 0;
    alloc_data(*c, *t);
    EnvPlan p = plan_left(*c, Lin->env, Lin->data, A->site, A->data, *W, t->env, t->data, zm, &B->site, B->data);
    run_env(*c, p);
    *Lout = t;
  });
}

int dmrj_env_right2(dmrj_ctx* c, const dmrj_tensor* Rin, const dmrj_tensor* B, const dmrj_tensor* A,
                    const dmrj_mpo* W, dmrj_tensor** Rout) {
  return guard(c, [&] {
    if (!Rin || !A || !B || !W || !Rout || Rin->kind != 2 || A->kind != 0 || A->site.d1 || B->kind != 0 ||
        B->site.d1)
      throw Error(DMRJ_ERR_ARG, "env_right2: bad handles");
    if (A->nq != B->nq) throw Error(DMRJ_ERR_ARG, "env_right2: bra and ket carry different charge types");
    auto* t = new_tensor(*c, 2);
    t->nq = A->nq;
    EnvMeta ro;
    ro.bra = B->site.left;
    ro.ket = A->site.left;
    ro.D = W->DL;
    ro.delta = W->dL;
    ro.build();
    t->env = ro;
    const int zm = (Rin->cplx ? 1 : 0) | (A->cplx ? 2 : 0) | (B->cplx ? 4 : 0);
    t->cplx = zm ? 1 : 0;
    alloc_data(*c, *t);
    EnvPlan p = plan_right(*c, Rin->env, Rin->data, A->site, A->data, *W, t->env, t->data, zm, &B->site, B->data);
    run_env(*c, p);
    *Rout = t;
  });
}

int dmrj_theta2(dmrj_ctx* c, const dmrj_tensor* psiL, const dmrj_tensor* psiR, dmrj_tensor** theta) {
  return guard(c, [&] {
    if (!psiL || !psiR || !theta) throw Error(DMRJ_ERR_ARG, "null pointer");
    *theta = make_theta(*c, *psiL, *psiR);
  });
}

int dmrj_heff2_apply(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_mpo* W1, const dmrj_mpo* W2,
                     const dmrj_tensor* R, const dmrj_tensor* x, int precision, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!L || !W1 || !W2 || !R || !x || !out || L->kind != 1 || R->kind != 2 || x->kind != 0 || !x->site.d1)
      throw Error(DMRJ_ERR_ARG, "heff2_apply: bad handles");
    const int prec = base_prec(precision);
    auto* y = new_tensor(*c, 0);
    y->nq = x->nq;
    y->site = x->site;
    const int zm = (L->cplx ? 1 : 0) | (x->cplx ? 2 : 0) | (R->cplx ? 4 : 0);
    y->cplx = zm ? 1 : 0;
    alloc_data(*c, *y);
    FusedMpo f = fuse_two(*W1, *W2);
    double *T1, *T3;
    HeffPlan p = plan_heff(*c, x->site, x->data, y->data, L->env, L->data, R->env, R->data, f, &T1, &T3, zm);
    run_heff(*c, p, prec);
    c->alloc.release(T1);
    c->alloc.release(T3);
    *out = y;
  });
}

int dmrj_heff1_apply(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_mpo* W, const dmrj_tensor* R,
                     const dmrj_tensor* x, int precision, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!L || !W || !R || !x || !out || L->kind != 1 || R->kind != 2 || x->kind != 0 || x->site.d1)
      throw Error(DMRJ_ERR_ARG, "heff1_apply: bad handles");
    const int prec = base_prec(precision);
    auto* y = new_tensor(*c, 0);
    y->nq = x->nq;
    y->site = x->site;
    const int zm = (L->cplx ? 1 : 0) | (x->cplx ? 2 : 0) | (R->cplx ? 4 : 0);
    y->cplx = zm ? 1 : 0;
    alloc_data(*c, *y);
    FusedMpo f = fuse_one(*W);
    double *T1, *T3;
    HeffPlan p = plan_heff(*c, x->site, x->data, y->data, L->env, L->data, R->env, R->data, f, &T1, &T3, zm);
    run_heff(*c, p, prec);
    c->alloc.release(T1);
    c->alloc.release(T3);
    *out = y;
  });
}

int dmrj_lanczos2(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* psiL,
                  const dmrj_tensor* psiR, const dmrj_mpo* W1, const dmrj_mpo* W2, int precision,
                  dmrj_tensor** AA, double* energy) {
  return guard(c, [&] {
    if (!L || !R || !psiL || !psiR || !W1 || !W2 || !AA) throw Error(DMRJ_ERR_ARG, "null pointer");
    *AA = lanczos2(*c, *L, *R, *psiL, *psiR, *W1, *W2, precision);
    const double e = read_scalar(*c, S_E);
    if (energy) *energy = e;
  });
}

int dmrj_krylov_expm2(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* AA,
                      const dmrj_mpo* W1, const dmrj_mpo* W2, int krylov, double prefactor_re,
                      double prefactor_im, int precision, dmrj_tensor** out, double* energy, int* kdim) {
  return guard(c, [&] {
    if (!L || !R || !AA || !W1 || !W2 || !out || L->kind != 1 || R->kind != 2 || AA->kind != 0 || !AA->site.d1)
      throw Error(DMRJ_ERR_ARG, "krylov_expm2: bad handles");
    *out = krylov_expm(*c, *L, *R, *AA, fuse_two(*W1, *W2), krylov, prefactor_re, prefactor_im, base_prec(precision),
                       energy, kdim, 1);
  });
}

int dmrj_krylov_expm1(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* A,
                      const dmrj_mpo* W, int krylov, double prefactor_re, double prefactor_im, int precision,
                      dmrj_tensor** out, double* energy, int* kdim) {
  return guard(c, [&] {
    if (!L || !R || !A || !W || !out || L->kind != 1 || R->kind != 2 || A->kind != 0 || A->site.d1)
      throw Error(DMRJ_ERR_ARG, "krylov_expm1: bad handles");
    *out = krylov_expm(*c, *L, *R, *A, fuse_one(*W), krylov, prefactor_re, prefactor_im, base_prec(precision), energy,
                       kdim, 2);
  });
}

int dmrj_lanczos1(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* A, const dmrj_mpo* W,
                  int krylov, int precision, dmrj_tensor** out, double* energy, int* kdim) {
  return guard(c, [&] {
    if (!L || !R || !A || !W || !out || L->kind != 1 || R->kind != 2 || A->kind != 0 || A->site.d1)
      throw Error(DMRJ_ERR_ARG, "lanczos1: bad handles");
    *out = krylov_expm(*c, *L, *R, *A, fuse_one(*W), krylov, 0.0, 0.0, base_prec(precision), energy, kdim, 2, true);
  });
}

int dmrj_lanczos_k(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, const dmrj_tensor* x, const dmrj_mpo* W1,
                   const dmrj_mpo* W2, int krylov, int precision, dmrj_tensor** out, double* energy, int* kdim) {
  return guard(c, [&] {
    if (!L || !R || !x || !W1 || !out || L->kind != 1 || R->kind != 2 || x->kind != 0)
      throw Error(DMRJ_ERR_ARG, "lanczos_k: bad handles");
    const bool two = x->site.d1 != 0;
    if (two && !W2) throw Error(DMRJ_ERR_ARG, "lanczos_k: a two-site tensor needs the second MPO tensor");
    *out = krylov_expm(*c, *L, *R, *x, two ? fuse_two(*W1, *W2) : fuse_one(*W1), krylov, 0.0, 0.0,
                       base_prec(precision), energy, kdim, two ? 1 : 2, true);
  });
}

int dmrj_expand3s(dmrj_ctx* c, const dmrj_tensor* E, const dmrj_tensor* A, const dmrj_mpo* W, double alpha, int dir,
                  dmrj_tensor** out) {
  return guard(c, [&] {
    if (!E || !A || !W || !out || (dir != 1 && dir != -1)) throw Error(DMRJ_ERR_ARG, "expand3s: bad arguments");
    if (!std::isfinite(alpha)) throw Error(DMRJ_ERR_ARG, "expand3s: alpha must be finite");
    *out = expand3s(*c, *E, *A, *W, alpha, dir);
  });
}

int dmrj_absorb3s(dmrj_ctx* c, const dmrj_tensor* F, const dmrj_tensor* N, int dir, dmrj_tensor** out) {
  return guard(c, [&] {
    if (!F || !N || !out || (dir != 1 && dir != -1)) throw Error(DMRJ_ERR_ARG, "absorb3s: bad arguments");
    *out = absorb3s(*c, *F, *N, dir);
  });
}

int dmrj_svd_trunc2(dmrj_ctx* c, const dmrj_tensor* AA, int64_t m, int64_t minm, double cutoff, int dir,
                    dmrj_tensor** psiL, dmrj_tensor** psiR, double* truncerr, int64_t* kept,
                    double* svals, int64_t cap) {
  return guard(c, [&] {
    if (!AA || !psiL || !psiR) throw Error(DMRJ_ERR_ARG, "null pointer");
    SvdResult r = svd_trunc(*c, *AA, m, minm, cutoff, dir);
    *psiL = r.L;
    *psiR = r.R;
    if (truncerr) *truncerr = r.truncerr;
    if (kept) *kept = (int64_t)r.svals.size();
    if (svals)
      for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)r.svals.size()); ++i) svals[i] = r.svals[i];
  });
}

int dmrj_normalize(dmrj_ctx* c, dmrj_tensor* x) {
  return guard(c, [&] {
    if (!x) throw Error(DMRJ_ERR_ARG, "null pointer");
    const double* xs[1] = {x->data};
    launch_dots(*c, 1, xs, xs, x->nstore(), c->scal + S_OC);
    launch_lz_scalar(*c, 0);
    launch_scale_dev(*c, x->data, x->nstore(), c->scal + S_INVOC);
  });
}

int dmrj_twostep(dmrj_ctx* c, const dmrj_tensor* L, const dmrj_tensor* R, dmrj_tensor* psiL, dmrj_tensor* psiR,
                 const dmrj_mpo* W1, const dmrj_mpo* W2, int dir, int64_t m, int64_t minm, double cutoff,
                 int precision, dmrj_tensor** newL, dmrj_tensor** newR, double* energy, double* truncerr,
                 int64_t* kept, double* svals, int64_t cap) {
  return guard(c, [&] {
    if (!L || !R || !psiL || !psiR || !W1 || !W2 || !newL || !newR) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (L->cplx || R->cplx || psiL->cplx || psiR->cplx)
      throw Error(DMRJ_ERR_ARG, "twostep: the DMRG step takes real tensors");
    dmrj_tensor* oc = dir > 0 ? psiL : psiR;
    {
      const double* xs[1] = {oc->data};
      launch_dots(*c, 1, xs, xs, oc->nelem(), c->scal + S_OC);
      launch_lz_scalar(*c, 0);
      launch_scale_dev(*c, oc->data, oc->nelem(), c->scal + S_INVOC);
    }
    dmrj_tensor* AA = lanczos2(*c, *L, *R, *psiL, *psiR, *W1, *W2, precision);
    HIP_OK(hipMemcpyAsync(c->hscal, c->scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SvdResult r;
    try {
      r = svd_trunc(*c, *AA, m, minm, cutoff, dir);  // synchronises the stream
    } catch (...) {
      c->alloc.release(AA->data);
      delete AA;
      throw;
    }
    if (energy) *energy = c->hscal[S_E];
    c->alloc.release(AA->data);
    delete AA;
    *newL = r.L;
    *newR = r.R;
    if (truncerr) *truncerr = r.truncerr;
    if (kept) *kept = (int64_t)r.svals.size();
    if (svals)
      for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)r.svals.size()); ++i) svals[i] = r.svals[i];
  });
}

int dmrj_dot(dmrj_ctx* c, const dmrj_tensor* x, const dmrj_tensor* y, double* out) {
  return guard(c, [&] {
    if (!x || !y || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (x->nstore() != y->nstore()) throw Error(DMRJ_ERR_ARG, "dot: size mismatch");
    const double* xs[1] = {x->data};
    const double* ys[1] = {y->data};
    launch_dots(*c, 1, xs, ys, x->nstore(), c->scal + S_TT);  // Re <x|y> for complex data
    *out = read_scalar(*c, S_TT);
  });
}

int dmrj_dots_noalloc(dmrj_ctx* c, int npairs, const dmrj_tensor* const* xs, const dmrj_tensor* const* ys,
                      double* out_dev) {
  return guard(c, [&] {
    if (!xs || !ys || !out_dev) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (npairs < 1 || npairs > kNoAllocPairs)
      throw Error(DMRJ_ERR_ARG, "dots_noalloc: npairs must be 1.." + std::to_string(kNoAllocPairs) +
                                    " (the scratch holds the partial sums of at most that many inner products)");
    for (int p = 0; p < npairs; ++p) {
      if (!xs[p] || !ys[p]) throw Error(DMRJ_ERR_ARG, "null tensor");
      if (xs[p]->nstore() != ys[p]->nstore()) throw Error(DMRJ_ERR_ARG, "dots_noalloc: size mismatch");
    }
    if (!c->noalloc) throw Error(DMRJ_ERR_STATE, "no scratch address for this device");
    // one pair per launch pair, as dmrj_dot: identical partial sums and
    // summation order, so each result equals dmrj_dot's bit for bit
    for (int p = 0; p < npairs; ++p) {
      const double* xp[1] = {xs[p]->data};
      const double* yp[1] = {ys[p]->data};
      launch_dots_buf(*c, 1, xp, yp, xs[p]->nstore(), out_dev + p, c->noalloc + (int64_t)p * kDotPartials);
    }
  });
}

int dmrj_site_apply_op(dmrj_ctx* c, const dmrj_tensor* A, const double* op, const int32_t* dqL, const int32_t* dqR,
                       dmrj_tensor** out) {
  return guard(c, [&] {
    if (!A || !op || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (A->kind != 0 || A->site.d1) throw Error(DMRJ_ERR_ARG, "site_apply_op: single-site MPS tensors only");
    const SiteMeta& x = A->site;
    const int nq = A->nq;
    const int P = x.P;
    Charge sl, sr;
    if (nq) {
      if (!dqL || !dqR) throw Error(DMRJ_ERR_ARG, "site_apply_op: charge shifts required for a charged tensor");
      for (int k = 0; k < nq; ++k) {
        sl.v[k] = dqL[k];
        sr.v[k] = dqR[k];
      }
    }
    // every non-zero O[k, j] must change the charge by q(k) - q(j) = sr - sl
    for (int j = 0; j < P; ++j)
      for (int k = 0; k < P; ++k) {
        const double v = op[k + (int64_t)P * j];
        if (!std::isfinite(v)) throw Error(DMRJ_ERR_NUMERIC, "site_apply_op: operator holds NaN/Inf");
        if (v != 0.0 && x.qp[k] - x.qp[j] != sr - sl)
          throw Error(DMRJ_ERR_ARG, "site_apply_op: operator entry changes the charge by other than dqR - dqL");
      }
    auto shifted = [](const Space& s, const Charge& d) {
      Space r = s;
      for (auto& q : r.q) q = q + d;
      r.finish();
      return r;
    };
    auto* n = new dmrj_tensor();
    n->kind = 0;
    n->nq = nq;
    n->cplx = A->cplx;
    n->site.P = P;
    n->site.qp = x.qp;
    n->site.left = shifted(x.left, sl);
    n->site.right = shifted(x.right, sr);
    n->site.build();
    try {
      alloc_data(*c, *n);
      // block (a, k) of the result = sum_j O[k, j] * block (a, j) of A: with
      // the shifts both blocks sit on the same right sector index
      std::vector<MixDst> dsts;
      std::vector<MixTerm> terms;
      const int planes = A->cplx ? 2 : 1;
      for (int pl = 0; pl < planes; ++pl)
        for (const SiteBlock& ob : n->site.blocks) {
          const int64_t ne = (int64_t)n->site.left.dim[ob.a] * n->site.right.dim[ob.b];
          MixDst d{n->data + pl * n->site.size + ob.off, ne, (int)terms.size(), 0};
          for (int j = 0; j < P; ++j) {
            const int xb = x.block(ob.a, j);
            if (xb < 0 || x.blocks[xb].b != ob.b) continue;
            const double cf = op[ob.s + (int64_t)P * j];
            if (cf == 0.0) continue;
            terms.push_back({A->data + pl * x.size + x.blocks[xb].off, cf});
            d.nt++;
          }
          dsts.push_back(d);
        }
      launch_mix(*c, dsts, terms);
    } catch (...) {
      if (n->data) c->alloc.release(n->data);
      delete n;
      throw;
    }
    *out = n;
  });
}

int dmrj_dotu(dmrj_ctx* c, const dmrj_tensor* x, const dmrj_tensor* y, double* out2) {
  return guard(c, [&] {
    if (!x || !y || !out2) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (x->nelem() != y->nelem()) throw Error(DMRJ_ERR_ARG, "dotu: size mismatch");
    const int64_t n = x->nelem();
    const double* xr = x->data;
    const double* yr = y->data;
    // sum x y without conjugation; a real operand has no imaginary plane
    double re = 0.0, im = 0.0;
    {
      const double* xs[1] = {xr};
      const double* ys[1] = {yr};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_TT);
      re = read_scalar(*c, S_TT);
    }
    if (x->cplx && y->cplx) {
      const double* xs[3] = {xr + n, xr, xr + n};
      const double* ys[3] = {yr + n, yr + n, yr};
      launch_dots(*c, 3, xs, ys, n, c->scal + S_A1);  // S_A1, S_CAXPY, S_B1
      re -= read_scalar(*c, S_A1);
      im = read_scalar(*c, S_CAXPY) + read_scalar(*c, S_B1);
    } else if (y->cplx) {
      const double* xs[1] = {xr};
      const double* ys[1] = {yr + n};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_A1);
      im = read_scalar(*c, S_A1);
    } else if (x->cplx) {
      const double* xs[1] = {xr + n};
      const double* ys[1] = {yr};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_A1);
      im = read_scalar(*c, S_A1);
    }
    out2[0] = re;
    out2[1] = im;
  });
}

int dmrj_dotc(dmrj_ctx* c, const dmrj_tensor* x, const dmrj_tensor* y, double* out2) {
  return guard(c, [&] {
    if (!x || !y || !out2) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (x->nelem() != y->nelem()) throw Error(DMRJ_ERR_ARG, "dotc: size mismatch");
    const int64_t n = x->nelem();
    const double* xr = x->data;
    const double* yr = y->data;
    // <x|y> = sum conj(x) y; a real operand has no imaginary plane
    double re = 0.0, im = 0.0;
    {
      const double* xs[1] = {xr};
      const double* ys[1] = {yr};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_TT);
      re = read_scalar(*c, S_TT);
    }
    if (x->cplx && y->cplx) {
      const double* xs[3] = {xr + n, xr, xr + n};
      const double* ys[3] = {yr + n, yr + n, yr};
      launch_dots(*c, 3, xs, ys, n, c->scal + S_A1);  // S_A1, S_CAXPY, S_B1
      re += read_scalar(*c, S_A1);
      im = read_scalar(*c, S_CAXPY) - read_scalar(*c, S_B1);
    } else if (y->cplx) {
      const double* xs[1] = {xr};
      const double* ys[1] = {yr + n};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_A1);
      im = read_scalar(*c, S_A1);
    } else if (x->cplx) {
      const double* xs[1] = {xr + n};
      const double* ys[1] = {yr};
      launch_dots(*c, 1, xs, ys, n, c->scal + S_A1);
      im = -read_scalar(*c, S_A1);
    }
    out2[0] = re;
    out2[1] = im;
  });
}

int dmrj_norm(dmrj_ctx* c, const dmrj_tensor* x, double* out) {
  return guard(c, [&] {
    if (!x || !out) throw Error(DMRJ_ERR_ARG, "null pointer");
    const double* xs[1] = {x->data};
    launch_dots(*c, 1, xs, xs, x->nstore(), c->scal + S_TT);
    *out = std::sqrt(read_scalar(*c, S_TT));
  });
}

int dmrj_scale(dmrj_ctx* c, dmrj_tensor* x, double a) {
  return guard(c, [&] {
    if (!x) throw Error(DMRJ_ERR_ARG, "null pointer");
    double* tmp = (double*)c->upload(&a, sizeof(double));
    launch_scale_dev(*c, x->data, x->nstore(), tmp);
  });
}

int dmrj_axpy(dmrj_ctx* c, double a, const dmrj_tensor* x, dmrj_tensor* y) {
  return guard(c, [&] {
    if (!x || !y) throw Error(DMRJ_ERR_ARG, "null pointer");
    if (x->nstore() != y->nstore()) throw Error(DMRJ_ERR_ARG, "axpy: size mismatch");
    double* tmp = (double*)c->upload(&a, sizeof(double));
    launch_axpy_dev(*c, y->data, x->data, x->nstore(), tmp);
  });
}

int dmrj_selftest_gemm(dmrj_ctx* c, int cfg, int64_t M, int64_t N, int64_t K, int tA, int tB,
                       const double* A, const double* B, double* Cout) {
  return guard(c, [&] {
    const int prec = (cfg >> 4) & 15;
    cfg &= 15;
    if (!A || !B || !Cout || M <= 0 || N <= 0 || K < 0 || cfg < 0 || cfg >= kNumGemmShapes)
      throw Error(DMRJ_ERR_ARG, "selftest_gemm: bad arguments");
    const int64_t na = std::max<int64_t>(M * K, 1), nb = std::max<int64_t>(K * N, 1);
    double* dA = (double*)c->alloc.alloc(na * sizeof(double));
    double* dB = (double*)c->alloc.alloc(nb * sizeof(double));
    double* dC = (double*)c->alloc.alloc(M * N * sizeof(double));
    HIP_OK(hipMemcpyAsync(dA, A, M * K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(dB, B, K * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<GemmTask> t{{dC, M, (int)M, (int)N, 0, 1, 0.0}};
    std::vector<GemmSeg> s{{dA, dB, tA ? K : M, tB ? N : K, (int)K, (tA ? 1 : 0) | (tB ? 2 : 0), 1.0}};
    if (prec > DMRJ_PREC_F16X3) throw Error(DMRJ_ERR_ARG, "selftest_gemm: bad precision");
    launch_gemm_arith(*c, t, s, cfg, prec);
    HIP_OK(hipMemcpyAsync(Cout, dC, M * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->release_transients();
    c->stage_reset();
    c->alloc.release(dA);
    c->alloc.release(dB);
    c->alloc.release(dC);
  });
}

int dmrj_bj_panel_width(int rows, int64_t lds_max, int64_t static_lds) {
  if (rows < 1 || lds_max < 0 || static_lds < 0) return DMRJ_ERR_ARG;
  return bj_panel_width(rows, (size_t)lds_max, (size_t)static_lds);
}

int dmrj_selftest_copy(dmrj_ctx* c, int64_t n) {
  return guard(c, [&] {
    if (n <= 0 || n % 4096 != 0 || n / 4096 > (1 << 30)) throw Error(DMRJ_ERR_ARG, "n must be a positive multiple of 4096");
    double* a = (double*)c->alloc.alloc(n * sizeof(double));
    double* b = (double*)c->alloc.alloc(n * sizeof(double));
    launch_fill(*c, a, n, 1.0);
    std::vector<CopyOp> ops{{a, b, 4096, 4096, 4096, (int)(n / 4096), nullptr, nullptr, 1.0}};
    launch_copy(*c, ops);
    HIP_OK(hipStreamSynchronize(c->stream));
    c->release_transients();
    c->stage_reset();
    c->alloc.release(a);
    c->alloc.release(b);
  });
}

int dmrj_selftest_mfma_f64(dmrj_ctx* c, double* out768) {
  return guard(c, [&] {
    if (!out768) throw Error(DMRJ_ERR_ARG, "null pointer");
    double* d = (double*)c->alloc.alloc(768 * sizeof(double));
    launch_mfma_probe(*c, d);
    HIP_OK(hipMemcpyAsync(out768, d, 768 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->alloc.release(d);
  });
}

}  // extern "C"
