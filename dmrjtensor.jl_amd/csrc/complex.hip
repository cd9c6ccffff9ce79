// ComplexF64 support on split planes.
//
// Complex tensors keep their block layout and store the real plane followed
// by the imaginary plane, so every kernel written for FP64 data serves both.
// A contraction plan is first built exactly as for real data; complexify_*
// then expands each product (A, B) into real products on the planes:
//     Re C = sum  Ar Br - sA sB Ai Bi,   Im C = sum  sB Ar Bi + sA Ai Br,
// with sX = -1 where X enters complex-conjugated.  The expanded segments of
// one output block stay in one task, so the whole complex contraction is
// still one grouped-GEMM launch.  Operands that are real (environments built
// from a real state, the MPO coefficients) have no imaginary plane and their
// imaginary products are dropped.
//
// Kernels here: the placement / extraction around the batched Hermitian
// eigensolve of the augmented matrices [[0, M], [M^H, 0]] used by the
// complex SVD.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dmrj.h"
#include "internal.h"

namespace dmrj {

#define HIP_OK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

const BufRef* BufMap::find(const double* p) const {
  for (const auto& r : b)
    if (p >= r.base && p < r.base + r.n) return &r;
  return nullptr;
}

static const BufRef& lookup(const BufMap& m, const double* p, const char* what) {
  const BufRef* r = m.find(p);
  if (!r) throw Error(DMRJ_ERR_STATE, std::string("complex plan: operand outside the known buffers (") + what + ")");
  return *r;
}

void complexify_gemm(std::vector<GemmTask>& tasks, std::vector<GemmSeg>& segs, const BufMap& m) {
  std::vector<GemmTask> nt;
  std::vector<GemmSeg> ns;
  nt.reserve(tasks.size() * 2);
  ns.reserve(segs.size() * 4);
  for (const GemmTask& t : tasks) {
    const BufRef& cb = lookup(m, t.C, "output");
    if (!cb.cplx) throw Error(DMRJ_ERR_STATE, "complex plan: output buffer is real");
    GemmTask re = t, im = t;
    im.C = t.C + cb.n;
    re.seg0 = (int)ns.size();
    re.nseg = 0;
    for (int q = 0; q < t.nseg; ++q) {
      const GemmSeg& g = segs[t.seg0 + q];
      const BufRef& ab = lookup(m, g.A, "A");
      const BufRef& bb = lookup(m, g.B, "B");
      const double sa = (g.flags & kSegConjA) ? -1.0 : 1.0, sb = (g.flags & kSegConjB) ? -1.0 : 1.0;
      GemmSeg x = g;
      x.flags = g.flags & 3;
      ns.push_back(x);  // Ar Br
      re.nseg++;
      if (ab.cplx && bb.cplx) {
        GemmSeg y = x;
        y.A = g.A + ab.n;
        y.B = g.B + bb.n;
        y.alpha = -sa * sb * g.alpha;  // - sA sB Ai Bi
        ns.push_back(y);
        re.nseg++;
      }
    }
    im.seg0 = (int)ns.size();
    im.nseg = 0;
    for (int q = 0; q < t.nseg; ++q) {
      const GemmSeg& g = segs[t.seg0 + q];
      const BufRef& ab = lookup(m, g.A, "A");
      const BufRef& bb = lookup(m, g.B, "B");
      const double sa = (g.flags & kSegConjA) ? -1.0 : 1.0, sb = (g.flags & kSegConjB) ? -1.0 : 1.0;
      GemmSeg x = g;
      x.flags = g.flags & 3;
      if (bb.cplx) {  // sB Ar Bi
        GemmSeg y = x;
        y.B = g.B + bb.n;
        y.alpha = sb * g.alpha;
        ns.push_back(y);
        im.nseg++;
      }
      if (ab.cplx) {  // sA Ai Br
        GemmSeg y = x;
        y.A = g.A + ab.n;
        y.alpha = sa * g.alpha;
        ns.push_back(y);
        im.nseg++;
      }
    }
    nt.push_back(re);
    nt.push_back(im);
  }
  tasks.swap(nt);
  segs.swap(ns);
}

void complexify_mix(std::vector<MixDst>& dsts, std::vector<MixTerm>& terms, const BufMap& m) {
  // real coefficients act on each plane separately
  const size_t nd = dsts.size();
  for (size_t i = 0; i < nd; ++i) {
    const MixDst d = dsts[i];
    const BufRef& db = lookup(m, d.dst, "mix output");
    if (!db.cplx) throw Error(DMRJ_ERR_STATE, "complex plan: mix output buffer is real");
    MixDst di{d.dst + db.n, d.n, (int)terms.size(), 0};
    for (int q = 0; q < d.nt; ++q) {
      const MixTerm tm = terms[d.t0 + q];
      const BufRef& sb = lookup(m, tm.src, "mix input");
      if (!sb.cplx) continue;
      terms.push_back({tm.src + sb.n, tm.coef});
      di.nt++;
    }
    dsts.push_back(di);
  }
}

// ------------------------------------------------------------------ kernels
// Z (N x N, complex interleaved) receives M in rows [0, R), columns [R, R+C):
// the upper triangle of [[0, M], [M^H, 0]].  Z is zero-filled beforehand.
__global__ void k_aug_place_z(const AugJobZ* __restrict__ jobs, const int2* __restrict__ work, int nwork) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const AugJobZ jb = jobs[w.x];
    const int col = w.y;  // column of M
    const int64_t RC = (int64_t)jb.R * jb.C;
    double2* z = reinterpret_cast<double2*>(jb.Z) + (int64_t)(jb.R + col) * jb.N;
    for (int r = threadIdx.x; r < jb.R; r += blockDim.x) {
      const int64_t e = (int64_t)r + (int64_t)col * jb.R;
      z[r] = make_double2(jb.M[e], jb.M[RC + e]);
    }
  }
}

// Eigenvector of eigenvalue sigma_i > 0 (index N-1-i, ascending order) is
// [u; v] / sqrt(2) with M v = sigma u: U[:, i] = u, Vt[i, :] = v^H, each
// half normalised separately.
__global__ __launch_bounds__(256) void k_aug_extract_z(const AugJobZ* __restrict__ jobs,
                                                       const int2* __restrict__ work, int nwork) {
  __shared__ double red[2][4];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const AugJobZ jb = jobs[w.x];
    const int i = w.y;
    const int col = jb.N - 1 - i;
    const double2* v = reinterpret_cast<const double2*>(jb.Z) + (int64_t)col * jb.N;
    double nt = 0.0, nb = 0.0;
    for (int r = threadIdx.x; r < jb.R; r += 256) nt += v[r].x * v[r].x + v[r].y * v[r].y;
    for (int r = threadIdx.x; r < jb.C; r += 256) {
      const double2 q = v[jb.R + r];
      nb += q.x * q.x + q.y * q.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
      nt += __shfl_down(nt, o, 64);
      nb += __shfl_down(nb, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = nt;
      red[1][threadIdx.x >> 6] = nb;
    }
    __syncthreads();
    const double st = sqrt(red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    const double sb = sqrt(red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    __syncthreads();
    const double it = st > 0.0 ? 1.0 / st : 0.0, ib = sb > 0.0 ? 1.0 / sb : 0.0;
    const int64_t nu = (int64_t)jb.R * jb.kmin, nv = (int64_t)jb.kmin * jb.C;
    for (int r = threadIdx.x; r < jb.R; r += 256) {
      const int64_t e = (int64_t)r + (int64_t)i * jb.R;
      jb.U[e] = v[r].x * it;
      jb.U[nu + e] = v[r].y * it;
    }
    for (int r = threadIdx.x; r < jb.C; r += 256) {
      const int64_t e = (int64_t)i + (int64_t)r * jb.kmin;
      const double2 q = v[jb.R + r];
      jb.Vt[e] = q.x * ib;
      jb.Vt[nv + e] = -q.y * ib;  // conjugate transpose
    }
    if (threadIdx.x == 0) jb.S[i] = fabs(jb.W[col]);
  }
}

void launch_aug_place_z(Context& c, const std::vector<AugJobZ>& jobs) {
  if (jobs.empty()) return;
  std::vector<int2> work;
  for (size_t j = 0; j < jobs.size(); ++j)
    for (int q = 0; q < jobs[j].C; ++q) work.push_back(make_int2((int)j, q));
  if (work.empty()) return;
  auto* dj = (AugJobZ*)c.upload(jobs.data(), jobs.size() * sizeof(AugJobZ));
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  hipLaunchKernelGGL(k_aug_place_z, dim3(grid), dim3(256), 0, c.stream, dj, dw, (int)work.size());
  HIP_OK(hipGetLastError());
}

void launch_aug_extract_z(Context& c, const std::vector<AugJobZ>& jobs) {
  if (jobs.empty()) return;
  std::vector<int2> work;
  for (size_t j = 0; j < jobs.size(); ++j)
    for (int i = 0; i < jobs[j].kmin; ++i) work.push_back(make_int2((int)j, i));
  if (work.empty()) return;
  auto* dj = (AugJobZ*)c.upload(jobs.data(), jobs.size() * sizeof(AugJobZ));
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  hipLaunchKernelGGL(k_aug_extract_z, dim3(grid), dim3(256), 0, c.stream, dj, dw, (int)work.size());
  HIP_OK(hipGetLastError());
}

}  // namespace dmrj
