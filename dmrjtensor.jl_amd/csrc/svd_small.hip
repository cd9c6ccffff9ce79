// Batched one-sided Jacobi (Hestenes) SVD for small charge sectors.
//
// One workgroup per sector matrix M (R x C, column-major, ld R).  With
// n = min(R, C) <= 64 the working matrix X (rows x n, the columns of M or of
// M^T) and the accumulated rotation V (n x n) live in LDS.  Column pairs are
// rotated in the round-robin (circle) order, n/2 disjoint pairs per round,
// one group of G lanes per pair; sweeps repeat until no pair has
// |x_p . x_q| > eps sqrt(|x_p|^2 |x_q|^2).  Output has the layout of the LAPACK
// gesdd call it replaces for these sizes: S descending, U (R x n, ld R),
// Vt (n x C, ld n).
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.h"

namespace dmrj {

constexpr int kSvdThreads = 256;

__device__ inline double group_sum(double v, int G) {
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kSvdThreads) void k_svd_small(const SmallSvdJob* __restrict__ jobs,
                                                           int njobs) {
  extern __shared__ double lds[];
  const SmallSvdJob jb = jobs[blockIdx.x];
  const bool tr = jb.R < jb.C;             // work on M^T when M is wide
  const int rows = tr ? jb.C : jb.R;       // length of a working column
  const int n = tr ? jb.R : jb.C;          // number of working columns
  const int np = (n + 1) & ~1;             // even number of players
  double* X = lds;                          // [np][rows]
  double* V = lds + (size_t)np * rows;      // [np][n]
  __shared__ double sig[64];
  __shared__ int perm[64];
  __shared__ int rotated;
  const int tid = threadIdx.x;
  // load
  for (int e = tid; e < np * rows; e += kSvdThreads) {
    const int col = e / rows, r = e % rows;
    double v = 0.0;
    if (col < n) v = tr ? jb.M[(int64_t)col + (int64_t)r * jb.R] : jb.M[(int64_t)r + (int64_t)col * jb.R];
    X[e] = v;
  }
  for (int e = tid; e < np * n; e += kSvdThreads) V[e] = ((e / n) == (e % n)) ? 1.0 : 0.0;
  __syncthreads();
  const int P = np / 2;
  int G = 64;
  while (G > 1 && G * P > kSvdThreads) G >>= 1;
  const int grp = tid / G, gl = tid % G;
  const double eps = 2.220446049250313e-16;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int rd = 0; rd < np - 1; ++rd) {
      if (grp < P) {
        auto player = [&](int k) { return k == 0 ? 0 : ((k - 1 + rd) % (np - 1)) + 1; };
        int p = player(grp), q = player(np - 1 - grp);
        if (p > q) { const int t = p; p = q; q = t; }
        if (q < n) {
          double* xp = X + (size_t)p * rows;
          double* xq = X + (size_t)q * rows;
          double a = 0.0, b = 0.0, g = 0.0;
          for (int r = gl; r < rows; r += G) {
            const double u = xp[r], w = xq[r];
            a += u * u;
            b += w * w;
            g += u * w;
          }
          a = group_sum(a, G);
          b = group_sum(b, G);
          g = group_sum(g, G);
          if (fabs(g) > eps * sqrt(a * b) && g != 0.0) {
            const double zeta = (b - a) / (2.0 * g);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
            for (int r = gl; r < rows; r += G) {
              const double u = xp[r], w = xq[r];
              xp[r] = c * u - s * w;
              xq[r] = s * u + c * w;
            }
            double* vp = V + (size_t)p * n;
            double* vq = V + (size_t)q * n;
            for (int r = gl; r < n; r += G) {
              const double u = vp[r], w = vq[r];
              vp[r] = c * u - s * w;
              vq[r] = s * u + c * w;
            }
            if (gl == 0) rotated = 1;
          }
        }
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  // singular values = column norms
  for (int col = grp; grp < P && col < n; col += P) {
    double a = 0.0;
    const double* x = X + (size_t)col * rows;
    for (int r = gl; r < rows; r += G) a += x[r] * x[r];
    a = group_sum(a, G);
    if (gl == 0) sig[col] = sqrt(a);
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 1; i < n; ++i) {  // stable insertion sort, descending
      const int pi = perm[i];
      int j = i - 1;
      while (j >= 0 && sig[perm[j]] < sig[pi]) {
        perm[j + 1] = perm[j];
        --j;
      }
      perm[j + 1] = pi;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += kSvdThreads) jb.S[i] = sig[perm[i]];
  // left vectors of the working matrix: X[:,k]/sigma_k (rows); right: V[:,k] (n)
  // M = X V^T (working) -> if !tr: U = X/sigma (R x n), Vt = V^T (n x C)
  //                       if  tr: M^T = X V^T -> M = V X^T: U = V (R x n), Vt = (X/sigma)^T (n x C)
  for (int e = tid; e < n * rows; e += kSvdThreads) {
    const int k = e / rows, r = e % rows;
    const int src = perm[k];
    const double sg = sig[src];
    const double val = sg > 0.0 ? X[(size_t)src * rows + r] / sg : 0.0;
    if (!tr)
      jb.U[(int64_t)r + (int64_t)k * jb.R] = val;      // U[r, k], ld R
    else
      jb.Vt[(int64_t)k + (int64_t)r * n] = val;        // Vt[k, r], ld n
  }
  for (int e = tid; e < n * n; e += kSvdThreads) {
    const int k = e / n, r = e % n;
    const double val = V[(size_t)perm[k] * n + r];
    if (!tr)
      jb.Vt[(int64_t)k + (int64_t)r * n] = val;        // Vt[k, r] = V[r, k]
    else
      jb.U[(int64_t)r + (int64_t)k * jb.R] = val;      // U[r, k] = V[r, k]
  }
}

// Singular triplets from the eigenpairs of the augmented symmetric matrix
// [[0, M], [M^T, 0]] (size N >= R + C, zero padded): eigenvalue +sigma_i with
// eigenvector (u_i; v_i)/sqrt(2).  Eigenvalues are ascending, so sigma_i is
// column N-1-i.  Each half is normalised by its own norm.
__global__ __launch_bounds__(256) void k_aug_extract(const AugJob* __restrict__ jobs,
                                                     const int2* __restrict__ work, int nwork) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const AugJob jb = jobs[w.x];
    const int i = w.y;
    const int col = jb.N - 1 - i;
    const double* v = jb.A + (int64_t)col * jb.N;
    double nt = 0.0, nb = 0.0;
    for (int r = threadIdx.x; r < jb.R; r += 256) nt += v[r] * v[r];
    for (int r = threadIdx.x; r < jb.C; r += 256) nb += v[jb.R + r] * v[jb.R + r];
    __shared__ double red[2][4];
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
    for (int r = threadIdx.x; r < jb.R; r += 256) jb.U[(int64_t)r + (int64_t)i * jb.R] = v[r] * it;
    for (int r = threadIdx.x; r < jb.C; r += 256) jb.Vt[(int64_t)i + (int64_t)r * jb.kmin] = v[jb.R + r] * ib;
    if (threadIdx.x == 0) jb.S[i] = fabs(jb.W[col]);
  }
}

void launch_aug_extract(Context& c, const std::vector<AugJob>& jobs) {
  if (jobs.empty()) return;
  std::vector<int2> work;
  for (size_t j = 0; j < jobs.size(); ++j)
    for (int i = 0; i < jobs[j].kmin; ++i) work.push_back(make_int2((int)j, i));
  auto* dj = (AugJob*)c.upload(jobs.data(), jobs.size() * sizeof(AugJob));
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  hipLaunchKernelGGL(k_aug_extract, dim3(grid), dim3(256), 0, c.stream, dj, dw, (int)work.size());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
}

size_t small_svd_lds(int R, int C) {
  const int rows = R < C ? C : R;
  const int n = R < C ? R : C;
  const int np = (n + 1) & ~1;
  return ((size_t)np * rows + (size_t)np * n) * sizeof(double);
}

// Launches on the context's side stream after everything queued on the main
// stream so far (inputs, the job table upload); the caller joins with
// hipStreamWaitEvent(c.stream, c.side_out) before reading the results.
// Returns whether a kernel was launched.
bool launch_svd_small(Context& c, const std::vector<SmallSvdJob>& jobs, size_t lds_bytes) {
  if (jobs.empty()) return false;
  auto* dj = (SmallSvdJob*)c.upload(jobs.data(), jobs.size() * sizeof(SmallSvdJob));
  auto ok = [](hipError_t e) {
    if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
  };
  ok(hipEventRecord(c.side_in, c.stream));
  ok(hipStreamWaitEvent(c.side, c.side_in, 0));
  hipLaunchKernelGGL(k_svd_small, dim3((unsigned)jobs.size()), dim3(kSvdThreads), lds_bytes, c.side, dj,
                     (int)jobs.size());
  ok(hipGetLastError());
  ok(hipEventRecord(c.side_out, c.side));
  return true;
}

}  // namespace dmrj
