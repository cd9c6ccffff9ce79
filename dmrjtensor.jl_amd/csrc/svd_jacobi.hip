// Batched block one-sided Jacobi for the larger charge sectors of the
// two-site SVD.
//
// The working matrix X (rows x ncols, column-major, ld rows) is M when the
// left factor is wanted orthonormal (split (U, D V^T)) and M^T otherwise.
// One-sided (Hestenes) Jacobi rotates column pairs of X until all columns are
// mutually orthogonal: X W = U D with W orthogonal, never formed.  Columns are
// grouped into blocks of bw; one kernel launch per round of a round-robin
// (circle) schedule over the blocks pairs every block once, nblk/2 disjoint
// block pairs per sector and all sectors in the same launch.  A work-group
// loads its block pair (rows x 2bw) into LDS and orthogonalises those columns
// among themselves with inner circle rounds (one group of lanes per column
// pair, 3 dot products and a plane rotation), then writes them back.  Rounds
// repeat as outer sweeps until a sweep rotates nothing beyond
// |x_p . x_q| <= tol sqrt(|x_p|^2 |x_q|^2), tol = rows eps (one of the
// relative thresholds of LAPACK's dgesvj; the rounding of a computed dot
// product of two orthogonal columns is of that order, so a stricter threshold
// keeps rotating pairs that are already orthogonal).  Columns whose norm fell to the rounding
// level of the whole tensor (8 eps sqrt(rows) |AA|) are numerically zero and
// take no further rotations.  Before every sweep the host orders the columns
// by decreasing norm (de Rijk's pivoting: about half the sweeps on DMRG
// sectors); the order enters as a permutation, the data does not move.  A
// sector that does not converge in 60 sweeps is reported as a solver error.  Column norms are then the singular values
// and the normalised columns the singular vectors of that side; the other
// factor comes from one grouped GEMM with M (U^T M = D V^T or M V = U D), so
// W is never accumulated.
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

constexpr int kBjThreads = 256;

__device__ inline double bj_group_sum(double v, int G) {
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// circle schedule: round s of n players (n even), pair k
__device__ inline void circle_pair(int n, int s, int k, int& p, int& q) {
  if (k == 0) {
    p = s;
    q = n - 1;
  } else {
    p = (s + k) % (n - 1);
    q = (s - k + (n - 1)) % (n - 1);
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

__global__ __launch_bounds__(kBjThreads) void k_bj_step(const BjJob* __restrict__ jobs, const int2* __restrict__ work,
                                                        int nwork, int step, int inner_max) {
  extern __shared__ double panel[];  // [2bw][rows]
  __shared__ int rot_any;
  const int tid = threadIdx.x;
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const BjJob jb = jobs[w.x];
    if (step >= jb.nblk - 1) continue;
    int bp, bq;
    circle_pair(jb.nblk, step, w.y, bp, bq);
    const int p0 = bp * jb.bw, q0 = bq * jb.bw;
    const int np_ = max(0, min(jb.bw, jb.ncols - p0)), nq_ = max(0, min(jb.bw, jb.ncols - q0));
    if (np_ == 0 || nq_ == 0) continue;  // paired with an empty block this round
    const int nc = np_ + nq_;
    const int rows = jb.rows;
    // columns 0..np_-1 from block p, np_..nc-1 from block q
    for (int e = tid; e < nc * rows; e += kBjThreads) {
      const int c = e / rows, r = e % rows;
      const int gc = jb.perm[c < np_ ? p0 + c : q0 + (c - np_)];
      panel[e] = jb.X[(int64_t)r + (int64_t)gc * rows];
    }
    __syncthreads();
    const int ncp = (nc + 1) & ~1;  // even number of players
    const int P = ncp / 2;
    int G = 64;
    while (G > 1 && G * P > kBjThreads) G >>= 1;
    const int grp = tid / G, gl = tid % G;
    // rotate only pairs whose cosine exceeds rows * eps: a computed dot
    // product of two orthogonal columns carries rounding of that order, so a
    // stricter threshold keeps rotating already orthogonal pairs
    const double tol = 2.220446049250313e-16 * (double)rows;
    bool any_step = false;
    for (int sweep = 0; sweep < inner_max; ++sweep) {
      if (tid == 0) rot_any = 0;
      __syncthreads();
      for (int rd = 0; rd < ncp - 1; ++rd) {
        if (grp < P) {
          int p, q;
          circle_pair(ncp, rd, grp, p, q);
          if (q < nc) {
            double* xp = panel + (size_t)p * rows;
            double* xq = panel + (size_t)q * rows;
            double a = 0.0, b = 0.0, g = 0.0;
            for (int r = gl; r < rows; r += G) {
              const double u = xp[r], v = xq[r];
              a += u * u;
              b += v * v;
              g += u * v;
            }
            a = bj_group_sum(a, G);
            b = bj_group_sum(b, G);
            g = bj_group_sum(g, G);
            if (g != 0.0 && fabs(g) > tol * sqrt(a * b) && fmin(a, b) > jb.tiny2) {
              const double zeta = (b - a) / (2.0 * g);
              // t = tan of the rotation angle; 1/(2 zeta) where zeta^2 would overflow
              const double t = fabs(zeta) > 1e150
                                   ? 0.5 / zeta
                                   : (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
              const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
              for (int r = gl; r < rows; r += G) {
                const double u = xp[r], v = xq[r];
                xp[r] = c * u - s * v;
                xq[r] = s * u + c * v;
              }
              if (gl == 0) rot_any = 1;
            }
          }
        }
        __syncthreads();
      }
      const int any = rot_any;
      __syncthreads();
      if (!any) break;
      any_step = true;
    }
    for (int e = tid; e < nc * rows; e += kBjThreads) {
      const int c = e / rows, r = e % rows;
      const int gc = jb.perm[c < np_ ? p0 + c : q0 + (c - np_)];
      jb.X[(int64_t)r + (int64_t)gc * rows] = panel[e];
    }
    if (any_step && tid == 0) atomicOr(jb.flag, 1);
    __syncthreads();
  }
}

// column norms of X into S (unsorted)
__global__ __launch_bounds__(kBjThreads) void k_bj_norms(const BjJob* __restrict__ jobs, const int2* __restrict__ work,
                                                         int nwork) {
  __shared__ double red[4];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const BjJob jb = jobs[w.x];
    const double* x = jb.X + (int64_t)w.y * jb.rows;
    double a = 0.0;
    for (int r = threadIdx.x; r < jb.rows; r += kBjThreads) a += x[r] * x[r];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) jb.S[w.y] = sqrt(red[0] + red[1] + red[2] + red[3]);
    __syncthreads();
  }
}

// normalised column perm[k] of X -> column k of D (trans = 0: D[r + k ldd])
// or row k of D (trans = 1: D[k + r ldd])
__global__ __launch_bounds__(kBjThreads) void k_bj_gather(const BjGather* __restrict__ g, int ng) {
  for (int wi = blockIdx.x; wi < ng; wi += gridDim.x) {
    const BjGather o = g[wi];
    const double inv = o.sigma > 0.0 ? 1.0 / o.sigma : 0.0;
    for (int r = threadIdx.x; r < o.rows; r += kBjThreads) {
      const double v = o.src[r] * inv;
      if (o.trans)
        o.dst[(int64_t)r * o.ldd] = v;
      else
        o.dst[r] = v;
    }
  }
}

__global__ __launch_bounds__(kBjThreads) void k_transpose(const double* __restrict__ M, int R, int C,
                                                           double* __restrict__ X) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int tilesR = (R + 31) / 32, tilesC = (C + 31) / 32;
  for (int t = blockIdx.x; t < tilesR * tilesC; t += gridDim.x) {
    const int r0 = (t % tilesR) * 32, c0 = (t / tilesR) * 32;
    for (int j = ty; j < 32; j += 8) {
      const int r = r0 + tx, cc = c0 + j;
      tile[j][tx] = (r < R && cc < C) ? M[(int64_t)r + (int64_t)cc * R] : 0.0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
      const int cc = c0 + tx, r = r0 + j;  // X[cc + r*C] = M[r + cc*R]
      if (cc < C && r < R) X[(int64_t)cc + (int64_t)r * C] = tile[tx][j];
    }
    __syncthreads();
  }
}

void launch_transpose(Context& c, const double* M, int R, int C, double* X) {
  if (R <= 0 || C <= 0) return;
  const int tiles = ((R + 31) / 32) * ((C + 31) / 32);
  hipLaunchKernelGGL(k_transpose, dim3(std::min(tiles, c.num_cu * 8)), dim3(kBjThreads), 0, c.stream, M, R, C, X);
  HIP_OK(hipGetLastError());
}

// diagonal entries d in [n, N) of the N x N matrix b of a strided batch
__global__ void k_pad_diag(double* __restrict__ Z, int N, const int2* __restrict__ jobs, int njobs, double v) {
  const int64_t NN = (int64_t)N * N;
  for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
    const int2 jb = jobs[j];
    for (int d = jb.y + threadIdx.x; d < N; d += blockDim.x) Z[jb.x * NN + d + (int64_t)d * N] = v;
  }
}

void launch_pad_diag(Context& c, double* Z, int N, const std::vector<int2>& jobs, double v) {
  if (jobs.empty()) return;
  auto* dj = (int2*)c.upload(jobs.data(), jobs.size() * sizeof(int2));
  hipLaunchKernelGGL(k_pad_diag, dim3((unsigned)std::min<size_t>(jobs.size(), 1024)), dim3(kBjThreads), 0, c.stream, Z,
                     N, dj, (int)jobs.size(), v);
  HIP_OK(hipGetLastError());
}

// Lower-triangular K x K matrix B written into columns [K, 2K), rows [0, K)
// of an N x N augmented matrix (the block the upper-triangle eigensolver
// reads): B[r][c] = X[r + c ldx] (the L of an LQ factorisation) or, with
// trans, X[c + r ldx] (the transpose of the R of a QR factorisation), zero
// above the diagonal (where the factorisation keeps its reflectors).
// One work item per (job, column c).
__global__ void k_tri_place(const TriJob* __restrict__ jobs, const int2* __restrict__ work, int nwork) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int2 w = work[wi];
    const TriJob jb = jobs[w.x];
    const int c = w.y;
    double* dst = jb.A + (int64_t)(jb.K + c) * jb.N;
    for (int r = threadIdx.x; r < jb.K; r += blockDim.x) {
      double v = 0.0;
      if (r >= c) v = jb.trans ? jb.X[c + (int64_t)r * jb.ldx] : jb.X[r + (int64_t)c * jb.ldx];
      dst[r] = v;
    }
  }
}

void launch_tri_place(Context& c, const std::vector<TriJob>& jobs) {
  if (jobs.empty()) return;
  std::vector<int2> work;
  for (size_t j = 0; j < jobs.size(); ++j)
    for (int col = 0; col < jobs[j].K; ++col) work.push_back(make_int2((int)j, col));
  auto* dj = (TriJob*)c.upload(jobs.data(), jobs.size() * sizeof(TriJob));
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  hipLaunchKernelGGL(k_tri_place, dim3((unsigned)std::min<size_t>(work.size(), (size_t)c.num_cu * 8)),
                     dim3(kBjThreads), 0, c.stream, dj, dw, (int)work.size());
  HIP_OK(hipGetLastError());
}

// Orthonormal completion of a kept canonical factor.  Singular vectors of
// numerically zero singular values are not determined to working precision
// by any of the solvers here (the two halves of an augmented eigenvector of
// the null space, or a Jacobi column at the rounding level, are not
// orthogonal to the rest), while the reference's LAPACK SVD returns an
// orthonormal completion.  Vectors c0 .. keep-1 (element r of vector c at
// Q[c vs + r es]) are replaced by an orthonormal basis of part of the
// orthogonal complement of vectors 0 .. c0-1.  They carry zero weight, so the
// state is unchanged; the factor becomes orthonormal.
//
// Blocked greedy construction, kCB vectors at a time:
//   * candidates are the unit vectors e_j with the largest residuals
//     d_j = 1 - sum_p Q_pj^2 against the vectors so far (tracked for all j;
//     the largest is at least (n - c) / n);
//   * block classical Gram-Schmidt, twice, against the vectors so far:
//     W -= Q (Q^T W), four grouped-GEMM launches over all sectors of a bond;
//   * within the block, modified Gram-Schmidt (twice) over the candidates in
//     order of decreasing d; a candidate whose residual fell below 1e-4, or
//     below a tenth of its d, is dropped;
//   * dropped places are filled one vector at a time by the exact greedy step
//     (e_j of largest d, two classical Gram-Schmidt passes against all vectors
//     so far), so every block completes exactly its vectors.
// One work-group per sector in k_complete_block; per sector scratch W (n x
// kCB), coef (keep x kCB), d (n), seld (kCB) and cand (kCB ints).
constexpr int kCB = 32;

__device__ double cb_block_sum(double a, double* red) {
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  const double t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

// index of the largest d_j over j not in excl[0 .. nex); ties go to the
// lowest index.  Requires n > nex.
__device__ int cb_argmax(const double* d, int n, const int* excl, int nex, double* bv, int* bi) {
  double mv = -2.0;
  int mi = n;
  for (int j = threadIdx.x; j < n; j += kBjThreads) {
    bool skip = false;
    for (int t = 0; t < nex; ++t) skip |= excl[t] == j;
    if (!skip && d[j] > mv) {
      mv = d[j];
      mi = j;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_down(mv, o, 64);
    const int oi = __shfl_down(mi, o, 64);
    if (ov > mv || (ov == mv && oi < mi)) {
      mv = ov;
      mi = oi;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    bv[threadIdx.x >> 6] = mv;
    bi[threadIdx.x >> 6] = mi;
  }
  __syncthreads();
  double best_v = bv[0];
  int best = bi[0];
  for (int w = 1; w < 4; ++w)
    if (bv[w] > best_v || (bv[w] == best_v && bi[w] < best)) {
      best_v = bv[w];
      best = bi[w];
    }
  __syncthreads();
  return best;
}

// Exact greedy step: v = e_j after two classical Gram-Schmidt passes against
// vectors 0 .. c-1 of Q; returns |v|.  Loops are ordered so that consecutive
// lanes read consecutive addresses for both layouts (es == 1: vectors are
// columns; vs == 1: rows).
__device__ double complete_candidate(const CompleteJob& jb, double* v, double* coef, int c, int j, double* red) {
  const int n = jb.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kWaves = kBjThreads / 64;
  for (int r = threadIdx.x; r < n; r += kBjThreads) v[r] = (r == j) ? 1.0 : 0.0;
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    if (jb.es == 1) {
      for (int p = wave; p < c; p += kWaves) {
        const double* q = jb.Q + (int64_t)p * jb.vs;
        double a = 0.0;
        for (int r = lane; r < n; r += 64) a += q[r] * v[r];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) coef[p] = a;
      }
    } else {
      for (int p = threadIdx.x; p < c; p += kBjThreads) {
        double a = 0.0;
        for (int r = 0; r < n; ++r) a += jb.Q[(int64_t)p * jb.vs + (int64_t)r * jb.es] * v[r];
        coef[p] = a;
      }
    }
    __syncthreads();
    if (jb.es == 1) {
      for (int r = threadIdx.x; r < n; r += kBjThreads) {
        double a = v[r];
        for (int p = 0; p < c; ++p) a -= coef[p] * jb.Q[(int64_t)p * jb.vs + r];
        v[r] = a;
      }
    } else {
      for (int r = wave; r < n; r += kWaves) {
        const double* q = jb.Q + (int64_t)r * jb.es;  // element r of every vector, contiguous over p
        double a = 0.0;
        for (int p = lane; p < c; p += 64) a += coef[p] * q[p];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) v[r] -= a;
      }
    }
    __syncthreads();
  }
  double a = 0.0;
  for (int r = threadIdx.x; r < n; r += kBjThreads) a += v[r] * v[r];
  return sqrt(cb_block_sum(a, red));
}

// Q vectors c .. c+s-1 = W columns 0 .. s-1; d_r -= sum of their squares.
__device__ void cb_store(const CompleteJob& jb, int c, int s) {
  const int n = jb.n;
  if (s <= 0) return;
  if (jb.es == 1) {
    for (int u = 0; u < s; ++u)
      for (int r = threadIdx.x; r < n; r += kBjThreads)
        jb.Q[(int64_t)(c + u) * jb.vs + r] = jb.W[(int64_t)u * n + r];
  } else {
    for (int64_t e = threadIdx.x; e < (int64_t)s * n; e += kBjThreads) {
      const int u = (int)(e % s), r = (int)(e / s);
      jb.Q[(int64_t)(c + u) * jb.vs + (int64_t)r * jb.es] = jb.W[(int64_t)u * n + r];
    }
  }
  for (int r = threadIdx.x; r < n; r += kBjThreads) {
    double a = jb.d[r];
    for (int u = 0; u < s; ++u) {
      const double q = jb.W[(int64_t)u * n + r];
      a -= q * q;
    }
    jb.d[r] = a;
  }
  __syncthreads();
}

// b < 0: residuals d of all unit vectors and the first candidate block.
// b >= 0: finish block b (its candidates in W have been projected against
// vectors 0 .. c0 + kCB b - 1), then set up the candidates of block b + 1.
__global__ __launch_bounds__(kBjThreads) void k_complete_block(const CompleteJob* __restrict__ jobs, int njobs,
                                                               int b) {
  __shared__ double red[4];
  __shared__ double bv[4];
  __shared__ int bi[4];
  __shared__ int cand_sh[kCB];
  for (int ji = blockIdx.x; ji < njobs; ji += gridDim.x) {
    const CompleteJob jb = jobs[ji];
    const int n = jb.n;
    int cnext;
    if (b < 0) {
      if (jb.es == 1) {
        for (int j = threadIdx.x; j < n; j += kBjThreads) {
          double a = 1.0;
          for (int p = 0; p < jb.c0; ++p) {
            const double q = jb.Q[(int64_t)p * jb.vs + j];
            a -= q * q;
          }
          jb.d[j] = a;
        }
      } else {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int j = wave; j < n; j += kBjThreads / 64) {
          const double* q = jb.Q + (int64_t)j * jb.es;  // contiguous over p
          double a = 0.0;
          for (int p = lane; p < jb.c0; p += 64) a += q[p] * q[p];
          for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
          if (lane == 0) jb.d[j] = 1.0 - a;
        }
      }
      __syncthreads();
      cnext = jb.c0;
    } else {
      const int cb = jb.c0 + kCB * b;
      if (cb >= jb.keep) continue;
      const int nb = min(kCB, jb.keep - cb);
      int s = 0;
      for (int t = 0; t < nb; ++t) {
        double* wt = jb.W + (int64_t)t * n;
        for (int pass = 0; pass < 2; ++pass)
          for (int u = 0; u < s; ++u) {
            const double* wu = jb.W + (int64_t)u * n;
            double a = 0.0;
            for (int r = threadIdx.x; r < n; r += kBjThreads) a += wu[r] * wt[r];
            a = cb_block_sum(a, red);
            for (int r = threadIdx.x; r < n; r += kBjThreads) wt[r] -= a * wu[r];
            __syncthreads();
          }
        double a = 0.0;
        for (int r = threadIdx.x; r < n; r += kBjThreads) a += wt[r] * wt[r];
        const double nrm2 = cb_block_sum(a, red);
        if (nrm2 >= 1e-4 && nrm2 >= 0.1 * jb.seld[t]) {
          const double inv = 1.0 / sqrt(nrm2);
          double* ws = jb.W + (int64_t)s * n;
          for (int r = threadIdx.x; r < n; r += kBjThreads) ws[r] = wt[r] * inv;
          __syncthreads();
          ++s;
        }
      }
      cb_store(jb, cb, s);
      // places left by dropped candidates: exact greedy steps against all
      // vectors so far (W is free now: column 0 serves as the work vector)
      for (int col = cb + s; col < cb + nb; ++col) {
        const int j = cb_argmax(jb.d, n, cand_sh, 0, bv, bi);
        const double nrm = complete_candidate(jb, jb.W, jb.coef, col, j, red);
        const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        for (int r = threadIdx.x; r < n; r += kBjThreads) jb.W[r] *= inv;
        __syncthreads();
        cb_store(jb, col, 1);
      }
      cnext = cb + nb;
    }
    const int nbn = min(kCB, jb.keep - cnext);
    if (nbn <= 0) continue;
    for (int t = 0; t < nbn; ++t) {
      const int j = cb_argmax(jb.d, n, cand_sh, t, bv, bi);
      if (threadIdx.x == 0) {
        cand_sh[t] = j;
        jb.cand[t] = j;
        jb.seld[t] = jb.d[j];
      }
      __syncthreads();
    }
    for (int64_t e = threadIdx.x; e < (int64_t)nbn * n; e += kBjThreads) jb.W[e] = 0.0;
    __syncthreads();
    if (threadIdx.x < nbn) jb.W[(int64_t)threadIdx.x * n + cand_sh[threadIdx.x]] = 1.0;
    __syncthreads();
  }
}

void complete_basis(Context& c, std::vector<CompleteJob>& jobs) {
  if (jobs.empty()) return;
  std::vector<double*> scratch;
  int nblk = 0;
  for (CompleteJob& j : jobs) {
    if (j.c0 < 0 || j.keep > j.n || j.c0 >= j.keep) throw Error(DMRJ_ERR_STATE, "completion job out of range");
    const int64_t sz = (int64_t)kCB * j.n + (int64_t)kCB * j.keep + j.n + 2 * kCB;
    double* p = (double*)c.alloc.alloc(sz * (int64_t)sizeof(double));
    scratch.push_back(p);
    j.W = p;
    j.coef = j.W + (int64_t)kCB * j.n;
    j.d = j.coef + (int64_t)kCB * j.keep;
    j.seld = j.d + j.n;
    j.cand = reinterpret_cast<int*>(j.seld + kCB);
    nblk = std::max(nblk, (j.keep - j.c0 + kCB - 1) / kCB);
  }
  auto* dj = (CompleteJob*)c.upload(jobs.data(), jobs.size() * sizeof(CompleteJob));
  const unsigned grid = (unsigned)std::min<size_t>(jobs.size(), 1024);
  hipLaunchKernelGGL(k_complete_block, dim3(grid), dim3(kBjThreads), 0, c.stream, dj, (int)jobs.size(), -1);
  HIP_OK(hipGetLastError());
  for (int b = 0; b < nblk; ++b) {
    // W -= Q (Q^T W) twice; op(Q^T) is cb x n, op(Q) n x cb
    std::vector<GemmTask> t1, t2;
    std::vector<GemmSeg> s1, s2;
    for (const CompleteJob& j : jobs) {
      const int cb = j.c0 + kCB * b;
      if (cb >= j.keep || cb == 0) continue;
      const int nb = std::min(kCB, j.keep - cb);
      const bool cols = j.es == 1;
      const int64_t ldq = cols ? j.vs : j.es;
      t1.push_back({j.coef, j.keep, cb, nb, (int)s1.size(), 1, 0.0});
      s1.push_back({j.Q, j.W, ldq, j.n, j.n, cols ? 1 : 0, 1.0});
      t2.push_back({j.W, j.n, j.n, nb, (int)s2.size(), 1, 1.0});
      s2.push_back({j.Q, j.coef, ldq, j.keep, cb, cols ? 0 : 1, -1.0});
    }
    if (!t1.empty())
      for (int pass = 0; pass < 2; ++pass) {
        launch_gemm(c, t1, s1, -1);
        launch_gemm(c, t2, s2, -1);
      }
    hipLaunchKernelGGL(k_complete_block, dim3(grid), dim3(kBjThreads), 0, c.stream, dj, (int)jobs.size(), b);
    HIP_OK(hipGetLastError());
  }
  for (double* p : scratch) c.alloc.release(p);
}

// Re-orthonormalisation of kept vectors c0 .. keep-1 of a canonical factor
// whose singular values are small (their vectors, taken from an augmented
// eigenvector or a Jacobi column, are orthogonal to the rest only to about
// eps / sigma, at most ~1e-6 above the completion threshold).  All k of them
// at once, per sector:
//   W = Q[:, c0:keep];  twice  W -= Q0 (Q0^T W)   (Q0 = the first c0 vectors)
//   three times  G = W^T W,  W <- W (I - F),  F = strict_upper(G) + diag(G - I) / 2
// The second step is the first-order form of the Cholesky-QR update
// W <- W R^-1 (G = R^T R, R = I + F + O((G-I)^2)): F + F^T = G - I, so
// (I - F)^T G (I - F) = I + O((G-I)^2) and each pass squares the deviation
// from orthonormality.  I - F is upper triangular, so vector j moves only
// along vectors before it and by its own norm, as in Gram-Schmidt.  Every
// product is one grouped-GEMM launch over all sectors.
struct ReorthJob {
  double* Q;
  int64_t vs, es;
  int n, c0, k;
  double* W;  // n x k (column-major)
  double* V;  // n x k, the other buffer of the W <- W (I - F) passes
  double* C;  // c0 x k
  double* G;  // k x k
  double* T;  // k x k
};
constexpr int kReorthPasses = 3;

// to_q = 0: W = Q[:, c0:keep]; to_q = 1: Q[:, c0:keep] = W.  Job blockIdx.y.
__global__ __launch_bounds__(256) void k_reorth_move(const ReorthJob* __restrict__ jobs, int to_q) {
  const ReorthJob jb = jobs[blockIdx.y];
  const int64_t tot = (int64_t)jb.n * jb.k;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    int u, r;
    if (jb.es == 1) {  // vectors are columns: walk rows fastest
      u = (int)(e / jb.n);
      r = (int)(e % jb.n);
    } else {  // vectors are rows (stride es between elements): walk vectors fastest
      r = (int)(e / jb.k);
      u = (int)(e % jb.k);
    }
    double* q = jb.Q + (int64_t)(jb.c0 + u) * jb.vs + (int64_t)r * jb.es;
    double* w = jb.W + (int64_t)u * jb.n + r;
    if (to_q)
      *q = *w;
    else
      *w = *q;
  }
}

// T = I - F from G (both k x k column-major): T[u][t] = -G[u][t] (u < t),
// 1.5 - G[t][t] / 2 (u = t), 0 (u > t).  Job blockIdx.y.
__global__ __launch_bounds__(256) void k_reorth_tri(const ReorthJob* __restrict__ jobs) {
  const ReorthJob jb = jobs[blockIdx.y];
  const int64_t tot = (int64_t)jb.k * jb.k;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int u = (int)(e % jb.k), t = (int)(e / jb.k);
    const double g = jb.G[e];
    jb.T[e] = u < t ? -g : (u == t ? 1.5 - 0.5 * g : 0.0);
  }
}

static unsigned reorth_chunks(const std::vector<ReorthJob>& rj, bool square) {
  int64_t mx = 1;
  for (const ReorthJob& j : rj) mx = std::max<int64_t>(mx, square ? (int64_t)j.k * j.k : (int64_t)j.n * j.k);
  return (unsigned)std::min<int64_t>((mx + 255) / 256, 256);
}

void reorth_basis(Context& c, std::vector<CompleteJob>& jobs) {
  if (jobs.empty()) return;
  if (jobs.size() > 65535) throw Error(DMRJ_ERR_STATE, "too many sectors for re-orthonormalisation");
  std::vector<ReorthJob> rj;
  std::vector<double*> scratch;
  for (const CompleteJob& j : jobs) {
    if (j.c0 < 1 || j.keep > j.n || j.c0 >= j.keep) throw Error(DMRJ_ERR_STATE, "re-orthonormalisation job out of range");
    const int k = j.keep - j.c0;
    const int64_t nk = (int64_t)j.n * k, kk = (int64_t)k * k;
    double* p = (double*)c.alloc.alloc((2 * nk + (int64_t)j.c0 * k + 2 * kk) * (int64_t)sizeof(double));
    scratch.push_back(p);
    rj.push_back({j.Q, j.vs, j.es, j.n, j.c0, k, p, p + nk, p + 2 * nk, p + 2 * nk + (int64_t)j.c0 * k,
                  p + 2 * nk + (int64_t)j.c0 * k + kk});
  }
  const unsigned nj = (unsigned)rj.size();
  auto* dj = (ReorthJob*)c.upload(rj.data(), rj.size() * sizeof(ReorthJob));
  hipLaunchKernelGGL(k_reorth_move, dim3(reorth_chunks(rj, false), nj), dim3(256), 0, c.stream, dj, 0);
  HIP_OK(hipGetLastError());
  {
    // W -= Q0 (Q0^T W), twice
    std::vector<GemmTask> t1, t2;
    std::vector<GemmSeg> s1, s2;
    for (const ReorthJob& j : rj) {
      const bool cols = j.es == 1;
      const int64_t ldq = cols ? j.vs : j.es;
      t1.push_back({j.C, j.c0, j.c0, j.k, (int)s1.size(), 1, 0.0});
      s1.push_back({j.Q, j.W, ldq, j.n, j.n, cols ? 1 : 0, 1.0});
      t2.push_back({j.W, j.n, j.n, j.k, (int)s2.size(), 1, 1.0});
      s2.push_back({j.Q, j.C, ldq, j.c0, j.c0, cols ? 0 : 1, -1.0});
    }
    for (int pass = 0; pass < 2; ++pass) {
      launch_gemm(c, t1, s1, -1);
      launch_gemm(c, t2, s2, -1);
    }
  }
  const unsigned tri_chunks = reorth_chunks(rj, true);
  for (int pass = 0; pass < kReorthPasses; ++pass) {
    // G = W^T W;  T = I - F;  V = W T;  then V and W trade places
    std::vector<GemmTask> tg, tv;
    std::vector<GemmSeg> sg, sv;
    for (const ReorthJob& j : rj) {
      tg.push_back({j.G, j.k, j.k, j.k, (int)sg.size(), 1, 0.0});
      sg.push_back({j.W, j.W, j.n, j.n, j.n, 1, 1.0});
      tv.push_back({j.V, j.n, j.n, j.k, (int)sv.size(), 1, 0.0});
      sv.push_back({j.W, j.T, j.n, j.k, j.k, 0, 1.0});
    }
    launch_gemm(c, tg, sg, -1);
    hipLaunchKernelGGL(k_reorth_tri, dim3(tri_chunks, nj), dim3(256), 0, c.stream, dj);
    HIP_OK(hipGetLastError());
    launch_gemm(c, tv, sv, -1);
    for (ReorthJob& j : rj) std::swap(j.W, j.V);
    dj = (ReorthJob*)c.upload(rj.data(), rj.size() * sizeof(ReorthJob));
  }
  hipLaunchKernelGGL(k_reorth_move, dim3(reorth_chunks(rj, false), nj), dim3(256), 0, c.stream, dj, 1);
  HIP_OK(hipGetLastError());
  for (double* p : scratch) c.alloc.release(p);
}

size_t bj_lds_bytes(int rows, int bw) { return (size_t)2 * bw * rows * sizeof(double); }

int bj_panel_width(int rows, size_t lds_max, size_t static_lds) {
  if (static_lds >= lds_max) return 0;
  const size_t dyn = lds_max - static_lds;
  int bw = 16;
  while (bw > 4 && bj_lds_bytes(rows, bw) > dyn) bw >>= 1;
  return bj_lds_bytes(rows, bw) <= dyn ? bw : 0;
}

size_t bj_static_lds() {
  // the compiled kernel's own static LDS, read once from its attributes
  static size_t bytes = [] {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, (const void*)k_bj_step) != hipSuccess) {
      (void)hipGetLastError();
      return (size_t)4096;  // a safe over-estimate if the attribute query fails
    }
    return (size_t)a.sharedSizeBytes;
  }();
  return bytes;
}

void launch_bj_sweep(Context& c, const BjJob* djobs, const std::vector<int2>& work, int nsteps, int inner_max,
                     size_t lds) {
  if (work.empty() || nsteps <= 0) return;
  const size_t static_lds = bj_static_lds();
  if (lds + static_lds > c.lds_max) throw Error(-2, "block Jacobi panel exceeds the work-group LDS");
  static bool attr_set = false;
  if (!attr_set) {  // allow dynamic LDS up to the device limit less the kernel's static LDS
    (void)hipFuncSetAttribute((const void*)k_bj_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(c.lds_max - static_lds));
    (void)hipGetLastError();
    attr_set = true;
  }
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 2);
  for (int s = 0; s < nsteps; ++s) {
    hipLaunchKernelGGL(k_bj_step, dim3(grid), dim3(kBjThreads), lds, c.stream, djobs, dw, (int)work.size(), s,
                       inner_max);
    HIP_OK(hipGetLastError());
  }
}

void launch_bj_norms(Context& c, const BjJob* djobs, const std::vector<int2>& work) {
  if (work.empty()) return;
  auto* dw = (int2*)c.upload(work.data(), work.size() * sizeof(int2));
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  hipLaunchKernelGGL(k_bj_norms, dim3(grid), dim3(kBjThreads), 0, c.stream, djobs, dw, (int)work.size());
  HIP_OK(hipGetLastError());
}

void launch_bj_gather(Context& c, const std::vector<BjGather>& g) {
  if (g.empty()) return;
  auto* dg = (BjGather*)c.upload(g.data(), g.size() * sizeof(BjGather));
  const int grid = (int)std::min<size_t>(g.size(), (size_t)c.num_cu * 8);
  hipLaunchKernelGGL(k_bj_gather, dim3(grid), dim3(kBjThreads), 0, c.stream, dg, (int)g.size());
  HIP_OK(hipGetLastError());
}

}  // namespace dmrj
