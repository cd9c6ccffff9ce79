// Device kernels of libdmrjhip (gfx950 / CDNA4).
//
//  * k_gemm_f64<BM,BN>: persistent grouped GEMM over a flat tile list.  Every
//    task is C = beta*C + sum over K-segments of alpha*op(A)*op(B) with its own
//    pointers, leading dimensions and transpose flags, so one launch covers all
//    symmetry blocks of one contraction step.  FP64 MFMA
//    v_mfma_f64_16x16x4f64, 256 threads = 4 waves in a 2x2 arrangement, LDS
//    double buffer with register prefetch (one barrier per K step).
//  * k_mix: small dense "label" contraction with the MPO coefficients
//    (HBM-bound weighted sums of equally shaped blocks).
//  * k_copy: batched strided block copy with optional row/column scaling
//    (SVD sector gather and split).
//  * deterministic two-pass dot products and the scalar steps of the
//    two-vector Lanczos, all reading and writing device scalars so that a
//    Lanczos application needs no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.h"

namespace dmrj {

typedef double double4_t __attribute__((ext_vector_type(4)));

#define HIP_OK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kThreads = 256;
constexpr int kBK = 16;
// LDS tiles are stored [row][k] with a pitch of kBK+1 doubles (34 banks): the
// 16 lanes of an MFMA operand read (consecutive rows, same k) and the 16-lane
// groups of the stores (consecutive k of one row, or consecutive rows of one k)
// each touch 16 distinct bank pairs.
constexpr int kPitch = kBK + 1;

// The 32x32 shape is held to 64 registers per lane (eight waves per SIMD,
// eight work-groups per CU as its LDS allows) without scratch; the larger
// shapes keep the compiler's allocation.
template <int BM, int BN>
__global__ __launch_bounds__(kThreads, (BM * BN <= 1024 ? 8 : 1)) void k_gemm_f64(const GemmTile* __restrict__ tiles,
                                                       int ntiles,
                                                       const GemmTask* __restrict__ tasks,
                                                       const GemmSeg* __restrict__ segs) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int AE = BM * kBK / kThreads, BE = BN * kBK / kThreads;
  __shared__ double sA[2][BM * kPitch];
  __shared__ double sB[2][BN * kPitch];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  int buf = 0;

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GemmTile tl = tiles[t];
    const GemmTask tk = tasks[tl.task];
    const int m0 = tl.m0, n0 = tl.n0;
    double4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

    for (int sidx = 0; sidx < tk.nseg; ++sidx) {
      const GemmSeg sg = segs[tk.seg0 + sidx];
      const bool tA = sg.flags & 1, tB = sg.flags & 2;
      const int K = sg.K;
      double ra[AE], rb[BE];
      // Per-thread element offsets of the first K step (k0 = 0); a K step adds
      // k0 * (stride along k).  A row (column) outside the task reads the
      // task's first row (column) instead and is written to LDS as zero, so
      // a K step that lies wholly inside K loads without per-element checks.
      const int64_t ska = tA ? 1 : sg.lda, skb = tB ? sg.ldb : 1;
      int64_t oa[AE], ob[BE];
      int ka[AE], kb[BE];
      bool va[AE], vb[BE];
#pragma unroll
      for (int r = 0; r < AE; ++r) {
        const int idx = tid + r * kThreads;
        int m, k;
        if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kBK; m = idx / kBK; }
        va[r] = m0 + m < tk.M;
        const int gm = va[r] ? m0 + m : m0;
        ka[r] = k;
        oa[r] = tA ? (int64_t)k + (int64_t)gm * sg.lda : (int64_t)gm + (int64_t)k * sg.lda;
      }
#pragma unroll
      for (int r = 0; r < BE; ++r) {
        const int idx = tid + r * kThreads;
        int n, k;
        if (!tB) { k = idx % kBK; n = idx / kBK; } else { n = idx % BN; k = idx / BN; }
        vb[r] = n0 + n < tk.N;
        const int gn = vb[r] ? n0 + n : n0;
        kb[r] = k;
        ob[r] = tB ? (int64_t)gn + (int64_t)k * sg.ldb : (int64_t)k + (int64_t)gn * sg.ldb;
      }
      auto load = [&](int k0) {
        const double* pa = sg.A + (int64_t)k0 * ska;
        const double* pb = sg.B + (int64_t)k0 * skb;
        if (k0 + kBK <= K) {
#pragma unroll
          for (int r = 0; r < AE; ++r) ra[r] = ld_global(pa + oa[r]);
#pragma unroll
          for (int r = 0; r < BE; ++r) rb[r] = ld_global(pb + ob[r]);
        } else {
#pragma unroll
          for (int r = 0; r < AE; ++r) ra[r] = k0 + ka[r] < K ? ld_global(pa + oa[r]) : 0.0;
#pragma unroll
          for (int r = 0; r < BE; ++r) rb[r] = k0 + kb[r] < K ? ld_global(pb + ob[r]) : 0.0;
        }
      };
      // alpha is applied when the prefetched values are written to LDS, after
      // the products of the current step, so nothing waits on the loads early
      auto store = [&](int b) {
#pragma unroll
        for (int r = 0; r < AE; ++r) {
          const int idx = tid + r * kThreads;
          int m, k;
          if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kBK; m = idx / kBK; }
          sA[b][m * kPitch + k] = va[r] ? ra[r] * sg.alpha : 0.0;
        }
#pragma unroll
        for (int r = 0; r < BE; ++r) {
          const int idx = tid + r * kThreads;
          int n, k;
          if (!tB) { k = idx % kBK; n = idx / kBK; } else { n = idx % BN; k = idx / BN; }
          sB[b][n * kPitch + k] = vb[r] ? rb[r] : 0.0;
        }
      };
      if (K <= 0) continue;
      load(0);
      for (int k0 = 0; k0 < K; k0 += kBK) {
        store(buf);
        __syncthreads();
        if (k0 + kBK < K) load(k0 + kBK);
#pragma unroll
        for (int kk = 0; kk < kBK / 4; ++kk) {
          double a[FM], b[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) a[i] = sA[buf][(wm * WM + i * 16 + lr) * kPitch + kk * 4 + lk];
#pragma unroll
          for (int j = 0; j < FN; ++j) b[j] = sB[buf][(wn * WN + j * 16 + lr) * kPitch + kk * 4 + lk];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        buf ^= 1;
      }
    }
    // epilogue: accumulator item r of lane l holds row 4r + l/16, column l%16
    // of each 16x16 block (gfx950 v_mfma_f64_16x16x4f64 output layout)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = n0 + wn * WN + j * 16 + lr;
        if (col >= tk.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WM + i * 16 + r * 4 + lk;
          if (row < tk.M) {
            double* p = tk.C + (int64_t)row + (int64_t)col * tk.ldc;
            double v = acc[i][j][r];
            if (tk.beta != 0.0) v += tk.beta * *p;
            *p = v;
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
constexpr int kMixChunk = 2048;

__global__ __launch_bounds__(kThreads) void k_mix(const MixWork* __restrict__ work, int nwork,
                                                  const MixDst* __restrict__ dsts,
                                                  const MixTerm* __restrict__ terms) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const MixWork w = work[wi];
    const MixDst d = dsts[w.dst];
    const int64_t e1 = min(w.e0 + (int64_t)kMixChunk, d.n);
    for (int64_t e = w.e0 + threadIdx.x; e < e1; e += kThreads) {
      double acc = 0.0;
      for (int t = 0; t < d.nt; ++t) {
        const MixTerm tm = terms[d.t0 + t];
        acc += tm.coef * tm.src[e];
      }
      d.dst[e] = acc;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_copy(const CopyWork* __restrict__ work, int nwork,
                                                   const CopyOp* __restrict__ ops) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const CopyWork w = work[wi];
    const CopyOp op = ops[w.op];
    const int64_t n = (int64_t)op.rows * op.cols;
    const int64_t e1 = min(w.e0 + (int64_t)kMixChunk, n);
    for (int64_t e = w.e0 + threadIdx.x; e < e1; e += kThreads) {
      const int r = (int)(e % op.rows), c = (int)(e / op.rows);
      double v = 0.0;
      if (op.src) {
        v = op.alpha * op.src[(int64_t)r + (int64_t)c * op.lds];
        if (op.rscale) v *= op.rscale[r];
        if (op.cscale) v *= op.cscale[c];
      }
      op.dst[(int64_t)r + (int64_t)c * op.ldd] = v;
    }
  }
}

__global__ void k_fill(double* p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
constexpr int kDotBlocks = 256;

struct DotArgs {
  const double* x[4];
  const double* y[4];
};

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kThreads) void k_dot_partial(DotArgs args, int64_t n, double* partial) {
  const int pair = blockIdx.y;
  const double* __restrict__ x = args.x[pair];
  const double* __restrict__ y = args.y[pair];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    s += x[i] * y[i];
  __shared__ double red[kThreads / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) t += red[i];
    partial[pair * kDotBlocks + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kThreads) void k_dot_final(const double* partial, double* out) {
  const int pair = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += kThreads) s += partial[pair * kDotBlocks + i];
  __shared__ double red[kThreads / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) t += red[i];
    out[pair] = t;
  }
}

// Block classical Gram-Schmidt pass of the Krylov reorthogonalisation:
// coef_i = <v_i, w> for all i < k in one pair of launches (deterministic
// two-level reduction), then w -= sum_i coef_i v_i in one launch.  Real:
// vectors of n elements.  Complex (split planes, stride nv >= 2n): <v, w> =
// (vr.wr + vi.wi) + i (vr.wi - vi.wr) and w -= c v on both planes.
__global__ __launch_bounds__(kThreads) void k_vdots_partial(const double* __restrict__ V, int64_t nv,
                                                            const double* __restrict__ w, int64_t n, int cplx,
                                                            double* __restrict__ partial) {
  const double* __restrict__ v = V + (int64_t)blockIdx.y * nv;
  double re = 0.0, im = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads) {
    if (cplx) {
      const double vr = v[e], vi = v[n + e], wr = w[e], wi = w[n + e];
      re += vr * wr + vi * wi;
      im += vr * wi - vi * wr;
    } else {
      re += v[e] * w[e];
    }
  }
  __shared__ double red[2][kThreads / 64];
  re = wave_sum(re);
  im = wave_sum(im);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = re;
    red[1][threadIdx.x >> 6] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) {
      a += red[0][i];
      b += red[1][i];
    }
    const int64_t o = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
    partial[o] = a;
    partial[o + 1] = b;
  }
}

__global__ __launch_bounds__(kThreads) void k_vdots_final(const double* __restrict__ partial, int nblk,
                                                          double* __restrict__ coef) {
  const int i = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kThreads) {
    re += partial[((int64_t)i * nblk + b) * 2];
    im += partial[((int64_t)i * nblk + b) * 2 + 1];
  }
  __shared__ double red[2][kThreads / 64];
  re = wave_sum(re);
  im = wave_sum(im);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = re;
    red[1][threadIdx.x >> 6] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int t = 0; t < kThreads / 64; ++t) {
      a += red[0][t];
      b += red[1][t];
    }
    coef[2 * i] = a;
    coef[2 * i + 1] = b;
  }
}

__global__ void k_cgs_sub(double* __restrict__ w, const double* __restrict__ V, int64_t nv, int64_t n, int k,
                          int cplx, const double* __restrict__ coef) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < k; ++i) {
      const double* v = V + (int64_t)i * nv;
      const double cr = coef[2 * i];
      if (cplx) {
        const double ci = coef[2 * i + 1], vr = v[e], vi = v[n + e];
        ar += cr * vr - ci * vi;
        ai += cr * vi + ci * vr;
      } else {
        ar += cr * v[e];
      }
    }
    w[e] -= ar;
    if (cplx) w[n + e] -= ai;
  }
}

// One Krylov iteration after the H_eff application, in one work-group, for
// short vectors (the latency of a dozen dependent launches dominates there):
// alpha_j = Re <v_j, w> -> sc[j]; unless last: two classical Gram-Schmidt
// passes of w against v_0..v_j (coefficients of a pass from the same w, then
// one subtraction, as the multi-launch passes), beta_j = |w| -> sc[64 + j],
// v_{j+1} = w / beta_j (zero when beta_j <= 1e-10) written to V and to the
// fixed H_eff input xin.  Complex vectors: real plane then imaginary plane of
// n elements each, <v, w> = sum conj(v) w.
__global__ __launch_bounds__(kThreads) void k_krylov_tail(double* __restrict__ V, int64_t nv, int j,
                                                          double* __restrict__ w, int64_t n, int cplx,
                                                          double* __restrict__ sc, double* __restrict__ xin,
                                                          int last) {
  __shared__ double red[2][kThreads / 64];
  __shared__ double coef[2 * 64];
  const int tid = threadIdx.x;
  const int64_t ne = cplx ? 2 * n : n;
  auto block_sum2 = [&](double a, double b, double& A, double& B) {
    a = wave_sum(a);
    b = wave_sum(b);
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = a;
      red[1][tid >> 6] = b;
    }
    __syncthreads();
    A = 0.0;
    B = 0.0;
    for (int i = 0; i < kThreads / 64; ++i) {
      A += red[0][i];
      B += red[1][i];
    }
    __syncthreads();
  };
  double A, B;
  {
    const double* vj = V + (int64_t)j * nv;
    double a = 0.0;
    for (int64_t e = tid; e < ne; e += kThreads) a += vj[e] * w[e];
    block_sum2(a, 0.0, A, B);
    if (tid == 0) sc[j] = A;
  }
  if (last) return;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i <= j; ++i) {
      const double* v = V + (int64_t)i * nv;
      double re = 0.0, im = 0.0;
      for (int64_t e = tid; e < n; e += kThreads) {
        if (cplx) {
          const double vr = v[e], vi = v[n + e], wr = w[e], wi = w[n + e];
          re += vr * wr + vi * wi;
          im += vr * wi - vi * wr;
        } else {
          re += v[e] * w[e];
        }
      }
      block_sum2(re, im, A, B);
      if (tid == 0) {
        coef[2 * i] = A;
        coef[2 * i + 1] = B;
      }
    }
    __syncthreads();
    for (int64_t e = tid; e < n; e += kThreads) {
      double ar = 0.0, ai = 0.0;
      for (int i = 0; i <= j; ++i) {
        const double* v = V + (int64_t)i * nv;
        const double cr = coef[2 * i];
        if (cplx) {
          const double ci = coef[2 * i + 1], vr = v[e], vi = v[n + e];
          ar += cr * vr - ci * vi;
          ai += cr * vi + ci * vr;
        } else {
          ar += cr * v[e];
        }
      }
      w[e] -= ar;
      if (cplx) w[n + e] -= ai;
    }
    __syncthreads();
  }
  double s2 = 0.0;
  for (int64_t e = tid; e < ne; e += kThreads) s2 += w[e] * w[e];
  block_sum2(s2, 0.0, A, B);
  const double b = sqrt(A);
  const double invb = b > 1e-10 ? 1.0 / b : 0.0;
  if (tid == 0) sc[64 + j] = b;
  double* vn = V + (int64_t)(j + 1) * nv;
  for (int64_t e = tid; e < ne; e += kThreads) {
    const double x = w[e] * invb;
    vn[e] = x;
    xin[e] = x;
  }
}

void launch_krylov_tail(Context& c, double* V, int64_t nv, int j, double* w, int64_t n, int cplx, double* sc,
                        double* xin, int last) {
  if (j < 0 || j >= 64) throw Error(-1, "krylov tail: vector index out of range");
  hipLaunchKernelGGL(k_krylov_tail, dim3(1), dim3(kThreads), 0, c.stream, V, nv, j, w, n, cplx, sc, xin, last);
  HIP_OK(hipGetLastError());
}

__global__ void k_axpy_dev(double* y, const double* x, int64_t n, const double* coef) {
  const double c = *coef;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += c * x[i];
}

__global__ void k_scale_dev(double* x, int64_t n, const double* coef) {
  const double c = *coef;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= c;
}

__global__ void k_comb_dev(double* out, const double* a, const double* b, int64_t n,
                           const double* ca, const double* cb) {
  const double u = *ca, v = *cb;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = u * a[i] + v * b[i];
}

// Scalar steps of algs/DMRG.jl:57-90 on device scalars.
__global__ void k_lz_scalar(double* s, int stage) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (stage == 0) {            // normalise the orthogonality-centre tensor
    s[S_INVOC] = 1.0 / sqrt(s[S_OC]);
  } else if (stage == 1) {     // alpha1 = <AA|Hpsi>, AA = theta/|theta|
    const double nrm = sqrt(s[S_TT]);
    s[S_NRM] = nrm;
    s[S_A1] = s[S_TH] / nrm;
    s[S_CAXPY] = -s[S_A1] / nrm;   // psi2 = Hpsi - alpha1*theta/|theta|
  } else if (stage == 2) {     // beta1 = |psi2|
    const double b = sqrt(s[S_PP]);
    s[S_B1] = b;
    const bool go = b > 1e-10;
    s[S_FLAG] = go ? 1.0 : 0.0;
    s[S_INVB] = go ? 1.0 / b : 0.0;
  } else if (stage == 3) {     // 2x2 Rayleigh-Ritz, lowest eigenpair
    const double a1 = s[S_A1];
    if (s[S_FLAG] == 0.0) {
      s[S_E] = a1;
      s[S_CA] = 1.0 / s[S_NRM];
      s[S_CB] = 0.0;
      return;
    }
    const double b = s[S_B1], a2 = s[S_PHP];
    s[S_A2] = a2;
    const double dd = 0.5 * (a2 - a1);
    const double r = sqrt(dd * dd + b * b);
    const double lam = 0.5 * (a1 + a2) - r;
    double v1, v2;
    if (dd >= 0.0) { v1 = -(dd + r); v2 = b; } else { v1 = b; v2 = -(r - dd); }
    const double nv = sqrt(v1 * v1 + v2 * v2);
    v1 /= nv;
    v2 /= nv;
    s[S_E] = lam;
    s[S_CA] = v1 / s[S_NRM];
    s[S_CB] = v2;
  } else if (stage == 4) {     // FP64 Rayleigh quotient of the Ritz vector
    s[S_ELP] = s[S_E];
    const double e = s[S_RH] / s[S_RR];
    s[S_E] = e;
    s[S_NEGE] = -e;
  }
}

// Lane-layout probes of v_mfma_f64_16x16x4f64 (one wave, in-bounds):
//  mode 0: a = lane+1, b = 1      -> each output is a row sum of A
//  mode 1: a = 1, b = lane+1      -> each output is a column sum of B
//  mode 2: A[:,0] = i+1, A[:,1] = 1, B[0,:] = 1, B[1,:] = 1000(j+1) under the
//          layout A[i][k] <- lane i+16k, B[k][j] <- lane j+16k, giving
//          D[i][j] = (i+1) + 1000(j+1), found in item r of lane l for
//          i = 4r + l/16, j = l%16
__global__ void k_mfma_probe(double* out) {
  const int l = threadIdx.x;
  const int r16 = l & 15, k = l >> 4;
  for (int mode = 0; mode < 3; ++mode) {
    double a, b;
    if (mode == 0) {
      a = (double)(l + 1);
      b = 1.0;
    } else if (mode == 1) {
      a = 1.0;
      b = (double)(l + 1);
    } else {
      a = k == 0 ? (double)(r16 + 1) : (k == 1 ? 1.0 : 0.0);
      b = k == 0 ? 1.0 : (k == 1 ? 1000.0 * (r16 + 1) : 0.0);
    }
    double4_t acc = double4_t{0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[mode * 256 + l * 4 + r] = acc[r];
  }
}

void launch_mfma_probe(Context& c, double* out) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, c.stream, out);
  HIP_OK(hipGetLastError());
}

// out[i] = sum_j c[j] * V[j*ld + i] (Krylov vector combination)
__global__ void k_combine(double* out, const double* V, int64_t ld, int64_t n, const double* c, int k) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += c[j] * V[(int64_t)j * ld + i];
    out[i] = acc;
  }
}

// Coefficients of the Krylov exponential / Ritz vector on the device (one
// wave; the Krylov dimension is at most 64, typically 2..12):
//   k = the Krylov dimension used (first beta_j <= 1e-10 ends the space),
//   T = Q diag(lam) Q^T by parallel-order Jacobi (FP64, the tridiagonal of
//   alpha in sc[0..], beta in sc[64..]): each step rotates k/2 disjoint index
//   pairs of a round-robin schedule at once (columns, then rows), k-1 steps
//   per sweep, until the off-diagonal part is at the matrix's rounding level,
//   c = Q exp(pf lam) Q^T e1 with the complex prefactor pf = (pf_re, pf_im),
//   or (ritz) the eigenvector of the lowest lam signed to c_1 >= 0.
// coef: real Krylov vectors: c_re[0..r), c_im[r..2r); complex Krylov vectors
// (split planes, basis planes 2j = Re v_j, 2j+1 = Im v_j): the plane weights
// of Re out in [0..2r) and of Im out in [2r..4r).  Entries j >= k are zero.
// outs: {lowest Ritz value, k, 1 if an alpha/beta was non-finite}.
// Dynamic LDS: 2 r^2 doubles (the matrix and its eigenvectors).
__device__ inline void kc_pair(int n, int s, int i, int& p, int& q) {  // round robin of n (even) players
  if (i == 0) {
    p = s;
    q = n - 1;
  } else {
    p = (s + i) % (n - 1);
    q = (s - i + (n - 1)) % (n - 1);
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

__global__ __launch_bounds__(64) void k_krylov_coef(const double* __restrict__ sc, int r, double pf_re,
                                                     double pf_im, int ritz, int zv, double* __restrict__ coef,
                                                     double* __restrict__ outs) {
  extern __shared__ double kc_lds[];
  __shared__ int sk, sbad, simin;
  __shared__ double scs[32], ssn[32];
  __shared__ int sp[32], sq[32];
  const int t = threadIdx.x;
  if (t == 0) {
    int k = r, bad = 0;
    for (int j = 0; j + 1 < r; ++j) {
      if (!isfinite(sc[64 + j]) || !isfinite(sc[j])) bad = 1;
      if (sc[64 + j] <= 1e-10) {
        k = j + 1;
        break;
      }
    }
    sk = k;
    sbad = bad;
  }
  __syncthreads();
  const int k = sk, bad = sbad;
  double* A = kc_lds;
  double* Q = kc_lds + (int64_t)k * k;
  for (int e = t; e < k * k; e += 64) {
    A[e] = 0.0;
    Q[e] = 0.0;
  }
  __syncthreads();
  for (int i = t; i < k; i += 64) {
    A[i + i * k] = sc[i];
    if (i + 1 < k) A[i + (i + 1) * k] = A[(i + 1) + i * k] = sc[64 + i];
    Q[i + i * k] = 1.0;
  }
  __syncthreads();
  const int K = (k + 1) & ~1;  // even number of players; player k of an odd k is idle
  const int npair = K / 2;
  for (int sweep = 0; sweep < 100 && !bad && k > 1; ++sweep) {
    double off = 0.0, dia = 0.0;  // off: both triangles
    for (int e = t; e < k * k; e += 64) {
      const double v = A[e];
      if (e % k == e / k)
        dia += v * v;
      else
        off += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      dia += __shfl_xor(dia, o, 64);
    }
    // lane 0's sums decide for all lanes (the butterfly may round differently
    // per lane), so the loop and its barriers stay uniform
    off = __shfl(off, 0, 64);
    dia = __shfl(dia, 0, 64);
    if (off <= 2e-30 * (dia + off) || off < 1e-300) break;
    for (int st = 0; st < K - 1; ++st) {
      if (t < npair) {
        int p, q;
        kc_pair(K, st, t, p, q);
        double cs = 1.0, sn = 0.0;
        if (q < k) {
          const double apq = A[p + q * k];
          if (apq != 0.0) {
            const double app = A[p + p * k], aqq = A[q + q * k];
            const double theta = (aqq - app) / (2.0 * apq);
            const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            cs = 1.0 / sqrt(tt * tt + 1.0);
            sn = tt * cs;
          }
        }
        sp[t] = p;
        sq[t] = q;
        scs[t] = cs;
        ssn[t] = sn;
      }
      __syncthreads();
      for (int e = t; e < npair * k; e += 64) {  // columns p, q of A and of Q
        const int pi = e / k, rr = e % k;
        const int p = sp[pi], q = sq[pi];
        if (q >= k) continue;
        const double cs = scs[pi], sn = ssn[pi];
        double x = A[rr + p * k], y = A[rr + q * k];
        A[rr + p * k] = cs * x - sn * y;
        A[rr + q * k] = sn * x + cs * y;
        x = Q[rr + p * k];
        y = Q[rr + q * k];
        Q[rr + p * k] = cs * x - sn * y;
        Q[rr + q * k] = sn * x + cs * y;
      }
      __syncthreads();
      for (int e = t; e < npair * k; e += 64) {  // rows p, q of A
        const int pi = e / k, rr = e % k;
        const int p = sp[pi], q = sq[pi];
        if (q >= k) continue;
        const double cs = scs[pi], sn = ssn[pi];
        const double x = A[p + rr * k], y = A[q + rr * k];
        A[p + rr * k] = cs * x - sn * y;
        A[q + rr * k] = sn * x + cs * y;
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    double emin = A[0];
    int imin = 0;
    for (int i = 1; i < k; ++i)
      if (A[i + i * k] < emin) {
        emin = A[i + i * k];
        imin = i;
      }
    simin = imin;
    outs[0] = emin;
    outs[1] = (double)k;
    outs[2] = bad ? 1.0 : 0.0;
  }
  __syncthreads();
  const int imin = simin;
  const int nc = zv ? 4 * r : 2 * r;
  for (int e = t; e < nc; e += 64) coef[e] = 0.0;
  __syncthreads();
  for (int jj = t; jj < k; jj += 64) {
    double cre = 0.0, cim = 0.0;
    if (ritz) {
      const double sg = Q[0 + imin * k] >= 0.0 ? 1.0 : -1.0;
      cre = sg * Q[jj + imin * k];
    } else {
      for (int i = 0; i < k; ++i) {
        const double lam = A[i + i * k];
        const double mag = exp(pf_re * lam) * Q[0 + i * k];
        cre += Q[jj + i * k] * (mag * cos(pf_im * lam));
        cim += Q[jj + i * k] * (mag * sin(pf_im * lam));
      }
    }
    if (!zv) {
      coef[jj] = cre;
      coef[r + jj] = cim;
    } else {
      coef[2 * jj] = cre;           // Re out += Re c Re v - Im c Im v
      coef[2 * jj + 1] = -cim;
      coef[2 * r + 2 * jj] = cim;   // Im out += Im c Re v + Re c Im v
      coef[2 * r + 2 * jj + 1] = cre;
    }
  }
}

// ---------------------------------------------------------------------------
static int grid_for(int64_t n, int per = kThreads * 4) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

int choose_gemm_cfg(const std::vector<GemmTask>& tasks, int num_cu) {
  // Heuristic used when the autotuner is off: the largest square tile whose
  // tile count still fills the chip ~2x; small blocks fall back to smaller
  // tiles.
  for (int cfg = 2; cfg >= 1; --cfg) {
    int64_t tiles = 0, waste = 0, useful = 0;
    for (const auto& t : tasks) {
      const int bm = kGemmShapes[cfg].bm, bn = kGemmShapes[cfg].bn;
      const int64_t tm = (t.M + bm - 1) / bm, tn = (t.N + bn - 1) / bn;
      tiles += tm * tn;
      useful += (int64_t)t.M * t.N;
      waste += tm * bm * tn * bn;
    }
    if (tiles >= 2 * num_cu && useful * 10 >= waste * 7) return cfg;
  }
  return 1;
}

double gemm_flops(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs) {
  double f = 0.0;
  for (const auto& t : tasks)
    for (int s = 0; s < t.nseg; ++s) f += 2.0 * t.M * (double)t.N * segs[t.seg0 + s].K;
  return f;
}

double gemm_bytes(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs) {
  double bytes = 0.0;  // compulsory: every operand block read once, C written once
  for (const auto& t : tasks) {
    bytes += 8.0 * t.M * (double)t.N;
    for (int q = 0; q < t.nseg; ++q) bytes += 8.0 * segs[t.seg0 + q].K * ((double)t.M + t.N);
  }
  return bytes;
}

template <int BM, int BN>
static void gemm_f64_kernel_launch(Context& c, const GemmLaunch& g) {
  if (c.ext_a)
    hipExtLaunchKernelGGL((k_gemm_f64<BM, BN>), dim3(g.grid), dim3(kThreads), 0, c.stream, c.ext_a, c.ext_b, 0,
                          g.tiles, g.ntiles, g.tasks, g.segs);
  else
    hipLaunchKernelGGL((k_gemm_f64<BM, BN>), dim3(g.grid), dim3(kThreads), 0, c.stream, g.tiles, g.ntiles,
                       g.tasks, g.segs);
}

static const void* gemm_f64_kernel_ptr(int cfg) {
  switch (cfg) {
    case 0: return (const void*)k_gemm_f64<32, 32>;
    case 1: return (const void*)k_gemm_f64<64, 64>;
    case 2: return (const void*)k_gemm_f64<128, 128>;
    case 3: return (const void*)k_gemm_f64<64, 128>;
    default: return (const void*)k_gemm_f64<128, 64>;
  }
}

int gemm_occupancy_f64(int cfg) {
  int nb = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_f64_kernel_ptr(cfg), kThreads, 0));
  return std::max(nb, 1);
}

void issue_gemm_f64(Context& c, const GemmLaunch& g) {
  if (g.ntiles <= 0) return;
  switch (g.cfg) {
    case 0: gemm_f64_kernel_launch<32, 32>(c, g); break;
    case 1: gemm_f64_kernel_launch<64, 64>(c, g); break;
    case 2: gemm_f64_kernel_launch<128, 128>(c, g); break;
    case 3: gemm_f64_kernel_launch<64, 128>(c, g); break;
    default: gemm_f64_kernel_launch<128, 64>(c, g); break;
  }
  HIP_OK(hipGetLastError());
}

void launch_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs,
                 int cfg) {
  launch_gemm_arith(c, tasks, segs, cfg, 0);
}

MixLaunch prepare_mix(Context& c, const std::vector<MixDst>& dsts, const std::vector<MixTerm>& terms) {
  MixLaunch m;
  if (dsts.empty()) return m;
  std::vector<MixWork> work;
  for (size_t i = 0; i < dsts.size(); ++i)
    for (int64_t e = 0; e < dsts[i].n; e += kMixChunk) work.push_back({(int)i, 0, e});
  if (work.empty()) return m;
  UploadPack up;
  const size_t ow = up.add(work.data(), work.size() * sizeof(MixWork));
  const size_t od = up.add(dsts.data(), dsts.size() * sizeof(MixDst));
  const size_t ot = up.add(terms.data(), terms.size() * sizeof(MixTerm));
  char* base = up.commit(c);
  m.work = (const MixWork*)(base + ow);
  m.dsts = (const MixDst*)(base + od);
  m.terms = (const MixTerm*)(base + ot);
  m.nwork = (int)work.size();
  m.grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  for (const auto& d : dsts) {
    m.bytes += 8.0 * d.n * (1.0 + d.nt);
    m.flops += 2.0 * d.n * d.nt;
  }
  return m;
}

void issue_mix(Context& c, const MixLaunch& m) {
  if (m.nwork <= 0) return;
  hipLaunchKernelGGL(k_mix, dim3(m.grid), dim3(kThreads), 0, c.stream, m.work, m.nwork, m.dsts, m.terms);
  HIP_OK(hipGetLastError());
}

// Region marker for kernel traces: a one-wave kernel that does nothing; its
// dispatches bracket a region of interest in a profiler's dispatch list.
__global__ void k_region_marker(int id, double* sink) {
  if (id < 0 && threadIdx.x == 0) sink[0] = (double)id;
}

// The marker's id is also its work-group count, so kernel traces (which
// record grid sizes, not arguments) tell the markers apart.
void launch_marker(Context& c, int id) {
  hipLaunchKernelGGL(k_region_marker, dim3(std::max(1, id)), dim3(64), 0, c.stream, id, c.partial);
  HIP_OK(hipGetLastError());
}

void launch_mix(Context& c, const std::vector<MixDst>& dsts, const std::vector<MixTerm>& terms) {
  MixLaunch m = prepare_mix(c, dsts, terms);
  if (m.nwork <= 0) return;
  const int ps = c.prof_begin();
  issue_mix(c, m);
  if (ps >= 0) c.prof_end(ps, KIND_MIX, m.flops, m.bytes);
}

void launch_copy(Context& c, const std::vector<CopyOp>& ops) {
  if (ops.empty()) return;
  std::vector<CopyWork> work;
  for (size_t i = 0; i < ops.size(); ++i) {
    const int64_t n = (int64_t)ops[i].rows * ops[i].cols;
    for (int64_t e = 0; e < n; e += kMixChunk) work.push_back({(int)i, 0, e});
  }
  if (work.empty()) return;
  UploadPack up;
  const size_t ow = up.add(work.data(), work.size() * sizeof(CopyWork));
  const size_t oo = up.add(ops.data(), ops.size() * sizeof(CopyOp));
  char* base = up.commit(c);
  auto* dw = (CopyWork*)(base + ow);
  auto* dops = (CopyOp*)(base + oo);
  const int grid = (int)std::min<size_t>(work.size(), (size_t)c.num_cu * 8);
  const int ps = c.prof_begin();
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kThreads), 0, c.stream, dw, (int)work.size(), dops);
  HIP_OK(hipGetLastError());
  if (ps >= 0) {
    double bytes = 0.0;
    for (const auto& o : ops) bytes += 8.0 * o.rows * (double)o.cols * (o.src ? 2.0 : 1.0);
    c.prof_end(ps, KIND_COPY, 0.0, bytes);
  }
}

void launch_fill(Context& c, double* p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, p, n, v);
  HIP_OK(hipGetLastError());
}

static_assert(kDotBlocks == kDotPartials, "partials per pair");

void launch_dots_buf(Context& c, int npairs, const double* const* xs, const double* const* ys, int64_t n,
                     double* out, double* partial) {
  if (npairs < 1 || npairs > 4) throw Error(-2, "launch_dots: 1..4 pairs");
  DotArgs a{};
  for (int i = 0; i < npairs; ++i) {
    a.x[i] = xs[i];
    a.y[i] = ys[i];
  }
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks, npairs), dim3(kThreads), 0, c.stream, a, n, partial);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(npairs), dim3(kThreads), 0, c.stream, partial, out);
  HIP_OK(hipGetLastError());
}

void launch_dots(Context& c, int npairs, const double* const* xs, const double* const* ys,
                 int64_t n, double* out) {
  launch_dots_buf(c, npairs, xs, ys, n, out, c.partial);
}

void launch_axpy_dev(Context& c, double* y, const double* x, int64_t n, const double* coef) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy_dev, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, y, x, n, coef);
  HIP_OK(hipGetLastError());
}

void launch_scale_dev(Context& c, double* x, int64_t n, const double* coef) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_dev, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, x, n, coef);
  HIP_OK(hipGetLastError());
}

void launch_comb_dev(Context& c, double* out, const double* a, const double* b, int64_t n,
                     const double* ca, const double* cb) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_comb_dev, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, out, a, b, n, ca,
                     cb);
  HIP_OK(hipGetLastError());
}

void launch_cgs_pass(Context& c, double* w, const double* V, int64_t nv, int64_t n, int k, int cplx,
                     double* partial, double* coef) {
  if (n <= 0 || k <= 0) return;
  const int nblk = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + kThreads - 1) / kThreads));
  hipLaunchKernelGGL(k_vdots_partial, dim3(nblk, k), dim3(kThreads), 0, c.stream, V, nv, w, n, cplx, partial);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_vdots_final, dim3(k), dim3(kThreads), 0, c.stream, partial, nblk, coef);
  HIP_OK(hipGetLastError());
  hipLaunchKernelGGL(k_cgs_sub, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, w, V, nv, n, k, cplx, coef);
  HIP_OK(hipGetLastError());
}

void launch_krylov_coef(Context& c, const double* sc, int r, double pf_re, double pf_im, int ritz, int zv,
                        double* coef, double* outs) {
  if (r < 1 || r > 64) throw Error(-1, "launch_krylov_coef: Krylov dimension must be 1..64");
  const size_t lds = (size_t)2 * r * r * sizeof(double);
  hipLaunchKernelGGL(k_krylov_coef, dim3(1), dim3(64), lds, c.stream, sc, r, pf_re, pf_im, ritz, zv, coef, outs);
  HIP_OK(hipGetLastError());
}

void launch_combine(Context& c, double* out, const double* V, int64_t ld, int64_t n, const double* coef, int k) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(kThreads), 0, c.stream, out, V, ld, n, coef, k);
  HIP_OK(hipGetLastError());
}

void launch_lz_scalar(Context& c, int stage) {
  hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(64), 0, c.stream, c.scal, stage);
  HIP_OK(hipGetLastError());
}

}  // namespace dmrj
