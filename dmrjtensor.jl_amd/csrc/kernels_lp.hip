// Split-precision grouped GEMM on the BF16 matrix cores of gfx950.
//
// Same task / K-segment / tile structure as the FP64 kernel; operands stay
// FP64 in HBM.  Each element is split on load into a high and a low BF16 part
// (x = hi + lo + O(2^-16 |x|)) and the product is formed as
//     hi*hi + hi*lo + lo*hi
// with v_mfma_f32_16x16x32_bf16 (FP32 accumulation), i.e. about 16 correct
// bits per product.  A 32-deep K step of the tile is one MFMA per product.
// LDS holds hi and lo of both operands, row-major in k with a 40-element
// pitch (the 16 lanes of a 16-byte ds_read hit disjoint banks).
//
// The FP16 form (k_gemm_f16x3) does the same with FP16 halves on
// v_mfma_f32_16x16x32_f16 (11-bit significands: about 22 correct bits per
// product instead of 16) and per-block scaling: every K-segment's operands
// are scaled by powers of two that bring max |alpha A| and max |B| of that
// segment to [2^13, 2^14) before the split (k_seg_absmax measures them right
// before the GEMM), the segment's FP32 partial sum is unscaled exactly by
// 2^-(eA+eB) and added to the tile's FP32 total.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.h"

namespace dmrj {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kLpThreads = 256;
constexpr int kLpBK = 32;
constexpr int kLpPitch = 40;

__device__ inline void split_bf16(double x, __bf16& hi, __bf16& lo) {
  const float xf = (float)x;
  hi = (__bf16)xf;
  const double r = x - (double)(float)hi;
  lo = (__bf16)(float)r;
}

template <int BM, int BN>
__global__ __launch_bounds__(kLpThreads) void k_gemm_bf16x3(const GemmTile* __restrict__ tiles, int ntiles,
                                                             const GemmTask* __restrict__ tasks,
                                                             const GemmSeg* __restrict__ segs) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int AE = BM * kLpBK / kLpThreads, BE = BN * kLpBK / kLpThreads;
  __shared__ __attribute__((aligned(16))) __bf16 sAh[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[2][BN * kLpPitch];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[2][BN * kLpPitch];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  int buf = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GemmTile tl = tiles[t];
    const GemmTask tk = tasks[tl.task];
    const int m0 = tl.m0, n0 = tl.n0;
    f32x4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int sidx = 0; sidx < tk.nseg; ++sidx) {
      const GemmSeg sg = segs[tk.seg0 + sidx];
      const bool tA = sg.flags & 1, tB = sg.flags & 2;
      const int K = sg.K;
      if (K <= 0) continue;
      double ra[AE], rb[BE];
      auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < AE; ++r) {
          const int idx = tid + r * kLpThreads;
          int m, k;
          if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kLpBK; m = idx / kLpBK; }
          const int gm = m0 + m, gk = k0 + k;
          double v = 0.0;
          if (gm < tk.M && gk < K)
            v = ld_global(tA ? sg.A + ((int64_t)gk + (int64_t)gm * sg.lda) : sg.A + ((int64_t)gm + (int64_t)gk * sg.lda));
          ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < BE; ++r) {
          const int idx = tid + r * kLpThreads;
          int n, k;
          if (!tB) { k = idx % kLpBK; n = idx / kLpBK; } else { n = idx % BN; k = idx / BN; }
          const int gn = n0 + n, gk = k0 + k;
          double v = 0.0;
          if (gn < tk.N && gk < K)
            v = ld_global(tB ? sg.B + ((int64_t)gn + (int64_t)gk * sg.ldb) : sg.B + ((int64_t)gk + (int64_t)gn * sg.ldb));
          rb[r] = v;
        }
      };
      auto store = [&](int b) {
#pragma unroll
        for (int r = 0; r < AE; ++r) {
          const int idx = tid + r * kLpThreads;
          int m, k;
          if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kLpBK; m = idx / kLpBK; }
          __bf16 h, l;
          split_bf16(ra[r] * sg.alpha, h, l);  // scaled here, not at the load: the loads stay in flight
          sAh[b][m * kLpPitch + k] = h;
          sAl[b][m * kLpPitch + k] = l;
        }
#pragma unroll
        for (int r = 0; r < BE; ++r) {
          const int idx = tid + r * kLpThreads;
          int n, k;
          if (!tB) { k = idx % kLpBK; n = idx / kLpBK; } else { n = idx % BN; k = idx / BN; }
          __bf16 h, l;
          split_bf16(rb[r], h, l);
          sBh[b][n * kLpPitch + k] = h;
          sBl[b][n * kLpPitch + k] = l;
        }
      };
      load(0);
      for (int k0 = 0; k0 < K; k0 += kLpBK) {
        store(buf);
        __syncthreads();
        if (k0 + kLpBK < K) load(k0 + kLpBK);
        bf16x8_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int row = wm * WM + i * 16 + lr;
          ah[i] = *reinterpret_cast<const bf16x8_t*>(&sAh[buf][row * kLpPitch + lk * 8]);
          al[i] = *reinterpret_cast<const bf16x8_t*>(&sAl[buf][row * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int col = wn * WN + j * 16 + lr;
          bh[j] = *reinterpret_cast<const bf16x8_t*>(&sBh[buf][col * kLpPitch + lk * 8]);
          bl[j] = *reinterpret_cast<const bf16x8_t*>(&sBl[buf][col * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
        buf ^= 1;
      }
    }
    // accumulator item r of lane l: row 4*(l/16) + r, column l%16
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = n0 + wn * WN + j * 16 + lr;
        if (col >= tk.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WM + i * 16 + lk * 4 + r;
          if (row < tk.M) {
            double* p = tk.C + (int64_t)row + (int64_t)col * tk.ldc;
            double v = (double)acc[i][j][r];
            if (tk.beta != 0.0) v += tk.beta * *p;
            *p = v;
          }
        }
      }
  }
}

// ------------------------------------------------------------------ FP16x3
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

__device__ inline void split_f16(double x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)(float)x;
  lo = (_Float16)(float)(x - (double)hi);
}

// 2^e max in [2^13, 2^14): FP16 (max 65504) holds the scaled hi part
__device__ inline int f16_scale_exp(double mx) {
  if (!(mx > 0.0) || !isfinite(mx)) return 0;
  int e;
  (void)frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1)
  return 14 - e;
}

// work item (segment, operand 0 = A / 1 = B, first column, end column) of the
// operand as stored (column-major with its leading dimension)
__global__ __launch_bounds__(256) void k_seg_absmax(const GemmTask* __restrict__ tasks, const GemmSeg* __restrict__ segs,
                                                    const int* __restrict__ seg_task, const int4* __restrict__ work,
                                                    int nwork, double* __restrict__ segmax) {
  __shared__ double red[4];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int4 w = work[wi];
    const GemmSeg sg = segs[w.x];
    const GemmTask tk = tasks[seg_task[w.x]];
    const bool tA = sg.flags & 1, tB = sg.flags & 2;
    const double* base = w.y == 0 ? sg.A : sg.B;
    const int64_t ld = w.y == 0 ? sg.lda : sg.ldb;
    const int rows = w.y == 0 ? (tA ? sg.K : tk.M) : (tB ? tk.N : sg.K);
    const int ncol = w.w - w.z;
    double m = 0.0;
    for (int64_t e = threadIdx.x; e < (int64_t)rows * ncol; e += 256) {
      const int r = (int)(e % rows), cc = w.z + (int)(e / rows);
      m = fmax(m, fabs(base[(int64_t)r + (int64_t)cc * ld]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      if (w.y == 0) v *= fabs(sg.alpha);
      // non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<unsigned long long*>(segmax + 2 * w.x + w.y), (unsigned long long)__double_as_longlong(v));
    }
    __syncthreads();
  }
}

template <int BM, int BN>
__global__ __launch_bounds__(kLpThreads) void k_gemm_f16x3(const GemmTile* __restrict__ tiles, int ntiles,
                                                            const GemmTask* __restrict__ tasks,
                                                            const GemmSeg* __restrict__ segs,
                                                            const double* __restrict__ segmax) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int AE = BM * kLpBK / kLpThreads, BE = BN * kLpBK / kLpThreads;
  __shared__ __attribute__((aligned(16))) _Float16 sAh[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sAl[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sBh[2][BN * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sBl[2][BN * kLpPitch];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  int buf = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GemmTile tl = tiles[t];
    const GemmTask tk = tasks[tl.task];
    const int m0 = tl.m0, n0 = tl.n0;
    f32x4_t tot[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) tot[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int sidx = 0; sidx < tk.nseg; ++sidx) {
      const int si = tk.seg0 + sidx;
      const GemmSeg sg = segs[si];
      const bool tA = sg.flags & 1, tB = sg.flags & 2;
      const int K = sg.K;
      if (K <= 0) continue;
      const int ea = f16_scale_exp(segmax[2 * si]), eb = f16_scale_exp(segmax[2 * si + 1]);
      const double fa = ldexp(sg.alpha, ea), fb = ldexp(1.0, eb);
      f32x4_t acc[FM][FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      double ra[AE], rb[BE];
      auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < AE; ++r) {
          const int idx = tid + r * kLpThreads;
          int m, k;
          if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kLpBK; m = idx / kLpBK; }
          const int gm = m0 + m, gk = k0 + k;
          double v = 0.0;
          if (gm < tk.M && gk < K)
            v = ld_global(tA ? sg.A + ((int64_t)gk + (int64_t)gm * sg.lda) : sg.A + ((int64_t)gm + (int64_t)gk * sg.lda));
          ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < BE; ++r) {
          const int idx = tid + r * kLpThreads;
          int n, k;
          if (!tB) { k = idx % kLpBK; n = idx / kLpBK; } else { n = idx % BN; k = idx / BN; }
          const int gn = n0 + n, gk = k0 + k;
          double v = 0.0;
          if (gn < tk.N && gk < K)
            v = ld_global(tB ? sg.B + ((int64_t)gn + (int64_t)gk * sg.ldb) : sg.B + ((int64_t)gk + (int64_t)gn * sg.ldb));
          rb[r] = v;
        }
      };
      auto store = [&](int b) {
#pragma unroll
        for (int r = 0; r < AE; ++r) {
          const int idx = tid + r * kLpThreads;
          int m, k;
          if (!tA) { m = idx % BM; k = idx / BM; } else { k = idx % kLpBK; m = idx / kLpBK; }
          _Float16 h, l;
          split_f16(ra[r] * fa, h, l);  // scaled here, not at the load: the loads stay in flight
          sAh[b][m * kLpPitch + k] = h;
          sAl[b][m * kLpPitch + k] = l;
        }
#pragma unroll
        for (int r = 0; r < BE; ++r) {
          const int idx = tid + r * kLpThreads;
          int n, k;
          if (!tB) { k = idx % kLpBK; n = idx / kLpBK; } else { n = idx % BN; k = idx / BN; }
          _Float16 h, l;
          split_f16(rb[r] * fb, h, l);
          sBh[b][n * kLpPitch + k] = h;
          sBl[b][n * kLpPitch + k] = l;
        }
      };
      load(0);
      for (int k0 = 0; k0 < K; k0 += kLpBK) {
        store(buf);
        __syncthreads();
        if (k0 + kLpBK < K) load(k0 + kLpBK);
        half8_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int row = wm * WM + i * 16 + lr;
          ah[i] = *reinterpret_cast<const half8_t*>(&sAh[buf][row * kLpPitch + lk * 8]);
          al[i] = *reinterpret_cast<const half8_t*>(&sAl[buf][row * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int col = wn * WN + j * 16 + lr;
          bh[j] = *reinterpret_cast<const half8_t*>(&sBh[buf][col * kLpPitch + lk * 8]);
          bl[j] = *reinterpret_cast<const half8_t*>(&sBl[buf][col * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
        buf ^= 1;
      }
      const float unscale = ldexpf(1.0f, -(ea + eb));  // exact power of two
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) tot[i][j] += acc[i][j] * unscale;
    }
    // accumulator item r of lane l: row 4*(l/16) + r, column l%16
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = n0 + wn * WN + j * 16 + lr;
        if (col >= tk.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WM + i * 16 + lk * 4 + r;
          if (row < tk.M) {
            double* p = tk.C + (int64_t)row + (int64_t)col * tk.ldc;
            double v = (double)tot[i][j][r];
            if (tk.beta != 0.0) v += tk.beta * *p;
            *p = v;
          }
        }
      }
  }
}

// Operand planes.  Work item (region, r0, k0): the 64 x 32 block of rows
// r0.. and k0.. of one region, read along the source's contiguous index into
// LDS, scaled by the region's power of two (alpha folded into A), split into
// FP16 hi / lo exactly as the on-load split of k_gemm_f16x3 does, and written
// k-contiguous (pairs of halves per 32-bit store).  k >= K is written as 0.
// Maximum of |alpha A| or |B| over every region: work item (region, first
// stored column, end column) over the operand as stored (columns of K
// elements when tr, else of rows elements).
__global__ __launch_bounds__(256) void k_region_absmax(const F16Region* __restrict__ regions,
                                                      const int4* __restrict__ work, int nwork,
                                                      double* __restrict__ regionmax) {
  __shared__ double red[4];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int4 w = work[wi];
    const F16Region rg = regions[w.x];
    const int srows = rg.tr ? rg.K : rg.rows;
    const int ncol = w.z - w.y;
    double m = 0.0;
    for (int64_t e = threadIdx.x; e < (int64_t)srows * ncol; e += 256) {
      const int r = (int)(e % srows), cc = w.y + (int)(e / srows);
      m = fmax(m, fabs(rg.src[(int64_t)r + (int64_t)cc * rg.ld]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      if (rg.operand == 0) v *= fabs(rg.alpha);
      // non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<unsigned long long*>(regionmax + w.x), (unsigned long long)__double_as_longlong(v));
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_f16_presplit(const F16Region* __restrict__ regions,
                                                     const int4* __restrict__ work, int nwork,
                                                     const double* __restrict__ regionmax) {
  __shared__ double tile[64][33];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const int4 w = work[wi];
    const F16Region rg = regions[w.x];
    const int r0 = w.y, k0 = w.z;
    const double f = ldexp(rg.alpha, f16_scale_exp(regionmax[w.x]));
    for (int q = threadIdx.x; q < 64 * 32; q += 256) {
      int r, k;
      if (rg.tr) {
        k = q % 32;
        r = q / 32;
      } else {
        r = q % 64;
        k = q / 64;
      }
      const int gr = r0 + r, gk = k0 + k;
      double v = 0.0;
      if (gr < rg.rows && gk < rg.K)
        v = rg.tr ? rg.src[(int64_t)gk + (int64_t)gr * rg.ld] : rg.src[(int64_t)gr + (int64_t)gk * rg.ld];
      tile[r][k] = v * f;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 64 * 16; q += 256) {
      const int kp = q % 16, r = q / 16;
      const int gr = r0 + r;
      if (gr < rg.rows) {
        _Float16 h0, l0, h1, l1;
        split_f16(tile[r][2 * kp], h0, l0);
        split_f16(tile[r][2 * kp + 1], h1, l1);
        const int64_t o = (int64_t)gr * rg.kpad + k0 + 2 * kp;
        const uint32_t hp = (uint32_t)__builtin_bit_cast(uint16_t, h0) | ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
        const uint32_t lp = (uint32_t)__builtin_bit_cast(uint16_t, l0) | ((uint32_t)__builtin_bit_cast(uint16_t, l1) << 16);
        *reinterpret_cast<uint32_t*>(rg.hi + o) = hp;
        *reinterpret_cast<uint32_t*>(rg.lo + o) = lp;
      }
    }
    __syncthreads();
  }
}

// FP16x3 grouped GEMM on operand planes: the same tiles, K steps, LDS layout
// and MFMA sequence as k_gemm_f16x3 (so the same bits).  Each K step the
// work-group copies the 16-byte chunks of its A and B plane rows into LDS
// (one copy per element and work-group instead of a load per wave that uses
// it), the next step's chunks are loaded into registers while the current
// step's products run.
template <int BM, int BN>
__global__ __launch_bounds__(kLpThreads) void k_gemm_f16p(const GemmTile* __restrict__ tiles, int ntiles,
                                                           const GemmTask* __restrict__ tasks,
                                                           const GemmSeg* __restrict__ segs,
                                                           const F16Seg* __restrict__ fsegs,
                                                           const double* __restrict__ regionmax) {
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int ACH = (BM * 4 + kLpThreads - 1) / kLpThreads;  // 16-byte chunks per thread and plane
  constexpr int BCH = (BN * 4 + kLpThreads - 1) / kLpThreads;
  __shared__ __attribute__((aligned(16))) _Float16 sAh[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sAl[2][BM * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sBh[2][BN * kLpPitch];
  __shared__ __attribute__((aligned(16))) _Float16 sBl[2][BN * kLpPitch];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  const half8_t zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  int buf = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GemmTile tl = tiles[t];
    const GemmTask tk = tasks[tl.task];
    const int m0 = tl.m0, n0 = tl.n0;
    f32x4_t tot[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) tot[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int sidx = 0; sidx < tk.nseg; ++sidx) {
      const int si = tk.seg0 + sidx;
      const int K = segs[si].K;
      if (K <= 0) continue;
      const F16Seg fs = fsegs[si];
      const int ea = f16_scale_exp(regionmax[fs.ra]), eb = f16_scale_exp(regionmax[fs.rb]);
      // chunk c of a plane tile: row c / 4, halves 8 (c % 4) .. +7 of the K step;
      // rows beyond the task read row 0 (always present) and are written to
      // LDS as zeros.  The loads go through global-address-space pointers and
      // the zero mask is applied at the LDS store, so the prefetch of the next
      // K step stays in flight across the current step's products.
      typedef __attribute__((address_space(1))) const half8_t ghalf8_t;
      half8_t rah[ACH], ral[ACH], rbh[BCH], rbl[BCH];
      // chunk offsets of the first K step; a K step adds k0
      int64_t oa[ACH], ob[BCH];
#pragma unroll
      for (int q = 0; q < ACH; ++q) {
        const int c = tid + q * kLpThreads;
        const int row = m0 + c / 4;
        const bool ok = c < BM * 4 && row < tk.M;
        oa[q] = (int64_t)(ok ? row : 0) * fs.kpad + (c % 4) * 8;
      }
#pragma unroll
      for (int q = 0; q < BCH; ++q) {
        const int c = tid + q * kLpThreads;
        const int col = n0 + c / 4;
        const bool ok = c < BN * 4 && col < tk.N;
        ob[q] = (int64_t)(ok ? col : 0) * fs.kpad + (c % 4) * 8;
      }
      auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < ACH; ++q) {
          rah[q] = *(ghalf8_t*)(fs.ah + k0 + oa[q]);
          ral[q] = *(ghalf8_t*)(fs.al + k0 + oa[q]);
        }
#pragma unroll
        for (int q = 0; q < BCH; ++q) {
          rbh[q] = *(ghalf8_t*)(fs.bh + k0 + ob[q]);
          rbl[q] = *(ghalf8_t*)(fs.bl + k0 + ob[q]);
        }
      };
      auto store = [&](int b) {
#pragma unroll
        for (int q = 0; q < ACH; ++q) {
          const int c = tid + q * kLpThreads;
          if (c < BM * 4) {
            const bool ok = m0 + c / 4 < tk.M;
            *reinterpret_cast<half8_t*>(&sAh[b][(c / 4) * kLpPitch + (c % 4) * 8]) = ok ? rah[q] : zero8;
            *reinterpret_cast<half8_t*>(&sAl[b][(c / 4) * kLpPitch + (c % 4) * 8]) = ok ? ral[q] : zero8;
          }
        }
#pragma unroll
        for (int q = 0; q < BCH; ++q) {
          const int c = tid + q * kLpThreads;
          if (c < BN * 4) {
            const bool ok = n0 + c / 4 < tk.N;
            *reinterpret_cast<half8_t*>(&sBh[b][(c / 4) * kLpPitch + (c % 4) * 8]) = ok ? rbh[q] : zero8;
            *reinterpret_cast<half8_t*>(&sBl[b][(c / 4) * kLpPitch + (c % 4) * 8]) = ok ? rbl[q] : zero8;
          }
        }
      };
      f32x4_t acc[FM][FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      load(0);
      for (int k0 = 0; k0 < K; k0 += kLpBK) {
        store(buf);
        __syncthreads();
        if (k0 + kLpBK < K) load(k0 + kLpBK);
        half8_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int row = wm * WM + i * 16 + lr;
          ah[i] = *reinterpret_cast<const half8_t*>(&sAh[buf][row * kLpPitch + lk * 8]);
          al[i] = *reinterpret_cast<const half8_t*>(&sAl[buf][row * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int col = wn * WN + j * 16 + lr;
          bh[j] = *reinterpret_cast<const half8_t*>(&sBh[buf][col * kLpPitch + lk * 8]);
          bl[j] = *reinterpret_cast<const half8_t*>(&sBl[buf][col * kLpPitch + lk * 8]);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
        buf ^= 1;
      }
      const float unscale = ldexpf(1.0f, -(ea + eb));  // exact power of two
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) tot[i][j] += acc[i][j] * unscale;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int col = n0 + wn * WN + j * 16 + lr;
        if (col >= tk.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WM + i * 16 + lk * 4 + r;
          if (row < tk.M) {
            double* p = tk.C + (int64_t)row + (int64_t)col * tk.ldc;
            double v = (double)tot[i][j][r];
            if (tk.beta != 0.0) v += tk.beta * *p;
            *p = v;
          }
        }
      }
  }
}

template <int BM, int BN>
static void gemm_f16p_kernel_launch(Context& c, const GemmLaunch& g) {
  if (c.ext_a)
    hipExtLaunchKernelGGL((k_gemm_f16p<BM, BN>), dim3(g.grid), dim3(kLpThreads), 0, c.stream, c.ext_a, c.ext_b, 0,
                          g.tiles, g.ntiles, g.tasks, g.segs, g.f16segs, g.regionmax);
  else
    hipLaunchKernelGGL((k_gemm_f16p<BM, BN>), dim3(g.grid), dim3(kLpThreads), 0, c.stream, g.tiles, g.ntiles,
                       g.tasks, g.segs, g.f16segs, g.regionmax);
}

static const void* gemm_f16p_kernel_ptr(int cfg) {
  switch (cfg) {
    case 0: return (const void*)k_gemm_f16p<32, 32>;
    case 1: return (const void*)k_gemm_f16p<64, 64>;
    case 2: return (const void*)k_gemm_f16p<128, 128>;
    case 3: return (const void*)k_gemm_f16p<64, 128>;
    default: return (const void*)k_gemm_f16p<128, 64>;
  }
}

int gemm_occupancy_f16p(int cfg) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_f16p_kernel_ptr(cfg), kLpThreads, 0);
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
  return std::max(nb, 1);
}

template <int BM, int BN>
static void gemm_f16_kernel_launch(Context& c, const GemmLaunch& g) {
  hipLaunchKernelGGL((k_gemm_f16x3<BM, BN>), dim3(g.grid), dim3(kLpThreads), 0, c.stream, g.tiles, g.ntiles,
                     g.tasks, g.segs, g.segmax);
}

static const void* gemm_f16_kernel_ptr(int cfg) {
  switch (cfg) {
    case 0: return (const void*)k_gemm_f16x3<32, 32>;
    case 1: return (const void*)k_gemm_f16x3<64, 64>;
    case 2: return (const void*)k_gemm_f16x3<128, 128>;
    case 3: return (const void*)k_gemm_f16x3<64, 128>;
    default: return (const void*)k_gemm_f16x3<128, 64>;
  }
}

int gemm_occupancy_f16(int cfg) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_f16_kernel_ptr(cfg), kLpThreads, 0);
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
  return std::max(nb, 1);
}

void issue_gemm_f16_prep(Context& c, const GemmLaunch& g) {
  if (g.ntiles <= 0) return;
  auto check = [] {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
  };
  if (g.presplit) {
    // maxima of every distinct operand block, then its planes
    launch_fill(c, g.regionmax, std::max(1, g.nregions), 0.0);
    if (g.nregionmaxwork > 0) {
      hipLaunchKernelGGL(k_region_absmax, dim3(std::min(g.nregionmaxwork, c.num_cu * 8)), dim3(256), 0, c.stream,
                         g.regions, g.regionmaxwork, g.nregionmaxwork, g.regionmax);
      check();
    }
    if (g.nsplitwork > 0) {
      hipLaunchKernelGGL(k_f16_presplit, dim3(std::min(g.nsplitwork, c.num_cu * 8)), dim3(256), 0, c.stream,
                         g.regions, g.splitwork, g.nsplitwork, g.regionmax);
      check();
    }
    return;
  }
  // operand maxima of every segment, measured right before the product
  launch_fill(c, g.segmax, std::max(2, 2 * g.nseg), 0.0);
  if (g.nmaxwork > 0) {
    hipLaunchKernelGGL(k_seg_absmax, dim3(std::min(g.nmaxwork, c.num_cu * 8)), dim3(256), 0, c.stream, g.tasks,
                       g.segs, g.segtask, g.maxwork, g.nmaxwork, g.segmax);
    check();
  }
}

void issue_gemm_f16_main(Context& c, const GemmLaunch& g) {
  if (g.ntiles <= 0) return;
  if (g.presplit) {
    switch (g.cfg) {
      case 0: gemm_f16p_kernel_launch<32, 32>(c, g); break;
      case 1: gemm_f16p_kernel_launch<64, 64>(c, g); break;
      case 2: gemm_f16p_kernel_launch<128, 128>(c, g); break;
      case 3: gemm_f16p_kernel_launch<64, 128>(c, g); break;
      default: gemm_f16p_kernel_launch<128, 64>(c, g); break;
    }
  } else {
    switch (g.cfg) {
      case 0: gemm_f16_kernel_launch<32, 32>(c, g); break;
      case 1: gemm_f16_kernel_launch<64, 64>(c, g); break;
      case 2: gemm_f16_kernel_launch<128, 128>(c, g); break;
      case 3: gemm_f16_kernel_launch<64, 128>(c, g); break;
      default: gemm_f16_kernel_launch<128, 64>(c, g); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
}

void issue_gemm_f16(Context& c, const GemmLaunch& g) {
  issue_gemm_f16_prep(c, g);
  issue_gemm_f16_main(c, g);
}

// Lane-layout probe of v_mfma_f32_16x16x32_bf16 (one wave, in-bounds):
//  mode 0: a = lane+1, b = 1; mode 1: a = 1, b = lane+1 (row / column sums)
//  mode 2: A[i][0] = i+1, A[i][1] = 1, B[0][j] = 1, B[1][j] = 16(j+1) under
//          A[i][k] <- lane i + 16(k/8) element k%8, B[k][j] <- lane j + 16(k/8)
__global__ void k_mfma_bf16_probe(float* out) {
  const int l = threadIdx.x;
  const int r16 = l & 15, kb = l >> 4;
  for (int mode = 0; mode < 3; ++mode) {
    bf16x8_t a, b;
    for (int e = 0; e < 8; ++e) {
      float av, bv;
      if (mode == 0) {
        av = (float)(l + 1);
        bv = 1.f;
      } else if (mode == 1) {
        av = 1.f;
        bv = (float)(l + 1);
      } else {
        const int k = kb * 8 + e;
        av = k == 0 ? (float)(r16 + 1) : (k == 1 ? 1.f : 0.f);
        bv = k == 0 ? 1.f : (k == 1 ? 16.f * (r16 + 1) : 0.f);
      }
      a[e] = (__bf16)av;
      b[e] = (__bf16)bv;
    }
    f32x4_t acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out[mode * 256 + l * 4 + r] = acc[r];
  }
}

void launch_mfma_bf16_probe(Context& c, float* out) {
  hipLaunchKernelGGL(k_mfma_bf16_probe, dim3(1), dim3(64), 0, c.stream, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
}

template <int BM, int BN>
static void gemm_lp_kernel_launch(Context& c, const GemmLaunch& g) {
  if (c.ext_a)
    hipExtLaunchKernelGGL((k_gemm_bf16x3<BM, BN>), dim3(g.grid), dim3(kLpThreads), 0, c.stream, c.ext_a, c.ext_b,
                          0, g.tiles, g.ntiles, g.tasks, g.segs);
  else
    hipLaunchKernelGGL((k_gemm_bf16x3<BM, BN>), dim3(g.grid), dim3(kLpThreads), 0, c.stream, g.tiles, g.ntiles,
                       g.tasks, g.segs);
}

static const void* gemm_lp_kernel_ptr(int cfg) {
  switch (cfg) {
    case 0: return (const void*)k_gemm_bf16x3<32, 32>;
    case 1: return (const void*)k_gemm_bf16x3<64, 64>;
    case 2: return (const void*)k_gemm_bf16x3<128, 128>;
    case 3: return (const void*)k_gemm_bf16x3<64, 128>;
    default: return (const void*)k_gemm_bf16x3<128, 64>;
  }
}

int gemm_occupancy_lp(int cfg) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, gemm_lp_kernel_ptr(cfg), kLpThreads, 0);
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
  return std::max(nb, 1);
}

void issue_gemm_lp(Context& c, const GemmLaunch& g) {
  if (g.ntiles <= 0) return;
  switch (g.cfg) {
    case 0: gemm_lp_kernel_launch<32, 32>(c, g); break;
    case 1: gemm_lp_kernel_launch<64, 64>(c, g); break;
    case 2: gemm_lp_kernel_launch<128, 128>(c, g); break;
    case 3: gemm_lp_kernel_launch<64, 128>(c, g); break;
    default: gemm_lp_kernel_launch<128, 64>(c, g); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e));
}

void launch_gemm_lp(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int cfg) {
  launch_gemm_arith(c, tasks, segs, cfg, 1);
}

}  // namespace dmrj
