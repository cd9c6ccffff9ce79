// Truncated SVD of a two-site tensor (fsvd! of the reference's twostep,
// algs/DMRG.jl:387-395): the tensor is cut into centre-charge sectors, every
// sector is decomposed by the selected backend (batched one-sided Jacobi for
// sectors up to 64, batched symmetric eigensolves of the augmented or Gram
// matrices, block one-sided Jacobi, or rocSOLVER's SVDs), the singular
// values of all sectors are sorted and truncated together, and the kept
// vectors are split into the left and right tensors.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "../../include/dmrj.h"
#include "internal.h"

#define HIP_OK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw dmrj::Error(-3, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

namespace dmrj {

// Cost model of one step of rocSOLVER's batched symmetric tridiagonalisation
// on MI355X, from its kernel trace on DMRG sectors: about 20 us of fixed
// latency per step (five small kernels) plus about 1e-6 us per element of
// the padded N x N matrices of the batch.  It only groups sectors into
// batches; every grouping gives a backward-stable result.
constexpr double kEigColUs = 20.0;
constexpr double kEigElemUs = 1.0e-6;

// Batches for a list of eigenproblem sizes sorted in descending order:
// contiguous classes, each solved as one batch padded to its first (largest)
// size N, chosen to minimise the sum over classes of
// N * (kEigColUs + kEigElemUs * count * N^2), so a class is split off only
// when the padding it no longer carries outweighs its extra steps.  Returns
// the first index of every class followed by n.size().
static std::vector<int> eig_classes(const std::vector<int>& n, int mode) {
  const int cnt = (int)n.size();
  std::vector<int> cls;
  if (mode != 1) {  // 0: one batch; 2: one batch per sector; 3: batches of two sectors
    const int step = mode == 0 ? std::max(cnt, 1) : mode == 2 ? 1 : 2;
    for (int i = 0; i < cnt; i += step) cls.push_back(i);
    cls.push_back(cnt);
    return cls;
  }
  std::vector<double> best(cnt + 1, 0.0);
  std::vector<int> cut(cnt + 1, cnt);
  for (int i = cnt - 1; i >= 0; --i) {
    const double Ni = (double)n[i];
    best[i] = 1e300;
    for (int j = i + 1; j <= cnt; ++j) {
      const double cost = Ni * (kEigColUs + kEigElemUs * (j - i) * Ni * Ni) + best[j];
      if (cost < best[i]) {
        best[i] = cost;
        cut[i] = j;
      }
    }
  }
  for (int i = 0; i < cnt; i = cut[i]) cls.push_back(i);
  cls.push_back(cnt);
  return cls;
}

SvdResult svd_trunc(Context& c, const Tensor& AA, int64_t m, int64_t minm, double cutoff, int dir) {
  const SiteMeta& x = AA.site;
  if (AA.kind != 0 || x.d1 == 0) throw Error(DMRJ_ERR_ARG, "svd_trunc2 needs a two-site tensor");
  // complex tensors: every array below holds a real and an imaginary plane
  const bool z = AA.cplx != 0;
  const int zf = z ? 2 : 1;
  const int d1 = x.d1, d2 = x.d2;
  const Space& S = x.left;
  const Space& T = x.right;
  // centre-charge sectors: rows (a, j1), cols (j2, b)
  struct Sec {
    Charge Q;
    std::vector<std::pair<int, int>> rows, cols;  // (a,j1) / (j2,b)
    std::vector<int64_t> roff, coff;
    int64_t R = 0, C = 0;
    double *M = nullptr, *U = nullptr, *Vt = nullptr, *s = nullptr, *E = nullptr;
    int kmin = 0, keep = 0;
    // sectors solved for one orthonormal factor (block Jacobi, backend 4, or
    // Gram eigensolve, backend 5); the other factor comes from a GEMM with M.
    // colsrc / coldiv: source column (rowsX elements) and divisor of the k-th
    // largest singular vector of that side
    double* X = nullptr;
    int bj = 0, rowsX = 0, ncolsX = 0;
    int qr = 0;  // one-factor sector solved through its LQ / QR triangle (X: K x K factor)
    int gram = 0;  // orthonormal factor = eigenvectors of the Gram matrix (orthonormal to working precision)
    std::vector<int> perm;
    std::vector<double> hn;
    std::vector<const double*> colsrc;
    std::vector<double> coldiv;
  };
  std::map<Charge, int> qidx;
  std::vector<Sec> secs;
  for (int a = 0; a < S.nsec(); ++a)
    for (int j1 = 0; j1 < d1; ++j1) {
      const Charge Q = S.q[a] + x.qp1[j1];
      auto it = qidx.find(Q);
      if (it == qidx.end()) {
        it = qidx.insert({Q, (int)secs.size()}).first;
        secs.push_back(Sec());
        secs.back().Q = Q;
      }
      Sec& s = secs[it->second];
      s.rows.push_back({a, j1});
      s.roff.push_back(s.R);
      s.R += S.dim[a];
    }
  for (int j2 = 0; j2 < d2; ++j2)
    for (int b = 0; b < T.nsec(); ++b) {
      const Charge Q = T.q[b] - x.qp2[j2];
      auto it = qidx.find(Q);
      if (it == qidx.end()) continue;
      Sec& s = secs[it->second];
      s.cols.push_back({j2, b});
      s.coff.push_back(s.C);
      s.C += T.dim[b];
    }
  // sector order by charge
  std::vector<int> order;
  for (auto& kv : qidx) order.push_back(kv.second);
  std::vector<CopyOp> gather;
  // the singular values of all sectors in one array, in sector order, so that
  // they come to the host in one copy
  int64_t total_s = 0;
  for (int si : order) {
    Sec& s = secs[si];
    if (s.R == 0 || s.C == 0) continue;
    s.kmin = (int)std::min(s.R, s.C);
    total_s += s.kmin;
  }
  double* sall = (double*)c.alloc.alloc(std::max<int64_t>(total_s, 1) * sizeof(double));
  int64_t soff_all = 0;
  for (int si : order) {
    Sec& s = secs[si];
    if (s.R == 0 || s.C == 0) continue;
    s.M = (double*)c.alloc.alloc(s.R * s.C * zf * sizeof(double));
    s.U = (double*)c.alloc.alloc(s.R * s.kmin * zf * sizeof(double));
    s.Vt = (double*)c.alloc.alloc((int64_t)s.kmin * s.C * zf * sizeof(double));
    s.s = sall + soff_all;
    soff_all += s.kmin;
    s.E = (double*)c.alloc.alloc(std::max(1, s.kmin) * sizeof(double));
    for (size_t r = 0; r < s.rows.size(); ++r)
      for (size_t q = 0; q < s.cols.size(); ++q) {
        const int a = s.rows[r].first, j1 = s.rows[r].second;
        const int j2 = s.cols[q].first, b = s.cols[q].second;
        const int xb = x.block(a, j1 + d1 * j2);
        if (xb < 0 || x.blocks[xb].b != b) throw Error(DMRJ_ERR_STATE, "two-site block table inconsistent");
        gather.push_back({AA.data + x.blocks[xb].off, s.M + s.roff[r] + s.coff[q] * s.R, S.dim[a], s.R,
                          S.dim[a], T.dim[b], nullptr, nullptr, 1.0});
        if (z)
          gather.push_back({AA.data + x.size + x.blocks[xb].off, s.M + s.R * s.C + s.roff[r] + s.coff[q] * s.R,
                            S.dim[a], s.R, S.dim[a], T.dim[b], nullptr, nullptr, 1.0});
      }
  }
  launch_copy(c, gather);
  double aa2 = 0.0;  // |AA|^2
  {
    // never hand non-finite data to the solver
    const double* xs[1] = {AA.data};
    launch_dots(c, 1, xs, xs, AA.nstore(), c.scal + S_CHK);
    double chk = 0.0;
    HIP_OK(hipMemcpyAsync(&c.hscal[S_CHK], c.scal + S_CHK, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    chk = c.hscal[S_CHK];
    aa2 = chk;
    if (!std::isfinite(chk)) {
      for (auto& s : secs) {
        c.alloc.release(s.M);
        c.alloc.release(s.U);
        c.alloc.release(s.Vt);
        c.alloc.release(s.E);
      }
      c.alloc.release(sall);
      throw Error(DMRJ_ERR_NUMERIC, "two-site tensor holds NaN/Inf; SVD not attempted");
    }
  }
  auto* hb = (rocblas_handle)c.rocblas;
  rocblas_int* info = (rocblas_int*)c.alloc.alloc(sizeof(rocblas_int) * std::max<size_t>(1, secs.size()));
  int nsvd = 0;
  const int psvd = c.prof_begin();
  double svd_flops = 0.0;
  // small sectors: one batched Jacobi launch; large ones: rocSOLVER gesdd
  std::vector<SmallSvdJob> small;
  size_t small_lds = 0;
  std::vector<char> is_small(secs.size(), 0);
  for (int si : order) {
    Sec& s = secs[si];
    if (z || s.kmin == 0 || s.kmin > 64 || c.svd_backend == 1 || c.svd_backend == 2) continue;
    const size_t l = small_svd_lds((int)s.R, (int)s.C);
    if (l > kSmallSvdMaxLds) continue;
    is_small[si] = 1;
    small_lds = std::max(small_lds, l);
    small.push_back({s.M, s.U, s.s, s.Vt, (int)s.R, (int)s.C});
    const double mm = (double)std::max(s.R, s.C), nn = (double)std::min(s.R, s.C);
    svd_flops += 4.0 * mm * nn * nn + 8.0 * nn * nn * nn;
  }
  const bool small_on_side = launch_svd_small(c, small, small_lds);
  // every exit from here on, exceptions included, orders the main stream
  // after the side-stream kernel, so no later synchronisation of the main
  // stream can release memory that kernel still uses
  struct SideJoin {
    Context& c;
    bool on;
    ~SideJoin() {
      if (on) (void)hipStreamWaitEvent(c.stream, c.side_out, 0);
    }
  } side_join{c, small_on_side};
  // large sectors, backend 4: batched block one-sided Jacobi on M (dir > 0,
  // orthonormal U wanted) or on M^T (dir < 0, orthonormal V wanted); the
  // other factor is one grouped GEMM with M after the truncation
  int* bjflag = nullptr;
  double* bjnorm = nullptr;
  if (!z && c.svd_backend == 4) {
    std::vector<BjJob> jobs;
    std::vector<int> noff;  // offset of every job's columns in the shared norm / permutation arrays
    int64_t ncol_all = 0;
    const double tiny = 8.0 * 2.220446049250313e-16 * std::sqrt(aa2);
    for (int si : order) {
      Sec& s = secs[si];
      if (s.kmin == 0 || is_small[si]) continue;
      const int rows = (int)(dir > 0 ? s.R : s.C), ncols = (int)(dir > 0 ? s.C : s.R);
      // the kernel's static LDS comes on top of the panel
      const int bw = bj_panel_width(rows, c.lds_max, bj_static_lds());
      if (bw == 0) continue;  // too tall for LDS: augmented solver below
      s.X = (double*)c.alloc.alloc((int64_t)rows * ncols * sizeof(double));
      if (dir > 0)
        HIP_OK(hipMemcpyAsync(s.X, s.M, (int64_t)rows * ncols * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
      else
        launch_transpose(c, s.M, (int)s.R, (int)s.C, s.X);
      int nblk = (ncols + bw - 1) / bw;
      nblk += nblk & 1;
      s.bj = 1;
      s.rowsX = rows;
      s.ncolsX = ncols;
      is_small[si] = 1;  // handled
      jobs.push_back({s.X, nullptr, nullptr, nullptr, tiny * tiny * rows, rows, ncols, nblk, bw});
      noff.push_back((int)ncol_all);
      ncol_all += ncols;
      const double mm = (double)std::max(s.R, s.C), nn = (double)std::min(s.R, s.C);
      svd_flops += 4.0 * mm * nn * nn + 8.0 * nn * nn * nn;
    }
    if (!jobs.empty()) {
      const int nj = (int)jobs.size();
      bjflag = (int*)c.alloc.alloc(nj * sizeof(int));
      bjnorm = (double*)c.alloc.alloc(std::max<int64_t>(ncol_all, 1) * sizeof(double));
      size_t lds = 0;
      for (int j = 0; j < nj; ++j) {
        jobs[j].flag = bjflag + j;
        jobs[j].S = bjnorm + noff[j];
        lds = std::max(lds, bj_lds_bytes(jobs[j].rows, jobs[j].bw));
      }
      std::vector<int> hperm(ncol_all);
      std::vector<double> hnorm(ncol_all);
      std::vector<char> active(nj, 1);
      std::vector<int> hflag(nj, 0);
      auto norms_to_host = [&](bool all) {
        std::vector<int2> nw;
        for (int j = 0; j < nj; ++j)
          if (all || active[j])
            for (int col = 0; col < jobs[j].ncols; ++col) nw.push_back(make_int2(j, col));
        auto* dj = (BjJob*)c.upload(jobs.data(), nj * sizeof(BjJob));
        launch_bj_norms(c, dj, nw);
        HIP_OK(hipMemcpyAsync(hnorm.data(), bjnorm, ncol_all * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      };
      // de Rijk ordering: columns by decreasing norm before every sweep
      auto sort_active = [&]() {
        for (int j = 0; j < nj; ++j) {
          if (!active[j]) continue;
          int* p = hperm.data() + noff[j];
          const double* nrm = hnorm.data() + noff[j];
          std::iota(p, p + jobs[j].ncols, 0);
          std::stable_sort(p, p + jobs[j].ncols, [&](int a, int b) { return nrm[a] > nrm[b]; });
        }
      };
      norms_to_host(true);
      HIP_OK(hipStreamSynchronize(c.stream));
      sort_active();
      constexpr int kBjMaxSweeps = 60;
      int sweep = 0;
      for (; sweep < kBjMaxSweeps; ++sweep) {
        std::vector<int2> work;
        int nsteps = 0;
        const int* dperm = (const int*)c.upload(hperm.data(), ncol_all * sizeof(int));
        for (int j = 0; j < nj; ++j) {
          jobs[j].perm = dperm + noff[j];
          if (!active[j]) continue;
          for (int k = 0; k < jobs[j].nblk / 2; ++k) work.push_back(make_int2(j, k));
          nsteps = std::max(nsteps, jobs[j].nblk - 1);
        }
        if (work.empty()) break;
        auto* djobs = (BjJob*)c.upload(jobs.data(), nj * sizeof(BjJob));
        HIP_OK(hipMemsetAsync(bjflag, 0, nj * sizeof(int), c.stream));
        launch_bj_sweep(c, djobs, work, nsteps, 8, lds);
        norms_to_host(false);
        HIP_OK(hipMemcpyAsync(hflag.data(), bjflag, nj * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
        bool any = false;
        for (int j = 0; j < nj; ++j) {
          if (active[j] && hflag[j] == 0) active[j] = 0;  // a sweep without rotations: converged
          any = any || active[j];
        }
        if (!any) break;
        sort_active();
      }
      if (sweep == kBjMaxSweeps) throw Error(DMRJ_ERR_SOLVER, "block Jacobi SVD did not converge");
      c.bj_sweeps += sweep + 1;
      c.bj_calls += 1;
      // final column norms of every sector are in hnorm (the last sweep of
      // each sector rotated nothing, so its norms are current)
      int jj = 0;
      for (int si : order) {
        Sec& s = secs[si];
        if (!s.bj) continue;
        s.hn.assign(hnorm.begin() + noff[jj], hnorm.begin() + noff[jj] + s.ncolsX);
        ++jj;
      }
    }
  }
  // large sectors, backend 5: batched symmetric eigensolve of the Gram matrix
  // of the side that must come out orthonormal (M M^T for dir > 0, M^T M for
  // dir < 0: size n = R or C instead of R + C); eigenvalues sigma^2 carry an
  // absolute error ~eps |M|^2; the other factor by the grouped GEMM below
  double* gramZ = nullptr;
  double* gramW = nullptr;
  double* gramE = nullptr;
  rocblas_int* graminfo = nullptr;
  int ngram = 0;
  int64_t gramWtot = 0;
  std::vector<int> gramsec, gramNb;       // per Gram sector: its batch's padded size
  std::vector<int64_t> gramZoff, gramWoff;  // and its offsets in gramZ / gramW
  if (!z && c.svd_backend == 5) {
    for (int si : order) {
      Sec& s = secs[si];
      if (s.kmin == 0 || is_small[si]) continue;
      gramsec.push_back(si);
    }
    if (!gramsec.empty()) {
      auto gsize = [&](int si) { return (int)(dir > 0 ? secs[si].R : secs[si].C); };
      std::stable_sort(gramsec.begin(), gramsec.end(), [&](int a, int b) { return gsize(a) > gsize(b); });
      ngram = (int)gramsec.size();
      std::vector<int> gn(ngram);
      for (int b = 0; b < ngram; ++b) gn[b] = gsize(gramsec[b]);
      const std::vector<int> gcls = eig_classes(gn, c.eig_batching);  // batches by size
      gramNb.resize(ngram);
      gramZoff.resize(ngram);
      gramWoff.resize(ngram);
      int64_t ztot = 0;
      for (size_t k = 0; k + 1 < gcls.size(); ++k)
        for (int b = gcls[k]; b < gcls[k + 1]; ++b) {
          gramNb[b] = gn[gcls[k]];
          gramZoff[b] = ztot;
          gramWoff[b] = gramWtot;
          ztot += (int64_t)gramNb[b] * gramNb[b];
          gramWtot += gramNb[b];
        }
      gramZ = (double*)c.alloc.alloc(ztot * sizeof(double));
      gramW = (double*)c.alloc.alloc(gramWtot * sizeof(double));
      gramE = (double*)c.alloc.alloc(gramWtot * sizeof(double));
      graminfo = (rocblas_int*)c.alloc.alloc(sizeof(rocblas_int) * ngram);
      launch_fill(c, gramZ, ztot, 0.0);
      std::vector<GemmTask> tasks;
      std::vector<GemmSeg> segs;
      std::vector<std::vector<int2>> padjobs(gcls.size() - 1);
      for (size_t k = 0; k + 1 < gcls.size(); ++k)
        for (int b = gcls[k]; b < gcls[k + 1]; ++b) padjobs[k].push_back(make_int2(b - gcls[k], gn[b]));
      for (int b = 0; b < ngram; ++b) {
        Sec& s = secs[gramsec[b]];
        const int n = (int)(dir > 0 ? s.R : s.C);
        double* Zb = gramZ + gramZoff[b];
        tasks.push_back({Zb, gramNb[b], n, n, (int)segs.size(), 1, 0.0});
        if (dir > 0)
          segs.push_back({s.M, s.M, s.R, s.R, (int)s.C, 2, 1.0});  // M M^T
        else
          segs.push_back({s.M, s.M, s.R, s.R, (int)s.R, 1, 1.0});  // M^T M
        s.bj = 1;
        s.gram = 1;
        s.rowsX = n;
        is_small[gramsec[b]] = 1;  // handled
        const double nn = (double)n, kk = (double)(dir > 0 ? s.C : s.R);
        svd_flops += 2.0 * nn * nn * kk + 9.0 * nn * nn * nn;
      }
      launch_gemm(c, tasks, segs, -1);
      // padding rows / columns: -1 on the diagonal keeps their eigenvalues
      // below every sigma^2 >= 0 and their eigenvectors outside the sector
      for (size_t k = 0; k + 1 < gcls.size(); ++k) {
        const int b0 = gcls[k], Nk = gramNb[b0];
        launch_pad_diag(c, gramZ + gramZoff[b0], Nk, padjobs[k], -1.0);
        rocblas_status st = rocsolver_dsyevd_strided_batched(
            hb, rocblas_evect_original, rocblas_fill_upper, Nk, gramZ + gramZoff[b0], Nk, (int64_t)Nk * Nk,
            gramW + gramWoff[b0], Nk, gramE + gramWoff[b0], Nk, graminfo + b0, gcls[k + 1] - b0);
        if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver_dsyevd_strided_batched failed");
      }
    }
  }
  // large sectors, default backend: one batched symmetric eigensolve of the
  // augmented matrices [[0, M], [M^T, 0]] (eigenvalues +-sigma, absolute
  // accuracy eps |M| like a bidiagonal SVD, no squaring of M)
  double* aug = nullptr;
  double* augW = nullptr;
  double* augE = nullptr;
  rocblas_int* auginfo = nullptr;
  int naug = 0;
  // LQ / QR-reduced sectors of the default backend
  double *qrX = nullptr, *qrT = nullptr, *qaug = nullptr, *qW = nullptr, *qE = nullptr, *qVt = nullptr;
  rocblas_int* qinfo = nullptr;
  int nqr = 0;
  if (z) {
    // complex: every sector through batched Hermitian eigensolves of
    // [[0, M], [M^H, 0]] (eigenvalues +-sigma), batched by size (eig_classes)
    std::vector<int> big;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.kmin == 0) continue;
      big.push_back(si);
    }
    if (!big.empty()) {
      std::stable_sort(big.begin(), big.end(),
                       [&](int a, int b) { return secs[a].R + secs[a].C > secs[b].R + secs[b].C; });
      naug = (int)big.size();
      std::vector<int> nsz(naug);
      for (int b = 0; b < naug; ++b) nsz[b] = (int)(secs[big[b]].R + secs[big[b]].C);
      const std::vector<int> cls = eig_classes(nsz, c.eig_batching);
      const int ncls = (int)cls.size() - 1;
      std::vector<int64_t> aoff(ncls + 1, 0), voff(ncls + 1, 0);
      for (int k = 0; k < ncls; ++k) {
        const int64_t Nk = nsz[cls[k]], cnt = cls[k + 1] - cls[k];
        aoff[k + 1] = aoff[k] + 2 * Nk * Nk * cnt;  // interleaved complex
        voff[k + 1] = voff[k] + Nk * cnt;
      }
      aug = (double*)c.alloc.alloc(aoff[ncls] * sizeof(double));
      augW = (double*)c.alloc.alloc(voff[ncls] * sizeof(double));
      augE = (double*)c.alloc.alloc(voff[ncls] * sizeof(double));
      auginfo = (rocblas_int*)c.alloc.alloc(sizeof(rocblas_int) * naug);
      launch_fill(c, aug, aoff[ncls], 0.0);
      std::vector<AugJobZ> jobs;
      for (int k = 0; k < ncls; ++k) {
        const int Nk = nsz[cls[k]];
        const int64_t NNk = (int64_t)Nk * Nk;
        for (int b = cls[k]; b < cls[k + 1]; ++b) {
          Sec& s = secs[big[b]];
          const int64_t j = b - cls[k];
          jobs.push_back({s.M, aug + aoff[k] + 2 * j * NNk, augW + voff[k] + j * Nk, s.U, s.s, s.Vt, (int)s.R,
                          (int)s.C, Nk, s.kmin});
          is_small[big[b]] = 1;  // handled
          const double n2 = (double)Nk;
          svd_flops += 4.0 * 9.0 * n2 * n2 * n2;
        }
      }
      launch_aug_place_z(c, jobs);
      for (int k = 0; k < ncls; ++k) {
        const int Nk = nsz[cls[k]];
        rocblas_status st = rocsolver_zheevd_strided_batched(
            hb, rocblas_evect_original, rocblas_fill_upper, Nk, (rocblas_double_complex*)(aug + aoff[k]), Nk,
            (int64_t)Nk * Nk, augW + voff[k], Nk, augE + voff[k], Nk, auginfo + cls[k], cls[k + 1] - cls[k]);
        if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver_zheevd_strided_batched failed");
      }
      launch_aug_extract_z(c, jobs);
    }
  } else if (c.svd_backend == 0) {
    std::vector<int> big, qrs;
    int N = 0;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.kmin == 0 || is_small[si]) continue;
      // short / long side, the short one being the side that must come out
      // orthonormal (U for dir > 0, V for dir < 0)
      const int sh = (int)(dir > 0 ? s.R : s.C), lg = (int)(dir > 0 ? s.C : s.R);
      if (2 * (int64_t)lg >= 3 * (int64_t)sh) {
        qrs.push_back(si);
      } else {
        big.push_back(si);
        N = std::max<int>(N, (int)(s.R + s.C));
      }
    }
    // Sectors at least 1.5 times longer than the side that must come out
    // orthonormal (the enlarged tensors of 3S): Householder LQ (dir > 0,
    // M = L Q: U(M) = U(L)) or QR (dir < 0, M = Q R: V(M) = V(R) = U(R^T))
    // of the sector first, then the augmented eigensolve of the K x K
    // triangle B = L or R^T at size 2K instead of R + C.  Backward stable like
    // the full augmented form (sigma to eps |M| absolute); the other factor
    // comes from the grouped GEMM with M below, as for the one-factor backends.
    if (!qrs.empty()) {
      // batches by size as for the augmented form below (eig_classes on the
      // 2K augmented sizes), each padded to its own largest K and L
      const int sdir = dir;
      auto shortside = [&](int si) { return (int)(sdir > 0 ? secs[si].R : secs[si].C); };
      auto longside = [&](int si) { return (int)(sdir > 0 ? secs[si].C : secs[si].R); };
      std::stable_sort(qrs.begin(), qrs.end(), [&](int a, int b) { return shortside(a) > shortside(b); });
      nqr = (int)qrs.size();
      std::vector<int> n2k(nqr);
      for (int b = 0; b < nqr; ++b) n2k[b] = 2 * shortside(qrs[b]);
      const std::vector<int> qcls = eig_classes(n2k, c.eig_batching);
      const int nqc = (int)qcls.size() - 1;
      std::vector<int> cK(nqc), cL(nqc, 0);
      std::vector<int64_t> xo(nqc + 1, 0), to(nqc + 1, 0), ao(nqc + 1, 0), wo(nqc + 1, 0), vo(nqc + 1, 0);
      for (int k = 0; k < nqc; ++k) {
        const int cnt = qcls[k + 1] - qcls[k];
        cK[k] = shortside(qrs[qcls[k]]);
        for (int b = qcls[k]; b < qcls[k + 1]; ++b) cL[k] = std::max(cL[k], longside(qrs[b]));
        xo[k + 1] = xo[k] + (int64_t)cK[k] * cL[k] * cnt;
        to[k + 1] = to[k] + (int64_t)cK[k] * cnt;
        ao[k + 1] = ao[k] + (int64_t)4 * cK[k] * cK[k] * cnt;
        wo[k + 1] = wo[k] + (int64_t)2 * cK[k] * cnt;
        vo[k + 1] = vo[k] + (int64_t)cK[k] * cK[k] * cnt;
      }
      qrX = (double*)c.alloc.alloc(xo[nqc] * sizeof(double));
      qrT = (double*)c.alloc.alloc(to[nqc] * sizeof(double));
      launch_fill(c, qrX, xo[nqc], 0.0);
      std::vector<CopyOp> place;
      for (int k = 0; k < nqc; ++k) {
        const int ldq = dir > 0 ? cK[k] : cL[k];
        const int64_t QS = (int64_t)cK[k] * cL[k];
        for (int b = qcls[k]; b < qcls[k + 1]; ++b) {
          Sec& s = secs[qrs[b]];
          place.push_back({s.M, qrX + xo[k] + (b - qcls[k]) * QS, s.R, ldq, (int)s.R, (int)s.C, nullptr, nullptr, 1.0});
        }
      }
      launch_copy(c, place);
      for (int k = 0; k < nqc; ++k) {
        const int Kk = cK[k], Lk = cL[k], cnt = qcls[k + 1] - qcls[k];
        const int64_t QS = (int64_t)Kk * Lk;
        rocblas_status st =
            dir > 0 ? rocsolver_dgelqf_strided_batched(hb, Kk, Lk, qrX + xo[k], Kk, QS, qrT + to[k], Kk, cnt)
                    : rocsolver_dgeqrf_strided_batched(hb, Lk, Kk, qrX + xo[k], Lk, QS, qrT + to[k], Kk, cnt);
        if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver LQ/QR (strided batched) failed");
      }
      qaug = (double*)c.alloc.alloc(ao[nqc] * sizeof(double));
      qW = (double*)c.alloc.alloc(wo[nqc] * sizeof(double));
      qE = (double*)c.alloc.alloc(wo[nqc] * sizeof(double));
      qVt = (double*)c.alloc.alloc(vo[nqc] * sizeof(double));
      qinfo = (rocblas_int*)c.alloc.alloc(sizeof(rocblas_int) * nqr);
      launch_fill(c, qaug, ao[nqc], 0.0);
      std::vector<TriJob> tj;
      for (int k = 0; k < nqc; ++k) {
        const int Nq = 2 * cK[k], ldq = dir > 0 ? cK[k] : cL[k];
        const int64_t QS = (int64_t)cK[k] * cL[k], NNq = (int64_t)Nq * Nq;
        for (int b = qcls[k]; b < qcls[k + 1]; ++b) {
          Sec& s = secs[qrs[b]];
          tj.push_back({qrX + xo[k] + (b - qcls[k]) * QS, qaug + ao[k] + (b - qcls[k]) * NNq, ldq, s.kmin, Nq,
                        dir > 0 ? 0 : 1});
        }
      }
      launch_tri_place(c, tj);
      for (int k = 0; k < nqc; ++k) {
        const int Nq = 2 * cK[k];
        rocblas_status st = rocsolver_dsyevd_strided_batched(
            hb, rocblas_evect_original, rocblas_fill_upper, Nq, qaug + ao[k], Nq, (int64_t)Nq * Nq, qW + wo[k], Nq,
            qE + wo[k], Nq, qinfo + qcls[k], qcls[k + 1] - qcls[k]);
        if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver_dsyevd_strided_batched failed");
      }
      std::vector<AugJob> jobs;
      for (int k = 0; k < nqc; ++k) {
        const int Kk = cK[k], Nq = 2 * Kk;
        const int64_t NNq = (int64_t)Nq * Nq;
        for (int b = qcls[k]; b < qcls[k + 1]; ++b) {
          Sec& s = secs[qrs[b]];
          const int K = s.kmin;
          const int64_t j = b - qcls[k];
          s.X = (double*)c.alloc.alloc((int64_t)K * K * sizeof(double));
          jobs.push_back({qaug + ao[k] + j * NNq, qW + wo[k] + j * Nq, s.X, s.s, qVt + vo[k] + j * Kk * Kk, K, K, Nq,
                          K});
          s.bj = 1;
          s.qr = 1;
          s.rowsX = K;
          is_small[qrs[b]] = 1;  // handled
          const double kk = (double)K, ll = (double)longside(qrs[b]), n2 = 2.0 * kk;
          svd_flops += 2.0 * ll * kk * kk - 2.0 / 3.0 * kk * kk * kk + 9.0 * n2 * n2 * n2;
        }
      }
      launch_aug_extract(c, jobs);
    }
    if (!big.empty()) {
      // Size classes (eig_classes), each one batched eigensolve padded to its
      // own largest augmented size: the eigensolve takes about N steps, each
      // a fixed latency plus a pass over the padded matrices of the batch.
      std::stable_sort(big.begin(), big.end(),
                       [&](int a, int b) { return secs[a].R + secs[a].C > secs[b].R + secs[b].C; });
      naug = (int)big.size();
      std::vector<int> nsz(naug);
      for (int b = 0; b < naug; ++b) nsz[b] = (int)(secs[big[b]].R + secs[big[b]].C);
      const std::vector<int> cls = eig_classes(nsz, c.eig_batching);
      const int ncls = (int)cls.size() - 1;
      std::vector<int> cN(ncls);
      std::vector<int64_t> aoff(ncls + 1, 0), voff(ncls + 1, 0);
      for (int k = 0; k < ncls; ++k) {
        cN[k] = (int)(secs[big[cls[k]]].R + secs[big[cls[k]]].C);
        aoff[k + 1] = aoff[k] + (int64_t)cN[k] * cN[k] * (cls[k + 1] - cls[k]);
        voff[k + 1] = voff[k] + (int64_t)cN[k] * (cls[k + 1] - cls[k]);
      }
      aug = (double*)c.alloc.alloc(aoff[ncls] * sizeof(double));
      augW = (double*)c.alloc.alloc(voff[ncls] * sizeof(double));
      augE = (double*)c.alloc.alloc(voff[ncls] * sizeof(double));
      auginfo = (rocblas_int*)c.alloc.alloc(sizeof(rocblas_int) * naug);
      launch_fill(c, aug, aoff[ncls], 0.0);
      std::vector<CopyOp> place;
      std::vector<AugJob> jobs;
      for (int k = 0; k < ncls; ++k) {
        const int Nk = cN[k];
        const int64_t NNk = (int64_t)Nk * Nk;
        for (int b = cls[k]; b < cls[k + 1]; ++b) {
          Sec& s = secs[big[b]];
          double* a = aug + aoff[k] + (b - cls[k]) * NNk;
          place.push_back({s.M, a + s.R * (int64_t)Nk, s.R, Nk, (int)s.R, (int)s.C, nullptr, nullptr, 1.0});
          jobs.push_back({a, augW + voff[k] + (int64_t)(b - cls[k]) * Nk, s.U, s.s, s.Vt, (int)s.R, (int)s.C, Nk,
                          s.kmin});
          is_small[big[b]] = 1;  // handled
          const double n2 = (double)Nk;
          svd_flops += 9.0 * n2 * n2 * n2;
        }
      }
      launch_copy(c, place);
      for (int k = 0; k < ncls; ++k) {
        const int Nk = cN[k];
        rocblas_status st = rocsolver_dsyevd_strided_batched(
            hb, rocblas_evect_original, rocblas_fill_upper, Nk, aug + aoff[k], Nk, (int64_t)Nk * Nk, augW + voff[k],
            Nk, augE + voff[k], Nk, auginfo + cls[k], cls[k + 1] - cls[k]);
        if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver_dsyevd_strided_batched failed");
      }
      launch_aug_extract(c, jobs);
    }
  }
  for (int si : order) {
    Sec& s = secs[si];
    if (s.kmin == 0 || is_small[si]) continue;
    rocblas_status st;
    if (c.svd_backend == 2 || c.svd_backend == 3)
      st = rocsolver_dgesvd(hb, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)s.R,
                            (rocblas_int)s.C, s.M, (rocblas_int)s.R, s.s, s.U, (rocblas_int)s.R, s.Vt,
                            (rocblas_int)s.kmin, s.E, rocblas_outofplace, info + nsvd);
    else  // divide-and-conquer SVD, one call per centre-charge sector
      st = rocsolver_dgesdd(hb, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)s.R,
                            (rocblas_int)s.C, s.M, (rocblas_int)s.R, s.s, s.U, (rocblas_int)s.R, s.Vt,
                            (rocblas_int)s.kmin, info + nsvd);
    if (st != rocblas_status_success) throw Error(DMRJ_ERR_SOLVER, "rocsolver SVD failed");
    ++nsvd;
    // nominal Golub-Kahan-Reinsch count with both singular-vector sets: 4 m n^2 + 8 n^3 (m >= n)
    const double mm = (double)std::max(s.R, s.C), nn = (double)std::min(s.R, s.C);
    svd_flops += 4.0 * mm * nn * nn + 8.0 * nn * nn * nn;
  }
  if (small_on_side) HIP_OK(hipStreamWaitEvent(c.stream, c.side_out, 0));  // small sectors done
  c.prof_end(psvd, KIND_SVD, svd_flops, 0.0);
  // singular values to host: one copy into the context's pinned buffer (the
  // entries of one-factor sectors are filled from their own results below)
  std::vector<double> hs(std::max<int64_t>(total_s, 1));
  double* hsig = c.host_sigma(std::max<int64_t>(total_s, 1));
  if (total_s > 0)
    HIP_OK(hipMemcpyAsync(hsig, sall, (size_t)total_s * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  std::vector<double> hgram((size_t)std::max<int64_t>(1, gramWtot));
  std::vector<rocblas_int> hginfo(std::max(1, ngram));
  if (ngram) {
    HIP_OK(hipMemcpyAsync(hgram.data(), gramW, (size_t)gramWtot * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream));
    HIP_OK(hipMemcpyAsync(hginfo.data(), graminfo, ngram * sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  }
  std::vector<rocblas_int> hinfo(std::max(1, nsvd));
  if (nsvd) HIP_OK(hipMemcpyAsync(hinfo.data(), info, nsvd * sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  std::vector<rocblas_int> hainfo(std::max(1, naug));
  if (naug)
    HIP_OK(hipMemcpyAsync(hainfo.data(), auginfo, naug * sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  std::vector<rocblas_int> hqinfo(std::max(1, nqr));
  if (nqr)
    HIP_OK(hipMemcpyAsync(hqinfo.data(), qinfo, nqr * sizeof(rocblas_int), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  if (total_s > 0) std::memcpy(hs.data(), hsig, (size_t)total_s * sizeof(double));
  c.alloc.release(aug);
  c.alloc.release(augW);
  c.alloc.release(augE);
  c.alloc.release(auginfo);
  for (double* p : {qrX, qrT, qaug, qW, qE, qVt}) c.alloc.release(p);
  c.alloc.release(qinfo);
  for (int i = 0; i < naug; ++i)
    if (hainfo[i] != 0) throw Error(DMRJ_ERR_SOLVER, "rocsolver symmetric eigensolver did not converge");
  for (int i = 0; i < nqr; ++i)
    if (hqinfo[i] != 0) throw Error(DMRJ_ERR_SOLVER, "rocsolver symmetric eigensolver (LQ/QR triangle) did not converge");
  c.release_transients();
  c.stage_reset();
  for (int i = 0; i < nsvd; ++i)
    if (hinfo[i] != 0) throw Error(DMRJ_ERR_SOLVER, "rocsolver SVD did not converge");
  for (int i = 0; i < ngram; ++i)
    if (hginfo[i] != 0) throw Error(DMRJ_ERR_SOLVER, "rocsolver_dsyevd (Gram) did not converge");
  {
    // one-factor sectors: singular values descending and the column of the
    // orthonormal factor belonging to each
    std::map<int, int> gram_of;
    for (int b = 0; b < ngram; ++b) gram_of[gramsec[b]] = b;
    int64_t o = 0;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.kmin == 0) continue;
      auto g = gram_of.find(si);
      if (g != gram_of.end()) {
        // Gram sector: eigenvalues ascending, sigma_k = sqrt(lambda_{N-1-k}), unit eigenvectors
        const int b = g->second, Nb = gramNb[b];
        for (int l = 0; l < s.kmin; ++l) {
          const double lam = hgram[(size_t)gramWoff[b] + Nb - 1 - l];
          hs[o + l] = std::sqrt(std::max(lam, 0.0));
          s.colsrc.push_back(gramZ + gramZoff[b] + (int64_t)(Nb - 1 - l) * Nb);
          s.coldiv.push_back(1.0);
        }
      } else if (s.qr) {
        // LQ / QR-reduced sector: sigma descending from the extraction, the
        // orthonormal factor's vectors are the columns of X (K x K)
        for (int l = 0; l < s.kmin; ++l) {
          s.colsrc.push_back(s.X + (int64_t)l * s.kmin);
          s.coldiv.push_back(1.0);
        }
      } else if (s.bj) {
        // block-Jacobi sector: column norms sorted descending (stable), the
        // largest kmin are the singular values
        s.perm.resize(s.ncolsX);
        std::iota(s.perm.begin(), s.perm.end(), 0);
        std::stable_sort(s.perm.begin(), s.perm.end(), [&](int a, int b) { return s.hn[a] > s.hn[b]; });
        for (int l = 0; l < s.kmin; ++l) {
          hs[o + l] = s.hn[s.perm[l]];
          s.colsrc.push_back(s.X + (int64_t)s.perm[l] * s.rowsX);
          s.coldiv.push_back(hs[o + l]);
        }
      }
      o += s.kmin;
    }
  }
  // global sort: sigma descending, ties by sector order then position
  struct Ent {
    double v;
    int sec, loc;
  };
  std::vector<Ent> ents;
  {
    int64_t o = 0;
    int rank = 0;
    for (int si : order) {
      Sec& s = secs[si];
      for (int l = 0; l < s.kmin; ++l) ents.push_back({hs[o + l], rank, l});
      o += s.kmin;
      ++rank;
    }
  }
  for (auto& e : ents)
    if (!std::isfinite(e.v)) throw Error(DMRJ_ERR_NUMERIC, "non-finite singular value");
  std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    if (a.v != b.v) return a.v > b.v;
    if (a.sec != b.sec) return a.sec < b.sec;
    return a.loc < b.loc;
  });
  const int64_t n = (int64_t)ents.size();
  int64_t keep = n;
  double acc = 0.0;
  while (keep > 0 && acc + ents[keep - 1].v * ents[keep - 1].v < cutoff) {
    acc += ents[keep - 1].v * ents[keep - 1].v;
    --keep;
  }
  if (m > 0) keep = std::min(keep, m);
  keep = std::max(keep, std::min(minm, n));
  SvdResult res;
  double terr = 0.0, kept2 = 0.0;
  for (int64_t i = keep; i < n; ++i) terr += ents[i].v * ents[i].v;
  for (int64_t i = 0; i < keep; ++i) kept2 += ents[i].v * ents[i].v;
  res.truncerr = terr;
  const double renorm = kept2 > 0 ? 1.0 / std::sqrt(kept2) : 1.0;
  // kept counts per sector (values within a sector are descending, so the
  // kept set of a sector is a prefix)
  std::vector<int> rank_to_sec;
  for (int si : order) rank_to_sec.push_back(si);
  for (int64_t i = 0; i < keep; ++i) secs[rank_to_sec[ents[i].sec]].keep++;
  for (int64_t i = 0; i < keep; ++i) res.svals.push_back(ents[i].v * renorm);
  // centre space
  Space mid;
  std::vector<double> scal;  // renormalised kept sigma per sector, concatenated
  std::vector<int64_t> soff(secs.size(), 0);
  {
    int64_t hoff = 0;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.keep > 0) {
        mid.q.push_back(s.Q);
        mid.dim.push_back(s.keep);
        soff[si] = (int64_t)scal.size();
        for (int l = 0; l < s.keep; ++l) scal.push_back(hs[hoff + l] * renorm);
      }
      hoff += s.kmin;
    }
  }
  mid.finish();
  {
    // block-Jacobi sectors: normalised kept columns, then the other factor
    // U^T M = D V^T (dir > 0) or M V = U D (dir < 0) in one grouped GEMM
    std::vector<BjGather> gat;
    std::vector<GemmTask> tasks;
    std::vector<GemmSeg> segs;
    int64_t hoff = 0;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.bj && s.keep > 0) {
        for (int k = 0; k < s.keep; ++k) {
          const double* src = s.colsrc[k];
          if (dir > 0)
            gat.push_back({src, s.U + (int64_t)k * s.R, 1, s.coldiv[k], s.rowsX, 0});
          else
            gat.push_back({src, s.Vt + k, s.kmin, s.coldiv[k], s.rowsX, 1});
        }
        if (dir > 0) {
          tasks.push_back({s.Vt, s.kmin, s.keep, (int)s.C, (int)segs.size(), 1, 0.0});
          segs.push_back({s.U, s.M, s.R, s.R, (int)s.R, 1, 1.0});
        } else {
          tasks.push_back({s.U, s.R, (int)s.R, s.keep, (int)segs.size(), 1, 0.0});
          segs.push_back({s.M, s.Vt, s.R, s.kmin, (int)s.C, 2, 1.0});
        }
      }
      hoff += s.kmin;
    }
    launch_bj_gather(c, gat);
    // canonical factor: kept vectors of singular values below 1e-2 sigma_max
    // (extracted vectors are orthogonal to the rest only to about
    // eps sigma_max / sigma, so those above stay within ~1e-14; not in Gram
    // sectors, whose factor is a set of eigenvectors) re-orthonormalised in
    // place, then an orthonormal
    // completion for numerically zero singular values (sigma <= 1e-10
    // sigma_max); the latter carry no weight
    double smax = 0.0;
    for (int64_t i = 0; i < n; ++i) smax = std::max(smax, ents[i].v);
    const double szero = 1e-10 * smax;
    const double ssmall = 1e-2 * smax;
    std::vector<CompleteJob> cj, rj;
    hoff = 0;
    for (int si : order) {
      Sec& s = secs[si];
      if (s.keep > 0) {
        int c0 = s.keep;
        while (c0 > 0 && hs[hoff + c0 - 1] <= szero) --c0;
        int c1 = c0;
        if (!s.gram)  // Gram eigenvectors need no re-orthonormalisation
          while (c1 > 1 && hs[hoff + c1 - 1] < ssmall) --c1;
        const int nv = (int)(dir > 0 ? s.R : s.C);
        if (c1 < c0 && c1 >= 1) {
          if (dir > 0)
            rj.push_back({s.U, s.R, 1, nv, c1, c0});
          else
            rj.push_back({s.Vt, 1, s.kmin, nv, c1, c0});
        }
        if (c0 < s.keep) {
          if (dir > 0)
            cj.push_back({s.U, s.R, 1, nv, c0, s.keep});
          else
            cj.push_back({s.Vt, 1, s.kmin, nv, c0, s.keep});
        }
      }
      hoff += s.kmin;
    }
    if (!z) {
      reorth_basis(c, rj);
      complete_basis(c, cj);
    }
    launch_gemm(c, tasks, segs, -1);
  }
  double* dscal = (double*)c.alloc.alloc(std::max<size_t>(1, scal.size()) * sizeof(double));
  HIP_OK(hipMemcpyAsync(dscal, scal.data(), std::max<size_t>(1, scal.size()) * sizeof(double),
                        hipMemcpyHostToDevice, c.stream));
  auto* tl = new_tensor(c, 0);
  auto* tr = new_tensor(c, 0);
  tl->nq = tr->nq = AA.nq;
  tl->site.left = S;
  tl->site.right = mid;
  tl->site.P = d1;
  tl->site.qp = x.qp1;
  tl->site.build();
  tr->site.left = mid;
  tr->site.right = T;
  tr->site.P = d2;
  tr->site.qp = x.qp2;
  tr->site.build();
  tl->cplx = tr->cplx = z ? 1 : 0;
  alloc_data(c, *tl);
  alloc_data(c, *tr);
  std::vector<CopyOp> split;
  for (const SiteBlock& b : tl->site.blocks) {
    const Charge Q = S.q[b.a] + x.qp1[b.s];
    Sec& s = secs[qidx[Q]];
    int64_t ro = -1;
    for (size_t r = 0; r < s.rows.size(); ++r)
      if (s.rows[r].first == b.a && s.rows[r].second == b.s) ro = s.roff[r];
    if (s.bj)  // U D from the GEMM (dir < 0) or orthonormal U (dir > 0)
      split.push_back({s.U + ro, tl->data + b.off, s.R, S.dim[b.a], S.dim[b.a], s.keep, nullptr, nullptr,
                       dir < 0 ? renorm : 1.0});
    else
      split.push_back({s.U + ro, tl->data + b.off, s.R, S.dim[b.a], S.dim[b.a], s.keep, nullptr,
                       dir < 0 ? dscal + soff[qidx[Q]] : nullptr, 1.0});
    if (z)
      split.push_back({s.U + s.R * s.kmin + ro, tl->data + tl->site.size + b.off, s.R, S.dim[b.a], S.dim[b.a],
                       s.keep, nullptr, dir < 0 ? dscal + soff[qidx[Q]] : nullptr, 1.0});
  }
  for (const SiteBlock& b : tr->site.blocks) {
    const Charge Q = mid.q[b.a];
    Sec& s = secs[qidx[Q]];
    int64_t co = -1;
    for (size_t q = 0; q < s.cols.size(); ++q)
      if (s.cols[q].first == b.s && s.cols[q].second == b.b) co = s.coff[q];
    if (s.bj)  // D V^T from the GEMM (dir > 0) or orthonormal V^T (dir < 0)
      split.push_back({s.Vt + co * s.kmin, tr->data + b.off, s.kmin, s.keep, s.keep, T.dim[b.b], nullptr, nullptr,
                       dir > 0 ? renorm : 1.0});
    else
      split.push_back({s.Vt + co * s.kmin, tr->data + b.off, s.kmin, s.keep, s.keep, T.dim[b.b],
                       dir > 0 ? dscal + soff[qidx[Q]] : nullptr, nullptr, 1.0});
    if (z)
      split.push_back({s.Vt + (int64_t)s.kmin * s.C + co * s.kmin, tr->data + tr->site.size + b.off, s.kmin, s.keep,
                       s.keep, T.dim[b.b], dir > 0 ? dscal + soff[qidx[Q]] : nullptr, nullptr, 1.0});
  }
  launch_copy(c, split);
  for (auto& s : secs) {
    c.alloc.release(s.M);
    c.alloc.release(s.U);
    c.alloc.release(s.Vt);
    c.alloc.release(s.E);
    c.alloc.release(s.X);
  }
  c.alloc.release(sall);
  c.alloc.release(bjflag);
  c.alloc.release(bjnorm);
  c.alloc.release(gramZ);
  c.alloc.release(gramW);
  c.alloc.release(gramE);
  c.alloc.release(graminfo);
  c.alloc.release(info);
  c.alloc.release(dscal);
  res.L = tl;
  res.R = tr;
  return res;
}

}  // namespace dmrj
