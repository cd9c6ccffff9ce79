// Internal types of libdmrjhip: graded bond spaces, block-sparse tensor
// metadata, contraction plans and the per-context device runtime.
//
// Block-sparse layout (shared by dense and quantum-number runs; a dense bond
// is a space with one sector of charge 0):
//   * site tensor A[alpha, s, beta] with P physical states: one block per
//     (left sector a, physical state s) whose right sector b has charge
//     q(a) + q(s).  Blocks are stored as column-major n_a x n_b matrices,
//     grouped in "row panels": for fixed a all s-blocks are adjacent, so the
//     panel is itself an n_a x C_a column-major matrix (ld = n_a).
//   * environment E[bra, w, ket] (left) or E[bra, u, ket] (right, stored
//     transposed with respect to the reference's Renv[ket, u, bra]): one
//     block per (bra sector a', MPO channel w) whose ket sector a has charge
//     q(a') - delta(w); blocks of one bra sector are adjacent ("bra panel",
//     an n_a' x sum_w n_a matrix, ld = n_a').
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct dmrj_tensor;  // the C ABI's tensor handle (a dmrj::Tensor)

namespace dmrj {

constexpr int kMaxQ = 2;

struct Charge {
  int32_t v[kMaxQ] = {0, 0};
  bool operator<(const Charge& o) const {
    for (int i = 0; i < kMaxQ; ++i)
      if (v[i] != o.v[i]) return v[i] < o.v[i];
    return false;
  }
  bool operator==(const Charge& o) const {
    for (int i = 0; i < kMaxQ; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
  bool operator!=(const Charge& o) const { return !(*this == o); }
  Charge operator+(const Charge& o) const {
    Charge r;
    for (int i = 0; i < kMaxQ; ++i) r.v[i] = v[i] + o.v[i];
    return r;
  }
  Charge operator-(const Charge& o) const {
    Charge r;
    for (int i = 0; i < kMaxQ; ++i) r.v[i] = v[i] - o.v[i];
    return r;
  }
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Graded vector space of a bond: sectors sorted by charge.
struct Space {
  std::vector<Charge> q;
  std::vector<int> dim;
  std::vector<int64_t> off;  // offset of the sector in the sorted basis
  int64_t total = 0;
  int nsec() const { return (int)q.size(); }
  int find(const Charge& c) const;
  void finish();  // fills off/total
  bool same(const Space& o) const;
};

struct SiteBlock {
  int a, s, b;
  int64_t off;      // element offset in the tensor buffer
  int64_t colpos;   // first column of the block inside its row panel
};

struct SiteMeta {
  Space left, right;
  int P = 0;
  std::vector<Charge> qp;          // charge of each physical state
  int d1 = 0, d2 = 0;              // two-site tensor: P = d1*d2, s = j1 + d1*j2
  std::vector<Charge> qp1, qp2;
  std::vector<SiteBlock> blocks;   // ordered by (a, s)
  std::vector<int> bidx;           // a*P + s -> block index or -1
  std::vector<int64_t> panel_off;  // per left sector
  std::vector<int64_t> panel_cols; // per left sector
  int64_t size = 0;
  void build();
  int block(int a, int s) const { return bidx[(size_t)a * P + s]; }
};

struct EnvBlock {
  int ap, w, a;     // bra sector, channel, ket sector
  int64_t off;
  int64_t colpos;   // column offset inside the bra panel
};

struct EnvMeta {
  Space bra, ket;
  int D = 0;
  std::vector<Charge> delta;       // per channel
  std::vector<EnvBlock> blocks;    // ordered by (ap, w)
  std::vector<int> bidx;           // ap*D + w -> block index or -1
  std::vector<int64_t> panel_off;  // per bra sector
  std::vector<int64_t> panel_cols; // per bra sector
  int64_t size = 0;
  void build();
  int block(int ap, int w) const { return bidx[(size_t)ap * D + w]; }
};

// Grouped GEMM descriptors (device side).  C[MxN] = beta*C + sum_seg alpha *
// op(A_seg)[MxK] * op(B_seg)[KxN]; flags bit0: A stored KxM (transposed),
// bit1: B stored NxK (transposed).
struct GemmSeg {
  const double* A;
  const double* B;
  int64_t lda, ldb;
  int K;
  int flags;
  double alpha;
};
struct GemmTask {
  double* C;
  int64_t ldc;
  int M, N;
  int seg0, nseg;
  double beta;
};
struct GemmTile {
  int task, m0, n0, cfg;
};

struct MixTerm {
  const double* src;
  double coef;
};
struct MixDst {
  double* dst;
  int64_t n;
  int t0, nt;
};
struct MixWork {
  int dst;
  int pad;
  int64_t e0;
};

struct CopyOp {
  const double* src;   // null -> write zeros
  double* dst;
  int64_t lds, ldd;
  int rows, cols;
  const double* rscale;  // optional per-row factor
  const double* cscale;  // optional per-column factor
  double alpha;
};
struct CopyWork {
  int op;
  int pad;
  int64_t e0;
};

// ---- runtime ----
struct Allocator {
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> live_;
  size_t bytes_live = 0, bytes_cached = 0, peak = 0;
  void* alloc(size_t bytes);
  void release(void* p);
  void trim();
  ~Allocator();
};

struct HostStage {   // pinned staging memory for plan uploads
  char* base = nullptr;
  size_t cap = 0, used = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) for work that overlaps the main one:
  // the small-sector SVD kernel runs beside the batched eigensolve of the
  // large sectors; side_in / side_out order it against the main stream
  hipStream_t side = nullptr;
  hipEvent_t side_in = nullptr, side_out = nullptr;
  void* rocblas = nullptr;   // rocblas_handle
  // batching of the SVD's batched symmetric eigensolves: 1 = size classes by
  // the cost model, 0 = one batch padded to the largest sector, 2 = one batch
  // per sector (dmrj_ctx_set_eig_batching)
  int eig_batching = 1;
  Allocator alloc;
  HostStage stage;
  std::vector<void*> transient;  // device copies of plan tables, freed at sync points
  std::string err;
  double* scal = nullptr;       // device scalars for Lanczos
  double* hscal = nullptr;      // pinned mirror
  double* hsig = nullptr;       // pinned buffer for singular values coming to the host
  int64_t hsig_cap = 0;
  // the pinned singular-value buffer with room for n doubles (grown when
  // needed; only after a synchronisation of the stream that used it)
  double* host_sigma(int64_t n) {
    if (n > hsig_cap) {
      if (hsig) {
        if (hipStreamSynchronize(stream) != hipSuccess || hipHostFree(hsig) != hipSuccess)
          throw Error(-3, "HIP: pinned singular-value buffer");
        hsig = nullptr;
      }
      const int64_t cap = std::max<int64_t>(n, 2 * hsig_cap);
      if (hipHostMalloc((void**)&hsig, (size_t)cap * sizeof(double), hipHostMallocDefault) != hipSuccess)
        throw Error(-3, "HIP: pinned singular-value buffer");
      hsig_cap = cap;
    }
    return hsig;
  }
  double* partial = nullptr;    // reduction partials
  double* noalloc = nullptr;    // device address of dmrj_dots_noalloc's scratch on this device
  int num_cu = 256;
  size_t lds_max = 65536;  // dynamic LDS per work-group (device property)
  // tuning: preferred GEMM tile config per operation (-1 = automatic)
  int gemm_cfg_override = -1;
  // SVD backend: 0 = batched Jacobi for small sectors + batched symmetric
  // eigensolver of the augmented matrix for large ones, 1 = gesdd,
  // 2 = gesvd, 3 = batched Jacobi for small sectors + gesvd
  int svd_backend = 0;
  double bj_calls = 0, bj_sweeps = 0;  // block-Jacobi SVD calls and sweeps
  int env_prec = 0;  // precision of environment updates
  int last_gemm_cfg = -1;
  // grouped-GEMM autotuner: 1 = time the tile shapes per histogram key
  int autotune = 1;
  int xcd_order = 1;  // 1: grouped-GEMM tiles dealt to XCDs by shared panel
  // 1 = each Lanczos step (both H_eff applications, inner products and the
  // Ritz update) is captured into one hipGraph and launched as a unit
  int graph_mode = 0;
  // executable graphs by kernel sequence: 0 = Lanczos step, 1 = two-site
  // Krylov exponential, 2 = one-site Krylov exponential
  static constexpr int kGraphSlots = 3;
  hipGraphExec_t graph_exec[kGraphSlots] = {};
  hipEvent_t graph_done[kGraphSlots] = {};
  double graph_updates = 0, graph_instantiations = 0;
  std::unordered_map<uint64_t, int> tune_cache;
  int occ_cache[4][8] = {{0}};
  int f16_presplit = 1;  // FP16x3 GEMMs on pre-split operand planes (dmrj_ctx_set_f16_presplit)
  double tune_hits = 0, tune_misses = 0, tune_ms = 0, tune_skipped = 0;
  // statistics
  double flops_heff = 0.0, flops_env = 0.0;
  // per-kernel-kind HIP-event timing on this context's stream
  bool prof = false;
  struct ProfRec {
    hipEvent_t a, b;
    int kind;
    double flops, bytes;
  };
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  double prof_ms[8] = {0}, prof_flops[8] = {0}, prof_bytes[8] = {0}, prof_count[8] = {0};
  int prof_begin();                                // returns slot or -1
  void prof_end(int slot, int kind, double flops, double bytes);
  // Timing of one kernel by its own dispatch: prof_begin_ext() sets ext_a /
  // ext_b, which the next launch helper passes to hipExtLaunchKernelGGL
  // (start and stop stamped by that dispatch, as a kernel trace records it);
  // prof_end_ext() files the record and clears them.
  hipEvent_t ext_a = nullptr, ext_b = nullptr;
  int prof_begin_ext();
  void prof_end_ext(int slot, int kind, double flops, double bytes);
  void prof_resolve();                             // after a stream sync
  void* upload(const void* host, size_t bytes);   // staged async copy, returns device ptr
  void stage_reset();
  void release_transients();
};

struct MpoSite {
  int DL = 0, d = 0, DR = 0;
  std::vector<double> W;  // column-major W[w + DL*(j + d*(k + d*u))]
  std::vector<Charge> qp, dL, dR;
  double at(int w, int j, int k, int u) const {
    return W[(size_t)w + (size_t)DL * ((size_t)j + (size_t)d * ((size_t)k + (size_t)d * u))];
  }
};

struct Tensor {
  int kind = 0;  // 0 = site tensor, 1 = left env, 2 = right env
  int nq = 0;
  int cplx = 0;  // 1: ComplexF64 held as two planes, real part then imaginary part
  SiteMeta site;
  EnvMeta env;
  double* data = nullptr;
  int64_t nelem() const { return kind == 0 ? site.size : env.size; }   // elements per plane
  int64_t nstore() const { return nelem() * (cplx ? 2 : 1); }         // doubles stored
};

// ---- ComplexF64 on split planes (complex.hip) ----
// A complex operand is stored as its real plane followed by its imaginary
// plane.  A plan built for real data (pointers into real planes) is turned
// into the complex plan by expanding every product into its real products;
// operands that are real contribute no imaginary part.  GemmSeg flag bits
// kSegConjA / kSegConjB mark operands that enter complex-conjugated (the bra
// of an environment update); the GEMM kernels ignore them.
constexpr int kSegConjA = 4, kSegConjB = 8;
struct BufRef {
  const double* base;
  int64_t n;  // elements per plane
  bool cplx;
};
struct BufMap {
  std::vector<BufRef> b;
  void add(const double* base, int64_t n, bool cplx) { b.push_back({base, n < 1 ? 1 : n, cplx}); }
  const BufRef* find(const double* p) const;
};
void complexify_gemm(std::vector<GemmTask>& tasks, std::vector<GemmSeg>& segs, const BufMap& m);
void complexify_mix(std::vector<MixDst>& dsts, std::vector<MixTerm>& terms, const BufMap& m);
struct AugJobZ {
  const double* M;   // R x C, two planes, ld R
  void* Z;           // N x N complex (interleaved), ld N
  const double* W;   // eigenvalues, ascending
  double* U;         // R x kmin, two planes, ld R
  double* S;
  double* Vt;        // kmin x C, two planes, ld kmin
  int R, C, N, kmin;
};
void launch_aug_place_z(Context& c, const std::vector<AugJobZ>& jobs);

// ---- batched block one-sided Jacobi for large SVD sectors (svd_jacobi.hip) ----
struct BjJob {
  double* X;         // rows x ncols working matrix, ld rows, orthogonalised in place
  double* S;         // ncols column norms (output of launch_bj_norms)
  int* flag;         // set to 1 by a round that rotated
  const int* perm;   // sweep order of the columns: block b holds perm[b bw .. b bw + bw)
  double tiny2;      // columns with |x|^2 below this are numerically zero (rank deficiency)
  int rows, ncols, nblk, bw;  // nblk even (an empty block pads an odd count)
};
struct BjGather {
  const double* src;  // column of X
  double* dst;
  int64_t ldd;        // trans: stride between consecutive elements of dst
  double sigma;       // divisor
  int rows, trans;
};
size_t bj_lds_bytes(int rows, int bw);
// static shared memory of k_bj_step (from its compiled attributes), reserved
// out of the per-work-group LDS limit (lds_max counts static and dynamic LDS
// together)
size_t bj_static_lds();
// column-block width of the block-Jacobi panel (rows x 2bw doubles in
// dynamic LDS) for `rows` rows: 16, 8 or 4, or 0 when even 4 does not fit
int bj_panel_width(int rows, size_t lds_max, size_t static_lds);
void launch_bj_sweep(Context& c, const BjJob* djobs, const std::vector<int2>& work, int nsteps, int inner_max,
                     size_t lds);
void launch_bj_norms(Context& c, const BjJob* djobs, const std::vector<int2>& work);
void launch_bj_gather(Context& c, const std::vector<BjGather>& g);
void launch_transpose(Context& c, const double* M, int R, int C, double* X);
void launch_pad_diag(Context& c, double* Z, int N, const std::vector<int2>& jobs, double v);
struct TriJob {
  const double* X;  // LQ / QR output, leading dimension ldx
  double* A;        // N x N augmented matrix of this sector
  int64_t ldx;
  int K, N, trans;  // trans: B = R^T of a QR (else B = L of an LQ)
};
void launch_tri_place(Context& c, const std::vector<TriJob>& jobs);
struct CompleteJob {
  double* Q;          // canonical factor: element r of vector c at Q[c vs + r es]
  int64_t vs, es;
  int n, c0, keep;    // vectors c0 .. keep-1 are replaced
  // scratch, set by complete_basis
  double* W = nullptr;     // n x 32 candidates
  double* coef = nullptr;  // keep x 32
  double* d = nullptr;     // n residuals
  double* seld = nullptr;  // 32
  int* cand = nullptr;     // 32
};
// A new tensor handle of the given kind (0 site, 1 left env, 2 right env) and
// its device storage for the current block table (api.hip).
::dmrj_tensor* new_tensor(Context& c, int kind);
void alloc_data(Context& c, Tensor& t);
// Truncated two-site SVD (svd.hip): left / right tensors of the kept
// singular vectors (the centre's weights on the side dir points to).
struct SvdResult {
  ::dmrj_tensor* L = nullptr;
  ::dmrj_tensor* R = nullptr;
  double truncerr = 0.0;
  std::vector<double> svals;  // kept, descending
};
SvdResult svd_trunc(Context& c, const Tensor& AA, int64_t m, int64_t minm, double cutoff, int dir);
// Orthonormal completion (svd_jacobi.hip); allocates and releases its scratch.
void complete_basis(Context& c, std::vector<CompleteJob>& jobs);
// Re-orthonormalisation of the kept vectors c0 .. keep-1 against the vectors
// before them and among themselves (svd_jacobi.hip); c0 >= 1.
void reorth_basis(Context& c, std::vector<CompleteJob>& jobs);
void launch_aug_extract_z(Context& c, const std::vector<AugJobZ>& jobs);

// ---- kernel launchers (kernels.hip) ----
void launch_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs,
                 int cfg);
void launch_mix(Context& c, const std::vector<MixDst>& dsts, const std::vector<MixTerm>& terms);
void launch_copy(Context& c, const std::vector<CopyOp>& ops);
void launch_fill(Context& c, double* p, int64_t n, double v);
void launch_marker(Context& c, int id);
// One block classical Gram-Schmidt pass: w -= sum_{i<k} <v_i, w> v_i (v_i at
// V + i nv; complex: split planes of n elements).  partial: 2 k 256 doubles,
// coef: 2 k doubles.
void launch_cgs_pass(Context& c, double* w, const double* V, int64_t nv, int64_t n, int k, int cplx,
                     double* partial, double* coef);
// One Krylov iteration after the H_eff application in a single work-group
// (alpha_j to sc[j]; unless last: two Gram-Schmidt passes against v_0..v_j,
// beta_j to sc[64 + j], v_{j+1} to V and xin) for short vectors.
void launch_krylov_tail(Context& c, double* V, int64_t nv, int j, double* w, int64_t n, int cplx, double* sc,
                        double* xin, int last);
constexpr int64_t kKrylovFusedMax = 16384;  // stored elements of a Krylov vector for the one-work-group path
constexpr int kDotPartials = 256;  // per-pair partial sums of launch_dots
// launch_dots with the partials in `partial` (npairs * kDotPartials doubles)
void launch_dots_buf(Context& c, int npairs, const double* const* xs, const double* const* ys, int64_t n,
                     double* out, double* partial);
void launch_dots(Context& c, int npairs, const double* const* xs, const double* const* ys,
                 int64_t n, double* out);
void launch_axpy_dev(Context& c, double* y, const double* x, int64_t n, const double* coef);
void launch_scale_dev(Context& c, double* x, int64_t n, const double* coef);
void launch_comb_dev(Context& c, double* out, const double* a, const double* b, int64_t n,
                     const double* ca, const double* cb);
void launch_lz_scalar(Context& c, int stage);
void launch_combine(Context& c, double* out, const double* V, int64_t ld, int64_t n, const double* coef, int k);
// exp(pf T) e1 (or the lowest Ritz vector) coefficients of a Krylov space on
// the device: sc = {alpha[0..64), beta[64..128)}; coef receives 2r (real
// Krylov vectors) or 4r (complex) weights, outs {lowest Ritz value, k, flag}
void launch_krylov_coef(Context& c, const double* sc, int r, double pf_re, double pf_im, int ritz, int zv,
                        double* coef, double* outs);
void launch_mfma_probe(Context& c, double* out);
void launch_mfma_bf16_probe(Context& c, float* out);
void launch_gemm_lp(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int cfg);
int choose_gemm_cfg(const std::vector<GemmTask>& tasks, int num_cu);

// ---- grouped-GEMM tile shapes and the autotuner (autotune.hip) ----
constexpr int kNumGemmShapes = 5;
struct TileShape {
  int bm, bn;
};
extern const TileShape kGemmShapes[kNumGemmShapes];
// Descriptor tables of one grouped-GEMM launch, already on the device.
// arith: 0 = FP64 MFMA, 1 = BF16x3 split, 2 = FP16x3 split with per-segment
// power-of-two scaling (the values of DMRJ_PREC_F64 / _BF16X3 / _F16X3).
// FP16x3 operand planes (Context::f16_presplit): every distinct operand
// region of a launch is scaled by its segment's power of two and split once
// into FP16 hi / lo planes, rows x kpad each, k contiguous (kpad = K rounded
// up to 32, zero-filled), so the GEMM reads MFMA fragments directly.
struct F16Region {
  const double* src;
  int64_t ld;
  int rows, K, kpad;
  int tr;        // 1: element (r, k) at src[k + r ld]; 0: at src[r + k ld]
  int operand;   // 0 A, 1 B
  double alpha;  // folded into A (1 for B)
  uint16_t* hi;  // FP16 bit patterns
  uint16_t* lo;
};
struct F16Seg {
  const uint16_t* ah;
  const uint16_t* al;
  const uint16_t* bh;
  const uint16_t* bl;
  int kpad;
  int ra, rb;  // regions of A and B (their maxima scale the segment)
  int pad;
};
struct GemmLaunch {
  const GemmTile* tiles = nullptr;
  const GemmTask* tasks = nullptr;
  const GemmSeg* segs = nullptr;
  int ntiles = 0, grid = 0, cfg = 0, arith = 0, nseg = 0;
  double* segmax = nullptr;          // FP16x3: 2 per segment, max |alpha A| and max |B|
  const int4* maxwork = nullptr;     // FP16x3: work items of the max pre-pass
  const int* segtask = nullptr;      // FP16x3: task of every segment
  int nmaxwork = 0;
  bool presplit = false;             // FP16x3 with operand planes
  const F16Region* regions = nullptr;
  const int4* splitwork = nullptr;   // (region, first row, first k, 0)
  int nsplitwork = 0;
  const F16Seg* f16segs = nullptr;
  double* regionmax = nullptr;       // max |alpha A| or max |B| of every region
  const int4* regionmaxwork = nullptr;  // (region, first stored column, end column, 0)
  int nregionmaxwork = 0, nregions = 0;
};
GemmLaunch prepare_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs,
                        int cfg, int arith);
void issue_gemm_f64(Context& c, const GemmLaunch& g);
void issue_gemm_lp(Context& c, const GemmLaunch& g);
void issue_gemm_f16(Context& c, const GemmLaunch& g);
// FP16x3 in two parts (plane path): the operand maxima and the plane split,
// then the GEMM kernel alone (timed by its own dispatch when ext events are set)
void issue_gemm_f16_prep(Context& c, const GemmLaunch& g);
void issue_gemm_f16_main(Context& c, const GemmLaunch& g);
void issue_gemm(Context& c, const GemmLaunch& g);
int gemm_occupancy_f64(int cfg);
int gemm_occupancy_lp(int cfg);
int gemm_occupancy_f16(int cfg);
int gemm_occupancy_f16p(int cfg);
double gemm_flops(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs);
double gemm_bytes(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs);
// Tile shape for a task list: cached per block-size-histogram key; on a
// miss every shape is timed on the live operands (only for launches that are
// safe to repeat: beta = 0 and outputs disjoint from inputs).
int autotune_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int arith);
// override > autotuner > heuristic
int pick_gemm_cfg(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int arith);
// one grouped GEMM in the given arithmetic (tile shape picked when cfg < 0)
void launch_gemm_arith(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int cfg,
                       int arith);
// Descriptor tables of one mix launch, already on the device.
struct MixLaunch {
  const MixWork* work = nullptr;
  const MixDst* dsts = nullptr;
  const MixTerm* terms = nullptr;
  int nwork = 0, grid = 0;
  double flops = 0.0, bytes = 0.0;
};
MixLaunch prepare_mix(Context& c, const std::vector<MixDst>& dsts, const std::vector<MixTerm>& terms);
void issue_mix(Context& c, const MixLaunch& m);
struct SmallSvdJob {
  const double* M;
  double* U;
  double* S;
  double* Vt;
  int R, C;
};
// runs on c.side; join with hipStreamWaitEvent(c.stream, c.side_out) when it returns true
bool launch_svd_small(Context& c, const std::vector<SmallSvdJob>& jobs, size_t lds_bytes);
struct AugJob {
  const double* A;
  const double* W;
  double* U;
  double* S;
  double* Vt;
  int R, C, N, kmin;
};
void launch_aug_extract(Context& c, const std::vector<AugJob>& jobs);
size_t small_svd_lds(int R, int C);
constexpr size_t kSmallSvdMaxLds = 60 * 1024;  // bytes of dynamic LDS per sector
double gemm_flops(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs);

// kernel kinds for Context profiling
enum KernelKind { KIND_GEMM_F64 = 0, KIND_MIX = 1, KIND_COPY = 2, KIND_SVD = 3, KIND_GEMM_LP = 4, KIND_GRAPH = 5,
                  KIND_GEMM_F16 = 6, KIND_F16_PREP = 7 };

// Lanczos scalar slots in Context::scal
enum Slot {
  S_TT = 0, S_TH = 1, S_PP = 2, S_PHP = 3, S_OC = 4,
  S_NRM = 8, S_A1 = 9, S_CAXPY = 10, S_B1 = 11, S_INVB = 12, S_FLAG = 13,
  S_A2 = 14, S_E = 15, S_CA = 16, S_CB = 17, S_INVOC = 18, S_CHK = 19,
  S_RR = 20, S_RH = 21, S_NEGE = 22, S_RES = 23, S_ELP = 24,
  S_COUNT = 32
};

// Load through a global-address-space pointer: a global_load counts only
// against vmcnt, so the prefetch of a GEMM's next K step stays in flight
// across the LDS reads and MFMAs of the current one (a generic flat_load would
// also hold every lgkmcnt wait of those LDS reads).
__device__ __forceinline__ double ld_global(const double* p) {
  return *(const __attribute__((address_space(1))) double*)p;
}

// The host arrays of one launch (tile list, tasks, segments, work lists,
// FP16 plane tables) packed into one blob, each at a 256-byte aligned offset,
// so that they reach the device in a single staged copy instead of one per
// array.  add() returns an array's byte offset in the blob; commit() uploads
// the blob once and returns its device address.
struct UploadPack {
  std::vector<char> blob;
  size_t add(const void* p, size_t bytes) {
    const size_t o = blob.size();
    blob.resize(o + std::max<size_t>(256, (bytes + 255) / 256 * 256), 0);
    if (p && bytes) std::memcpy(blob.data() + o, p, bytes);
    return o;
  }
  char* commit(Context& c) { return (char*)c.upload(blob.data(), blob.size()); }
};

}  // namespace dmrj

// The C ABI's opaque handles are the library's own objects.
struct dmrj_ctx : dmrj::Context {};
struct dmrj_tensor : dmrj::Tensor {};
struct dmrj_mpo : dmrj::MpoSite {};
