// Composite hot-path operations on device tensors (C++ layer under the C ABI).
#pragma once

#include <vector>

#include "engine.hpp"

namespace mpx {

struct OpStats {
  int64_t flops = 0;   // algorithmic flops of the last plan executed (reference contraction order)
  int64_t bytes = 0;   // algorithmic bytes (inputs + output, FP64 storage)
};

// Last plan statistics (per thread).
OpStats& last_stats();

// The plan and input pointers of the last hot-path call on this thread (for per-kernel profiling; the plan stays
// valid while it is cached, the inputs while the caller keeps the tensors).
// The most recent hot-path call on this thread: the cache key of its plan and its input tensors with the data
// pointer and structure each had then.  last_plan() looks the plan up again and checks that every input still
// exists unchanged (error otherwise), so a plan evicted since, or a tensor freed or reshaped since, is never used.
struct LastRun {
  std::vector<uint64_t> key;
  uint64_t plan_id = 0;  // identity of the cached plan the call ran
  std::vector<const Tensor*> tensors;
  std::vector<const double*> inputs;
  std::vector<uint64_t> hashes;
};
LastRun& last_run();
void record_last_run(const std::vector<uint64_t>& key, std::initializer_list<const Tensor*> ins);
void record_last_run(const std::vector<uint64_t>& key, const std::vector<Tensor*>& ins);
// The plan of last_run() and its current input pointers; throws MPSXX_ERR_ARG if there is none or it went stale.
Plan& last_plan(std::vector<const double*>& inputs);

// Generic BTAS machinery
void op_contract(double alpha, Tensor& a, bool conj_a, const std::vector<int>& ca, Tensor& b, bool conj_b,
                 const std::vector<int>& cb, double beta, Tensor& c, hipStream_t s);
void op_gemm(int ta, int tb, double alpha, Tensor& a, bool conj_a, Tensor& b, bool conj_b, double beta, Tensor& c,
             hipStream_t s);
void op_tie(Tensor& x, const std::vector<int>& tie, Tensor& y, hipStream_t s);
void op_copy_upcast(Tensor& x, Tensor& y, hipStream_t s);

// BLAS-1
double op_dotc(Tensor& x, Tensor& y, hipStream_t s);
// The same dot product with its segment list, chunk list, partial sums and result in the device's static scratch
// array (noalloc.hip) instead of library-allocated buffers: no device allocation.  Calls that overlap on one device
// race on that array.  Throws MPSXX_ERR_ARG when the lists do not fit it.
double op_dotc_noalloc(Tensor& x, Tensor& y, hipStream_t s);
double op_nrm2(Tensor& x, hipStream_t s);
void op_axpy(double alpha, Tensor& x, Tensor& y, hipStream_t s);
void op_scal(double alpha, Tensor& x, hipStream_t s);
void op_copy(Tensor& x, Tensor& y, hipStream_t s);

// Hot path (reference: src/compute_sigma_vector.hpp, compute_diagonal_elements.hpp, renormalize.hpp)
// convert_mask (reduced-precision plans): bit i set = refresh the FP32 copy of input i (in argument order); a clear bit
// reuses the copy made by the previous call of the same plan
void op_sigma_twosite(Tensor& lmpo, Tensor& rmpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec,
                      hipStream_t s, unsigned convert_mask = ~0u);
void op_sigma_onesite(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec, hipStream_t s);
void op_sigma_merged(Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec, hipStream_t s,
                     unsigned convert_mask = ~0u);
void op_sigma_overlap(Tensor& lsov, Tensor& rsov, Tensor& wfn, Tensor& sgv, hipStream_t s);
void op_sigma_conj(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, hipStream_t s);
void op_diag_twosite(Tensor& lmpo, Tensor& rmpo, Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s);
void op_diag_onesite(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s);
void op_diag_merged(Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s);
void op_renormalize(bool fwd, Tensor& mpo, Tensor& opr0, Tensor& bra, Tensor& ket, Tensor& opr1, int prec,
                    hipStream_t s);
void op_renormalize_overlap(bool fwd, Tensor& sov0, Tensor& bra, Tensor& ket, Tensor& sov1, hipStream_t s);
void op_renormalize_merged(bool fwd, Tensor& opr0, Tensor& bra, Tensor& ket, Tensor& opr1, int prec,
                           hipStream_t s);

// Plan cache counters since the process started: lookups served from the cache, plans built (and the host time
// spent building them, device allocation included), plans evicted, and what the cache holds now.
struct PlanCacheStats {
  int64_t hits = 0, builds = 0, evictions = 0, plans = 0, device_bytes = 0;
  int64_t releases = 0, restores = 0;  // plans whose slot buffers were released under the budget / allocated again
  int64_t collisions = 0;              // lookups whose key held a plan of other operand structures
  double build_ms = 0.0;
};
PlanCacheStats plan_cache_stats();
// Drop every cached plan (their task lists, intermediates and graphs go back to the device pool).
void plan_cache_clear();
// Drop every cached merged-operator plan (merge.cpp).
void merged_plan_cache_clear();
// The most recently used cached plans are never evicted: a caller may hold references to at most this many plans
// from the accessors below at once (the Davidson driver holds one per operator group).
constexpr size_t kPlanCacheKeep = 16;
// Plan for a two-site sigma vector (exposed for the Davidson driver, which binds device pointers itself).  `copy`
// selects one of several cached plans of the same structures: reduced-precision plans hold FP32 copies of their
// operands, so operator groups of equal structure need plans of their own to keep their copies resident.
Plan& plan_sigma_twosite(const Layout& lmpo, const Layout& rmpo, const Layout& lopr, const Layout& ropr,
                         const Layout& wfn, const Layout* out, int prec, int copy = 0);
// Cached plan of the one-site sigma vector (compute_sigma_vector.hpp:10-22).
Plan& plan_sigma_onesite(const Layout& mpo, const Layout& lopr, const Layout& ropr, const Layout& wfn,
                         const Layout* out, int prec, int copy = 0);
// Cached plan of the merged 2-index sigma vector (compute_sigma_vector.hpp:86-96).
Plan& plan_sigma_merged(const Layout& lopr, const Layout& ropr, const Layout& wfn, const Layout* out, int prec,
                        int copy = 0);
// The same plan, uncached and owned by the caller (batched launches keep one per instance: its intermediates are
// that instance's scratch).
Plan build_plan_sigma_twosite(const Layout& lmpo, const Layout& rmpo, const Layout& lopr, const Layout& ropr,
                              const Layout& wfn, const Layout* out, int prec);

void clear_plan_cache();
void clear_merged_operator_cache();  // device copies of the scalar-MPO merged-operator plans

// Quantum-number merging (include/btas/QSPARSE/QSTmerge.h) and merged operators (src/optimize.hpp:273-368)
struct MergeInfo {
  std::vector<QN> qm;     // merged sector quantum numbers (first-appearance order)
  std::vector<int> dm;    // merged sector dims
  std::vector<int> p2m;   // packed index -> merged sector
  std::vector<int> poff;  // packed index -> offset inside its merged sector
  std::vector<int> pdim;  // packed index -> dim
  std::vector<int> radix; // sector counts of the merged indices (packing radix, first index slowest)
};
MergeInfo merge_info(const std::vector<std::vector<QN>>& qs, const std::vector<std::vector<int>>& ds);
Layout merge_layout(const Layout& X, int nr, int nc, std::vector<CopySeg>& segs);
void op_qst_merge(Tensor& x, int nr, int nc, Tensor& y, hipStream_t s);
void op_merged_operator(bool fwd, Tensor& xmpo, Tensor& xopr, Tensor& mopr, hipStream_t s);
// Host plan of the merged operator when every MPO block is a scalar (false otherwise): output structure, one task per
// contributing sub-block (slice x column), the environment slice table and the MPO columns (see merge.cpp).
struct MergedOpPlan {
  Layout Y;
  std::vector<MergedOpTask> tasks;
  std::vector<Chunk> chunks;
  std::vector<int32_t> slice_blk, col_begin, ent_w;
  std::vector<int64_t> ent_off, env_off;
  int nw = 0;
};
bool merged_operator_plan(bool fwd, const Layout& W, const Layout& E, MergedOpPlan& P);
void op_merged_overlap(bool fwd, Tensor& xmpo, Tensor& xsov, Tensor& msov, hipStream_t s);

// Truncation rule of the canonicalizations on this thread (MPSXX_TRUNCATION_STABLE by default); set_ returns the
// previous rule.
int truncation_rule();
int set_truncation_rule(int rule);
struct TruncationScope {
  int old;
  explicit TruncationScope(int rule) : old(set_truncation_rule(rule)) {}
  ~TruncationScope() { set_truncation_rule(old); }
};
// make_noise (src/canonicalize.hpp:11-24): w += noise * U(-1, 1) draws, in place (nothing for noise < 1e-8).
void op_make_noise(Tensor& w, double noise, hipStream_t s);
// Two-site canonicalization with truncation to M states (src/canonicalize.hpp:27-46); returns the discarded weight.
double op_canonicalize_twosite(bool forward, Tensor& wfnc, Tensor& lmps, Tensor& rmps, int M, hipStream_t s);
// One-site form (src/canonicalize.hpp:49-85, noise 0): forward (LeftArrow) lmps = U(l,n,s), rmps = s*Vt(s,r);
// backward (RightArrow) lmps = U*s(l,s), rmps = Vt(s,n,r).  The caller multiplies the 2-index factor into the
// neighbouring site tensor, as the reference does.
double op_canonicalize_onesite(bool forward, Tensor& wfnc, Tensor& lmps, Tensor& rmps, int M, hipStream_t s);
// The same from the merged wavefunction wm((l,n1), (n2,r)) = QSTmerge of a 4-index psi whose index structure is
// `shape4` (q, qshape, dshape; blocks unused): the SVD blocks are wm's blocks, the factors come out as the 3-index
// site tensors U(l,n1,s) and V(s,n2,r), exactly as op_canonicalize_twosite gives them for the expanded psi.
double op_canonicalize_merged(bool forward, Tensor& wm, const Layout& shape4, Tensor& lmps, Tensor& rmps, int M,
                              hipStream_t s);

// Davidson (src/diagonalize.hpp:310-449) over operator groups: two-site groups (lmpo, rmpo, lopr, ropr) on a
// 4-index wavefunction; one-site groups (lmpo = the site MPO, rmpo = null) on a 3-index wavefunction
// (optimize_onesite); merged groups (lmpo = rmpo = null; lopr, ropr the merged operators of optimize.hpp:273-368) on
// the merged 2-index wavefunction (optimize_twosite_merged, optimize.hpp:495-590).
struct DavGroup {
  Tensor *lmpo, *rmpo, *lopr, *ropr;
  bool merged() const { return lmpo == nullptr; }
  bool onesite() const { return lmpo != nullptr && rmpo == nullptr; }
  int wfn_rank() const { return merged() ? 2 : onesite() ? 3 : 4; }
};
struct DavidsonStats {
  std::vector<double> evals;   // lowest Ritz value after each subspace solve (diagonalize.hpp:363's micro iterations)
  std::vector<double> rnorms;  // ||r||^2 of the residual tested after each of them (diagonalize.hpp:402-408)
  int nsigma = 0;
  int nsigma_fp32 = 0;    // sigma vectors applied in FP32 (mixed precision)
  double sigma_ms = 0.0;  // device time spent in sigma vectors
  double last_rnorm = 0.0;
};
// Residual tolerance (||r||^2) of the FP32 phase of a mixed-precision Davidson solve, and of the split-BF16 phase
// (products good to ~2^-16 relative).
constexpr double kMixedTolerance = 1.0e-9;
constexpr double kMixedToleranceX3 = 1.0e-6;
// blocks with both sides <= this run as 32x32 tiles in the merged sigma's plans (the default policy's 64 elsewhere)
constexpr int kMergedSmallLimit = 32;
// A lower root k for excited states (diagonalize.hpp:36-89, Functor with N_ > 1): sigma = H|k> with the cross
// environments of the current root and root k, overlap = S|k> with the cross overlaps, energy = E_k.  The Davidson
// operator becomes v = Hc - sum_k [<H_k|c> S_k + <S_k|c> H_k - E_k <S_k|c> S_k] and the result is orthogonalized
// against every S_k and normalized (Functor::orthonormalize).
struct DavLowerRoot {
  Tensor *sigma, *overlap;
  double energy;
};
// The host-side check run before every rocSOLVER batched SVD call (rocSOLVER's documented argument bounds for an
// m x n column-major problem with singular vectors, and the extent of each buffer from its base for `batch` strided
// problems; bytesE < 0: no E array).  Throws Error(MPSXX_ERR_HIP) naming the failed rule.
void gesvd_argument_check(int m, int n, int lda, int64_t sA, int64_t sS, int ldu, int64_t sU, int ldv, int64_t sV,
                          int64_t sE, int batch, int64_t bytesA, int64_t bytesS, int64_t bytesU, int64_t bytesV,
                          int64_t bytesE);

Layout davidson_layout(const std::vector<DavGroup>& groups, const Layout& psi);
double davidson_twosite(const std::vector<DavGroup>& groups, Tensor& diag, Tensor& wfnc, double tol, int prec,
                        DavidsonStats& st, hipStream_t s, const std::vector<DavLowerRoot>& lower = {});

}  // namespace mpx
