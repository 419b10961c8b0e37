/*
 * mpsxx_amd.h -- C ABI of the MI355X (gfx950) DMRG effective-Hamiltonian hot path.
 *
 * Every entry point takes plain pointers and sizes, returns an int status (0 = ok) and, on failure, leaves a message
 * readable through mpsxx_last_error() (thread-local).  Tensors are opaque device-resident block-sparse arrays with
 * U(1)xU(1) quantum numbers (particle number, 2Sz).  The host data layout used by upload/download is canonical:
 * blocks in ascending row-major block-index order ("tag"), each block dense and row-major.
 *
 * Semantics follow the reference (naokin/mpsxx, paths relative to its root):
 *   - outputs ACCUMULATE (beta = 1): an empty output (mpsxx_tensor_clear'ed, or freshly created empty) is shaped from
 *     the inputs; a non-empty one must carry the same total QN and quantum shape or MPSXX_ERR_QN is returned
 *     (include/btas/QSPARSE/QSTBLAS.h:201-211);
 *   - the output block set is every QN-allowed block that receives at least one (A-block, B-block) pair
 *     (include/btas/SPARSE/STBLAS.h:316-356), applied step by step along the reference contraction order;
 *   - calls are asynchronous on the given stream (NULL = default stream) except those that return host values.
 */
#ifndef MPSXX_AMD_H
#define MPSXX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------------------------------- */
enum {
  MPSXX_OK = 0,
  MPSXX_ERR_SHAPE = 1,       /* dense-shape mismatch                                                                */
  MPSXX_ERR_QN = 2,          /* quantum-number / quantum-shape mismatch (reference: BTAS_THROW -> runtime_error)   */
  MPSXX_ERR_HIP = 3,         /* HIP runtime error                                                                   */
  MPSXX_ERR_OOM = 4,         /* device allocation failed                                                            */
  MPSXX_ERR_ARG = 5,         /* invalid argument                                                                    */
  MPSXX_ERR_UNSUPPORTED = 6, /* contraction form outside the supported set                                          */
  MPSXX_ERR_IO = 7           /* file missing, unreadable, truncated or corrupt (reference: boost archive exceptions) */
};

/* ---- arithmetic of the block contractions ------------------------------------------------------------------------- */
enum {
  MPSXX_FP64 = 0,   /* FP64 storage and MFMA (v_mfma_f64_16x16x4_f64)                                                */
  MPSXX_FP32 = 1,   /* FP32 storage and MFMA with FP32 accumulation; outputs accumulated into the FP64 tensor       */
  MPSXX_BF16X3 = 2  /* FP32 storage; blocks above the small-tile size multiply as split-BF16 pairs x = hi + lo, three */
                    /* v_mfma_f32_16x16x32_bf16 per product (hi*hi + hi*lo + lo*hi), FP32 accumulation               */
};

/* ---- structure descriptor ----------------------------------------------------------------------------------------- */
typedef struct mpsxx_qst_desc {
  int32_t rank;             /* number of indices (1..8)                                                           */
  int32_t qsym;             /* 1: QSTArray (blocks must satisfy sum_i qshape[i][idx_i] == q_total); 0: STArray      */
  int32_t q_total[2];       /* total quantum number (particle number, 2Sz); ignored when qsym == 0                  */
  const int32_t* nsec;      /* [rank] number of quantum-number sectors of each index                              */
  const int32_t* sec_q;     /* [sum(nsec)][2] sector quantum numbers, index 0 first (NULL allowed if qsym == 0)   */
  const int32_t* sec_dim;   /* [sum(nsec)] dense dimension of each sector                                          */
  int64_t nblocks;          /* number of stored blocks                                                             */
  const int32_t* block_idx; /* [nblocks][rank] sector index of each stored block (any order)                        */
} mpsxx_qst_desc;

typedef struct mpsxx_tensor_s* mpsxx_tensor;

/* ---- library / error ---------------------------------------------------------------------------------------------- */
int mpsxx_init(int device);                      /* select the HIP device (called implicitly with 0)              */
const char* mpsxx_last_error(void);
int mpsxx_synchronize(void* stream);

/* ---- tensors (reference: btas::QSTArray, include/btas/QSPARSE/QSTArray.h:16-328) ---------------------------------- */
int mpsxx_tensor_create(const mpsxx_qst_desc* desc, mpsxx_tensor* out);  /* device storage, zero-filled           */
int mpsxx_tensor_create_empty(mpsxx_tensor* out);                        /* size()==0 array, shaped by first use   */
int mpsxx_tensor_free(mpsxx_tensor t);
int mpsxx_tensor_clear(mpsxx_tensor t);                                  /* QSTArray::clear                        */
int mpsxx_tensor_upload(mpsxx_tensor t, const double* host, int64_t nelem, void* stream);
int mpsxx_tensor_download(mpsxx_tensor t, double* host, int64_t nelem, void* stream);
/* structure queries: sizes first, then the arrays (caller-allocated with the sizes returned by _info) */
int mpsxx_tensor_info(mpsxx_tensor t, int32_t* rank, int32_t* qsym, int32_t q_total[2], int64_t* nsec_total,
                      int64_t* nblocks, int64_t* nelem);
int mpsxx_tensor_layout(mpsxx_tensor t, int32_t* nsec, int32_t* sec_q, int32_t* sec_dim, int32_t* block_idx,
                        int64_t* block_off);
/* device pointer of the FP64 storage (for callers that own streams/graphs) */
int mpsxx_tensor_device_ptr(mpsxx_tensor t, double** dptr);
/* copy of x with identical structure and data (btas::Copy) */
int mpsxx_tensor_copy(mpsxx_tensor x, mpsxx_tensor y, void* stream);

/* ---- on-disk arrays (reference: load/save of src/fileio.h:9-25, boost binary archives; names of getfile,
 * src/fileio.cpp:7-13).  Own self-describing little-endian format: magic "MPSXXQST", version 1, the structure
 * (rank, qsym, q_total, nsec, sector quantum numbers and dimensions, blocks in ascending tag), the FP64 data in
 * canonical order and an FNV-1a-64 checksum; readers reject a wrong magic/version, out-of-range or disallowed
 * structure, a truncated or over-long file and a checksum mismatch with MPSXX_ERR_IO.  Files are written beside the
 * target and renamed into place.  Rank 0 on disk is an array with no structure (size() == 0). ----------------------- */
/* host-only (no device needed): desc's blocks in any order; data holds nelem values in canonical order (blocks by
 * ascending tag, each row-major), as mpsxx_tensor_upload takes them */
int mpsxx_qst_save(const char* path, const mpsxx_qst_desc* desc, const double* data, int64_t nelem);  /* rank 0: empty */
/* sizes first (as mpsxx_tensor_info), then the arrays (as mpsxx_tensor_layout, plus data of nelem values) */
int mpsxx_qst_file_info(const char* path, int32_t* rank, int32_t* qsym, int32_t q_total[2], int64_t* nsec_total,
                        int64_t* nblocks, int64_t* nelem);
int mpsxx_qst_file_read(const char* path, int32_t* nsec, int32_t* sec_q, int32_t* sec_dim, int32_t* block_idx,
                        double* data, int64_t nelem);
/* device tensors: save downloads and writes; load replaces t's structure and data with the file's.  Synchronous. */
int mpsxx_tensor_save(mpsxx_tensor t, const char* path, void* stream);
int mpsxx_tensor_load(const char* path, mpsxx_tensor t, void* stream);
/* "[prefix]/[name].[site].[istate].[jstate].0.tmp" into buf (cap bytes including the terminating NUL); *len = the
 * name's length (without NUL) also when cap is too small (then MPSXX_ERR_ARG).  The reference's last field is the
 * MPI rank; one process drives one device here, so it is 0. */
int mpsxx_getfile(const char* name, const char* prefix, int site, int istate, int jstate, char* buf, int64_t cap,
                  int64_t* len);

/* ---- BTAS machinery (include/btas/QSPARSE/QSTCONTRACT.h:16-57, QSTBLAS.h:183-226) --------------------------------- */
/* c = alpha * contract(a[contract_a], b[contract_b]) + beta * c */
int mpsxx_contract(double alpha, mpsxx_tensor a, int conj_a, const int32_t* contract_a, mpsxx_tensor b, int conj_b,
                   const int32_t* contract_b, int32_t K, double beta, mpsxx_tensor c, void* stream);
/* btas::Gemm with CblasNoTrans (0) / CblasTrans (1) for each operand */
int mpsxx_gemm(int trans_a, int trans_b, double alpha, mpsxx_tensor a, int conj_a, mpsxx_tensor b, int conj_b,
               double beta, mpsxx_tensor c, void* stream);
/* btas::Tie (include/btas/SPARSE/STREINDEX.h:95-147): y = x with indices tie[0..K) merged onto tie[0]; y is STArray */
int mpsxx_tie(mpsxx_tensor x, const int32_t* tie, int32_t K, mpsxx_tensor y, void* stream);
/* btas::Copy(x, y, UpCast=1) onto y's existing q/qshape (include/btas/SPARSE/STBLAS.h:48-60) */
int mpsxx_copy_upcast(mpsxx_tensor x, mpsxx_tensor y, void* stream);

/* ---- BLAS-1 on block-sparse arrays (QSTBLAS.h:40-93, 242-252); FP64 accumulation ---------------------------------- */
int mpsxx_dotc(mpsxx_tensor x, mpsxx_tensor y, double* result, void* stream);     /* synchronous               */
int mpsxx_nrm2(mpsxx_tensor x, double* result, void* stream);                     /* synchronous               */
/* mpsxx_dotc with no device allocation: no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The segment list,
 * chunk list, partial sums and result live in a static 16 MiB scratch array that every device holds until the process
 * ends (its address is resolved once per device, at mpsxx_init or on the first call); returns MPSXX_ERR_ARG when one
 * call's lists need more (about 420,000 block segments, or 10^9 elements in 1,024-element chunks).  Same kernels and
 * summation order as mpsxx_dotc, so the same result bit for bit.  Calls that overlap on one device race on the
 * scratch array.  Synchronous. */
int mpsxx_dotc_noalloc(mpsxx_tensor x, mpsxx_tensor y, double* result, void* stream);
int mpsxx_axpy(double alpha, mpsxx_tensor x, mpsxx_tensor y, void* stream);       /* inserts missing y blocks  */
int mpsxx_scal(double alpha, mpsxx_tensor x, void* stream);

/* ---- hot path: sigma vector (src/compute_sigma_vector.hpp) -------------------------------------------------------- */
/* two-site 4-index form: replaces compute_sigma_vector(lmpo,rmpo,lopr,ropr,wfn0,sgv0), compute_sigma_vector.hpp:54-70 */
int mpsxx_compute_sigma_vector_twosite(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                       mpsxx_tensor wfn0, mpsxx_tensor sgv0, int precision, void* stream);
/* one-site form: replaces compute_sigma_vector(mpo0,lopr,ropr,wfn0,sgv0), compute_sigma_vector.hpp:10-22 */
int mpsxx_compute_sigma_vector_onesite(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0,
                                       mpsxx_tensor sgv0, int precision, void* stream);
/* merged 2-index form: replaces compute_sigma_vector(lopr,ropr,wfn0,sgv0) with 3-index operators,
 * compute_sigma_vector.hpp:86-96 */
int mpsxx_compute_sigma_vector_merged(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                      int precision, void* stream);
/* Reduced-precision variants for iterative callers whose operators stay fixed between calls (no reference
 * counterpart; the reference recomputes nothing per precision): bit i of convert_mask set = refresh the FP32 copy of
 * input i (argument order, e.g. 1 << 4 = the wavefunction of the two-site form, 1 << 2 of the merged form); a clear
 * bit reuses the copy made by the previous call with the same structures and precision.  FP64 ignores the mask. */
int mpsxx_compute_sigma_vector_twosite_ex(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                          mpsxx_tensor wfn0, mpsxx_tensor sgv0, int precision, uint32_t convert_mask,
                                          void* stream);
int mpsxx_compute_sigma_vector_merged_ex(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                         int precision, uint32_t convert_mask, void* stream);
/* overlap forms S|psi> for 3-, 4- and 2-index wavefunctions: compute_sigma_vector.hpp:25-35, 73-83, 99-109 */
int mpsxx_compute_sigma_vector_overlap(mpsxx_tensor lsov, mpsxx_tensor rsov, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                       void* stream);
/* <psi|H^* form: replaces compute_sigma_vector_conj, compute_sigma_vector.hpp:38-51 */
int mpsxx_compute_sigma_vector_conj(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0,
                                    mpsxx_tensor sgv0, void* stream);

/* ---- hot path: diagonal of H_eff (src/compute_diagonal_elements.hpp) --------------------------------------------- */
/* diag must already carry the wavefunction's q/qshape (reference constructs it that way, optimize.hpp:183) */
int mpsxx_compute_diagonal_elements_twosite(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr,
                                            mpsxx_tensor ropr, mpsxx_tensor diag, void* stream);   /* :35-61  */
int mpsxx_compute_diagonal_elements_onesite(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                            mpsxx_tensor diag, void* stream);                      /* :11-32  */
int mpsxx_compute_diagonal_elements_merged(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor diag,
                                           void* stream);                                          /* :64-80  */

/* ---- hot path: environment renormalization (src/renormalize.hpp) -------------------------------------------------- */
int mpsxx_renormalize(int forward, mpsxx_tensor mpo0, mpsxx_tensor opr0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                      mpsxx_tensor opr1, int precision, void* stream);                              /* :10-32  */
int mpsxx_renormalize_overlap(int forward, mpsxx_tensor sov0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                              mpsxx_tensor sov1, void* stream);                    /* :36-53 and merged :78-95 */
int mpsxx_renormalize_merged(int forward, mpsxx_tensor opr0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                             mpsxx_tensor opr1, int precision, void* stream);                       /* :57-74  */

/* ---- merged operator (src/optimize.hpp:273-368) ------------------------------------------------------------------- */
/* forward: L~(Lb,w',Lk) from L(lb,w,lk) and W(w,nb,nk,w'); backward: R~(Rb,w,Rk) from W and R(rb,w',rk); one call per
 * operator group (the reference loops over groups, optimize.hpp:318-322, 355-359).  The output is overwritten. */
int mpsxx_compute_merged_operator(int forward, mpsxx_tensor xmpo, mpsxx_tensor xopr, mpsxx_tensor mopr,
                                  void* stream);
/* merged overlap S~ = merge(S (x) I_site) (optimize.hpp:326-328, 363-365); output overwritten */
int mpsxx_compute_merged_overlap(int forward, mpsxx_tensor xmpo, mpsxx_tensor xsov, mpsxx_tensor msov,
                                 void* stream);
/* btas::QSTmerge(rows, a, cols, b) merging the first nrow and the last ncol indices (include/btas/QSPARSE/QSTmerge.h,
 * QSTmergeInfo.h:44-95); nrow or ncol = 1 leaves that side as is.  Output overwritten. */
int mpsxx_qst_merge(mpsxx_tensor x, int32_t nrow, int32_t ncol, mpsxx_tensor y, void* stream);

/* ---- Davidson (src/diagonalize.hpp:310-449) ----------------------------------------------------------------------- */
/* Ground state of sum_g H_g with H_g given by the two-site operator groups; wfnc is the initial guess and receives the
 * converged vector; returns the eigenvalue in *energy.  max_ritz = 10, <= 4 macro iterations, convergence on
 * ||r||^2 < tolerance as in the reference. precision selects the block-GEMM arithmetic of the sigma vectors. */
int mpsxx_davidson_twosite(int ngroups, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                           const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, mpsxx_tensor diag, mpsxx_tensor wfnc,
                           double tolerance, int precision, double* energy, int32_t* n_sigma, void* stream);

/* The same with lower roots (excited states, src/diagonalize.hpp:36-89 Functor with more than one wavefunction):
 * lower_sigma[k] = H|k> (sigma of lower root k with the cross environments <current|H|k>), lower_overlap[k] = S|k>
 * (compute_sigma_vector with the cross overlaps), lower_energy[k] = E_k; the operator is
 * v = Hc - sum_k [<H_k|c> S_k + <S_k|c> H_k - E_k <S_k|c> S_k] and the result is orthogonalized against every S_k and
 * normalized (Functor::orthonormalize).  lmpo = rmpo = NULL selects merged groups (lopr, ropr the merged operators of
 * compute_merged_operator, wfnc the merged 2-index wavefunction; optimize_twosite_merged). */
int mpsxx_davidson_twosite_ex(int ngroups, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                              const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, mpsxx_tensor diag, mpsxx_tensor wfnc,
                              int nlower, const mpsxx_tensor* lower_sigma, const mpsxx_tensor* lower_overlap,
                              const double* lower_energy, double tolerance, int precision, double* energy,
                              int32_t* n_sigma, void* stream);
/* No reference counterpart: the Davidson solvers replay each sigma plan's launch sequence from a hipGraph captured on
 * its first use (operand bases from a per-plan device table refreshed when they move); enable = 0 launches every
 * kernel from the host instead (same kernels, same results).  Default: on unless MPSXX_SIGMA_GRAPH=0.  Returns the
 * previous setting (1 on, 0 off) and applies to the whole process. */
int mpsxx_sigma_graphs(int enable);

/* No reference counterpart (diagnostic, host only): the check made before every rocSOLVER batched SVD call of the
 * fallback path -- rocSOLVER's documented argument bounds for an m x n column-major problem with singular vectors
 * (lda >= m, strideA >= lda n, strideS >= min(m,n), ldu >= m, strideU >= ldu min(m,n), ldv >= min(m,n),
 * strideV >= ldv n, strideE >= min(m,n) - 1) and that each buffer of the given byte size holds `batch` strided
 * problems (bytesE < 0: no E array).  MPSXX_ERR_HIP with the failed rule in mpsxx_last_error when violated. */
int mpsxx_check_gesvd_args(int m, int n, int lda, int64_t strideA, int64_t strideS, int ldu, int64_t strideU, int ldv,
                           int64_t strideV, int64_t strideE, int batch, int64_t bytesA, int64_t bytesS,
                           int64_t bytesU, int64_t bytesV, int64_t bytesE);

/* ---- two-site canonicalization (src/canonicalize.hpp:27-46 -> QSTLAPACK.h:122-288) -------------------------------- */
/* forward (LeftArrow): lmps = U (total QN 0), rmps = s*Vt; backward (RightArrow): lmps = U*s, rmps = Vt (total 0).
 * Keeps every singular value >= the M-th largest (M <= 0: all >= 1e-16); returns the discarded weight. */
int mpsxx_canonicalize_twosite(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps, int max_states,
                               double* discarded, void* stream);
/* Truncation rules.  REFERENCE: exactly QSTLAPACK.h:209-222 (cutoff = the M-th largest singular value when
 * 0 < M < count, else 1e-16; every value >= cutoff kept; singular values as computed).  STABLE (the default of
 * mpsxx_canonicalize_twosite and of the sweeps): values within a relative 1e-12 of the M-th largest count as ties and
 * are kept, the cutoff never drops below 1e-16, and values below max(m, n) eps |A|_F of their block are exact zeros.
 * The two differ only on near-ties at the cutoff and on values at the rounding level, where the reference's own
 * choice depends on the rounding of its SVD. */
enum { MPSXX_TRUNCATION_REFERENCE = 0, MPSXX_TRUNCATION_STABLE = 1 };
/* canonicalize(forward, wfnc, lmps, rmps, M, noise) (src/canonicalize.hpp:27-46): make_noise on wfnc in place, as the
 * reference does, then the SVD with the given truncation rule. */
int mpsxx_canonicalize_twosite_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps,
                                  int max_states, double noise, int rule, double* discarded, void* stream);
/* one-site canonicalize(forward, wfnc, lmps, rmps, M, noise) (src/canonicalize.hpp:49-85).  Forward: lmps = U,
 * rmps (the right neighbour, in) <- s Vt rmps; backward: rmps = Vt, lmps (the left neighbour, in) <- lmps U s.  With
 * noise >= 1e-8 the two-site wavefunction wfnc rmps (lmps wfnc) goes through the two-site form with noise, as the
 * reference does (canonicalize.hpp:57-66). */
int mpsxx_canonicalize_onesite_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps,
                                  int max_states, double noise, int rule, double* discarded, void* stream);
/* merged canonicalize(forward, wfnc, lmat, rmat, M, noise) on a 2-index wavefunction (src/canonicalize.hpp:88-106):
 * make_noise on wfnc in place, then lmat = U, rmat = s Vt (forward) or lmat = U s, rmat = Vt (backward). */
int mpsxx_canonicalize_merged_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmat, mpsxx_tensor rmat,
                                 int max_states, double noise, int rule, double* discarded, void* stream);
/* make_noise (src/canonicalize.hpp:11-24): nothing for noise < 1e-8, else x += noise * p, where p holds every allowed
 * block of x's net dense shape (sectors without a stored block excluded), each filled row-major with the draws of a
 * fresh default-seeded std::mt19937 through std::uniform_real_distribution<double>(-1, 1), as the reference's
 * generator constructor fills them; missing blocks are inserted.  Synchronous. */
int mpsxx_make_noise(mpsxx_tensor x, double noise, void* stream);

/* ---- device-resident two-site sweeps (dmrg/make_sweep.cpp:12-254, src/make_random_mpss.cpp:97-272) --------------- */
typedef struct mpsxx_dmrg_s* mpsxx_dmrg;
typedef struct mpsxx_sweep_stats {
  double seconds;          /* wall time of the sweep                                                              */
  double sigma_ms;         /* device time in sigma vectors (HIP events)                                           */
  double davidson_s, svd_s, renorm_s, diag_s;
  int32_t nsigma, max_bond;
  double discarded_weight;
  int32_t nsigma_fp32;     /* sigma vectors applied in FP32 (precision MPSXX_FP32: FP32 phase + FP64 refinement)    */
  int32_t pad;
} mpsxx_sweep_stats;
/* init_mps: site tensors A(l,n,r) with total QN 0 except the last site, which carries the target (N, 2Sz); the driver
 * right-canonicalizes them (truncating to max_states) and builds every right environment on the device. */
int mpsxx_dmrg_create(int nsites, const mpsxx_tensor* mpo, const mpsxx_tensor* init_mps, int max_states,
                      mpsxx_dmrg* out, void* stream);
/* one forward + one backward two-site sweep; *energy = minimum site energy (make_sweep.cpp:115-118) */
int mpsxx_dmrg_sweep(mpsxx_dmrg d, int max_states, double tolerance, int precision, double* energy,
                     mpsxx_sweep_stats* stats, void* stream);
int mpsxx_dmrg_site(mpsxx_dmrg d, int site, mpsxx_tensor out, void* stream);
/* Form of the two-site step used by later sweeps: MPSXX_SWEEP_FOURINDEX (default; dmrg/make_sweep.cpp:113 with
 * src/optimize.hpp:164-265: 4-index sigma, diag and renormalize) or MPSXX_SWEEP_MERGED (src/make_sweep.cpp:112 with
 * optimize_twosite_merged, src/optimize.hpp:495-630: merged operators built once per site, merged sigma and diag,
 * SVD of the merged wavefunction, environment = U~^+ L~ U~).  Same energies; the merged form is the reference's
 * quantum-chemistry path and avoids the per-site planning of the 4-index MPO-coefficient steps for large D_w.
 * MPSXX_SWEEP_ONESITE: one-site sweeps (dmrg/make_sweep.cpp with ONESITE, optimize_onesite, dmrg/optimize.hpp:20-125),
 * the refinement dmrg/dmrg.cpp:76-114 runs after the two-site sweeps have converged. */
enum { MPSXX_SWEEP_FOURINDEX = 0, MPSXX_SWEEP_MERGED = 1, MPSXX_SWEEP_ONESITE = 2 };
int mpsxx_dmrg_set_form(mpsxx_dmrg d, int form);
/* Noise and truncation rule of later sweeps: DMRGInput::noise (src/input.h:43, default 0), added by make_noise before
 * every two-site or merged canonicalize (src/optimize.hpp:210,237,586,618 -> canonicalize.hpp:35,95); rule is
 * MPSXX_TRUNCATION_STABLE (default) or MPSXX_TRUNCATION_REFERENCE.  One-site sweeps reject noise > 0. */
int mpsxx_dmrg_set_truncation(mpsxx_dmrg d, double noise, int rule);
/* The Davidson solves of the last sweep, in order (the reference logs them, diagonalize.hpp:317,363): per solve the
 * site (two-site pair (site, site+1) or one-site centre), the direction, the eigenvalue and the number of micro
 * iterations; evals holds every solve's lowest Ritz value after each subspace solve, concatenated, and rnorms the
 * squared residual norm tested after it (the convergence test is rnorm < tolerance).  Call with null arrays for the
 * counts. */
int mpsxx_dmrg_sweep_trace(mpsxx_dmrg d, int32_t max_steps, int32_t* nsteps, int32_t* step_site,
                           int32_t* step_forward, double* step_energy, int32_t* step_nevals, int32_t max_evals,
                           int32_t* nevals, double* evals, double* rnorms);
/* Excited states (dmrg/dmrg.cpp:36-72 with iroot > 0, dmrg/make_sweep.cpp, dmrg/optimize.hpp:127-230): register the
 * converged state `root` (energy E_k) as a lower root of `d`.  Both must have the centre at site 0 (as after
 * mpsxx_dmrg_create or a sweep).  Later sweeps of d project every lower root out of its Davidson solves and keep
 * the cross environments <d|H|k>, <d|k> up to date; the sweep energy is then root d's.  Two-site 4-index and
 * one-site forms. */
int mpsxx_dmrg_add_lower_root(mpsxx_dmrg d, mpsxx_dmrg root, double energy, void* stream);
/* Adjoint of the sigma-vector / renormalization contractions with respect to the MPO: grads[k] = <psi|dH/dW_k|psi>
 * for the current state (W_k's block indices, quantum numbers negated).  The centre moves across the chain without
 * truncation and back; sweeps can continue afterwards.  No reference counterpart. */
int mpsxx_dmrg_mpo_gradient(mpsxx_dmrg d, const mpsxx_tensor* grads, void* stream);
/* <psi|O|psi>/<psi|psi> for psi = prod_i mps[i] by left-to-right environment chains (src/expectation.hpp:55-114) */
int mpsxx_expectation(int nsites, const mpsxx_tensor* mpo, const mpsxx_tensor* mps, double* value, void* stream);
/* Sweep state on disk under the reference's site-file names (src/make_random_mpss.cpp:155-217 and
 * src/make_sweep.cpp:46-247 keep the state as "lmps" left of the centre, "wave" at it and "rmps" right of it, at
 * getfile(name, prefix, site, iroot)): one file per site plus the record getfile("dmrg", prefix, 0, iroot, 0) holding
 * nsites and the centre. */
int mpsxx_dmrg_save(mpsxx_dmrg d, const char* prefix, int iroot, void* stream);
/* Restart: reads the record and the site files and builds every right environment of the MPO on the device.  A state
 * saved with the centre at site 0 (after mpsxx_dmrg_create or a sweep) is used as saved; any other centre is brought
 * to site 0 as mpsxx_dmrg_create does (SVDs from the right, truncating to max_states).  Settings (form, noise, lower
 * roots) are not part of the state and start at their defaults. */
int mpsxx_dmrg_load(int nsites, const mpsxx_tensor* mpo, const char* prefix, int iroot, int max_states,
                    mpsxx_dmrg* out, void* stream);
int mpsxx_dmrg_free(mpsxx_dmrg d);
const char* mpsxx_dmrg_last_error(void);

/* ---- planner introspection (host only, no device work) ---------------------------------------------------------- */
/* Counters of the plan cache since the library was loaded: lookups served from the cache, plans built and the host
 * milliseconds spent building them (device allocation included), plans evicted, the plans and device bytes it holds
 * now, and how often a plan released its intermediate buffers to keep the cache's device memory within
 * MPSXX_PLAN_CACHE_GB (keeping its task lists) and allocated them again when next used.  Any pointer may be NULL.  No
 * reference counterpart (the reference recomputes every contraction's structure on each call). */
int mpsxx_plan_cache_stats(int64_t* hits, int64_t* builds, int64_t* evictions, double* build_ms, int64_t* plans,
                           int64_t* device_bytes, int64_t* releases, int64_t* restores);
/* Cached plans are keyed by 64-bit structure hashes of their operands and taken on a hit only when the operands'
 * block structures equal the ones the plan was built for; `collisions` = lookups whose key held a plan of other
 * structures (the plan is then looked up and built under an extended key).  mpsxx_set_structure_hash_bits keeps only
 * the low `bits` bits (1..64) of structure hashes computed from then on, so that tests can make distinct structures
 * collide; 64 (the default) restores full hashes.  Diagnostic; no reference counterpart. */
int mpsxx_plan_cache_collisions(int64_t* collisions);
int mpsxx_set_structure_hash_bits(int32_t bits);
/* Device allocations of the library since it was loaded: hipMalloc and hipFree calls and the host milliseconds they
 * took, buffers handed out again by the library's caching pool (MPSXX_DEVICE_CACHE_GB, default 31 % of device
 * memory) and the device synchronizations that took, the bytes in use and the bytes the pool holds.  Any pointer may
 * be NULL.  No reference counterpart (diagnostic). */
int mpsxx_device_memory_stats(int64_t* mallocs, int64_t* frees, double* malloc_ms, double* free_ms, int64_t* reuses,
                              int64_t* pool_syncs, int64_t* bytes_live, int64_t* cached_bytes);
/* Drop every cached plan (sigma, diagonal, renormalization and merged-operator plans, with their task lists,
 * intermediates and graphs) and return the pool's unused buffers to the device, after a device-wide
 * synchronization.  Later calls rebuild what they need, with the same results.  Call it when no other call of the
 * library is running (from any thread).  For callers that want the device memory back, and for timings that must
 * start without plans of an earlier run.  No reference counterpart (the reference keeps no plans). */
int mpsxx_release_caches(void);
/* Host threads of the contraction planner (threads <= 0: MPSXX_HOST_THREADS or the hardware concurrency, at most 16)
 * and the candidate pairs per thread below which a contraction stays on fewer threads.  Plans are identical for any
 * setting.  No reference counterpart (the reference enumerates pairs serially, STBLAS.h:316-356). */
int mpsxx_planner_threads(int threads, int64_t min_pairs_per_thread);
/* Block structure of btas::Contract(a[ca], b[cb]) (QSTCONTRACT.h:16-57 + STBLAS.h:298-359): the output blocks (indices
 * over a's free then b's free indices, ascending tag; a block exists iff it is quantum-number allowed and receives at
 * least one pair) and, per output block x, its (a block, b block) pairs pair_begin[x]..pair_begin[x+1] in accumulation
 * (contracted-tag) order.  Block numbers refer to each input's blocks in ascending tag order.  Call with block_idx =
 * NULL for the counts, then with arrays of nblocks*rank(c), nblocks+1 and npairs entries.  Blocks of zero size are
 * not supported in the pair numbering. */
int mpsxx_contract_structure(const mpsxx_qst_desc* a, int conj_a, const int32_t* ca, const mpsxx_qst_desc* b,
                             int conj_b, const int32_t* cb, int32_t k, int64_t* nblocks, int64_t* npairs,
                             int32_t* block_idx, int64_t* pair_begin, int64_t* pair_a, int64_t* pair_b);

/* ---- plan introspection (measurement) ---------------------------------------------------------------------------- */
/* Algorithmic cost of the two-site sigma vector for these inputs in the reference contraction order: flops = 2 m n k
 * summed over every accepted (A-block, B-block) pair of the four Contract steps, bytes = 8 * (|L|+|R|+|W|+|psi|+|sigma|)
 * counted over allocated blocks; nsteps = number of device steps of the plan. */
int mpsxx_sigma_twosite_stats(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                              mpsxx_tensor wfn0, int precision, int64_t* flops, int64_t* bytes, int32_t* nsteps);
/* Times every kernel launch of the two-site sigma plan separately (HIP events on the launch stream around each launch,
 * mean over `reps`, scratch output) and returns per launch: device ms, its share of the algorithmic flops and bytes,
 * the kernel (0 gemm_grouped_kernel 64x64, 1 gemm_small_kernel, 2 gemm_split_kernel, 3 axpy_grouped_kernel,
 * 4 gemm_x3_kernel, 5 gemm128_kernel 128x128, 6 gemm_x3_128_kernel) and the plan step it belongs to. */
int mpsxx_sigma_twosite_profile(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                mpsxx_tensor wfn0, int precision, int reps, int32_t max_entries, double* ms,
                                int64_t* flops, int64_t* bytes, int32_t* kernel, int32_t* step, int32_t* nentries,
                                void* stream);

/* Same per-launch profile for the plan of the most recent hot-path call on this thread (any form: merged, one-site,
 * renormalize, ...), re-run on the same inputs with the output redirected to scratch; the caller must still hold the
 * input tensors.  plan_flops = the plan's algorithmic flops. */
int mpsxx_last_plan_profile(int reps, int32_t max_entries, double* ms, int64_t* flops, int64_t* bytes,
                            int32_t* kernel, int32_t* step, int32_t* nentries, int64_t* plan_flops, void* stream);

/* Work items of the plan of the most recent hot-path call on this thread, summed over its steps: tiles[0] 64x64 tiles,
 * [1] 64x64 tiles of split reductions (partial-tile sums), [2] 32x32 tiles of the one-wave kernel, [3] 32x32 tiles of
 * the four-wave split kernel, [4] of those, tiles whose pair range was cut over several work-groups (partial-tile sums
 * of gemm_reduce32_kernel), [5] coefficient (AXPY) chunks, [6] 128x128 tiles of the largest blocks, [7] 128x128
 * tiles of split reductions.  Diagnostic; no reference counterpart. */
int mpsxx_last_plan_tiles(int64_t tiles[8]);

/* ---- quantum-chemistry MPO (src/gen_qc_operators.cpp:18-202, src/boundary_opinfo.cpp:47-140) ------------------- */
/* H = sum_{pq,s} h[p*K+q] a+_ps a_qs + 1/2 sum_{pqrs,st} v[((p*K+q)*K+r)*K+s] a+_ps a+_rt a_st a_qs + ecore, with v the
 * chemist-notation (pq|rs) (the reference's twoint(p,r,q,s), src/parsing_integral.cpp:139-147).  Bond operators are
 * the reference's (identity, H, normal singles of the left sites, complementary singles of the right sites, pair
 * operators of the smaller side when swap_sweep != 0, else of the left side).  Host-side: create, info and
 * site_entries need no GPU. */
typedef struct mpsxx_qcmpo_s* mpsxx_qcmpo;
int mpsxx_qc_mpo_create(int32_t norb, const double* oneint, const double* twoint, double ecore, int swap_sweep,
                        mpsxx_qcmpo* out);
/* bond_dims[0..norb]: number of bond operators left of site 0 (1), between sites, right of the last site (1) */
int mpsxx_qc_mpo_info(mpsxx_qcmpo m, int32_t* bond_dims, int64_t* nentries_total);
/* Site tensor W(wl, n_bra, n_ket, wr) as host arrays: two calls, the first with null buffers returns the counts.
 * lq/rq: (particle, 2Sz) of every left/right bond operator; per entry: wl, wr and the 4x4 (bra, ket) matrix. */
int mpsxx_qc_mpo_site_entries(mpsxx_qcmpo m, int32_t site, int32_t* nl, int32_t* nr, int64_t* nentries, int32_t* lq,
                              int32_t* rq, int32_t* wl, int32_t* wr, double* mats);
/* Site tensor as a device array (quantum shape (q_wl, q_n, -q_n, -q_wr), every non-zero element a 1x1x1x1 block) */
int mpsxx_qc_mpo_site(mpsxx_qcmpo m, int32_t site, mpsxx_tensor out, void* stream);
/* Energy derivatives from MPO gradients (mpsxx_dmrg_mpo_gradient): dh[p*K+q] = dE/dh_pq with h_pq and h_qp taken as
 * independent (= sum_s <a+_ps a_qs>, the 1-RDM), *dlambda = dE/dlambda for v -> lambda v at lambda = 1 (= <V>), *energy
 * = <H> rebuilt from the same derivatives (a consistency check).  No reference counterpart (the reference has no
 * derivatives); the 1-RDM it gives is the one src/opex.cpp:72-111 computes by expectation values. */
int mpsxx_qc_mpo_gradient(mpsxx_qcmpo m, int32_t nsites, const mpsxx_tensor* grads, double* dh, double* dlambda,
                          double* energy, void* stream);
int mpsxx_qc_mpo_free(mpsxx_qcmpo m);

/* ---- batched two-site sigma vectors (many independent instances per launch) ------------------------------------- */
/* No single reference counterpart: it runs compute_sigma_vector (src/compute_sigma_vector.hpp:54-70) for n independent
 * problems (different chains, bond dimensions and block structures) with ONE grouped launch per contraction step and
 * optional hipGraph replay.  The batch keeps pointers to the caller's tensors, which must outlive it; their block
 * structures are fixed at creation (a changed input structure makes run fail with MPSXX_ERR_SHAPE).  Not re-entrant. */
typedef struct mpsxx_batch_s* mpsxx_batch;
int mpsxx_batch_sigma_twosite_create(int32_t n, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                                     const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, const mpsxx_tensor* wfn0,
                                     const mpsxx_tensor* sgv0, int precision, mpsxx_batch* out, void* stream);
/* sgv0[i] += H_i wfn0[i] for all i (accumulate, as compute_sigma_vector).  FP32 batches refresh the FP32 copies of
 * the inputs whose bit is set in convert_mask (1 lmpo, 2 rmpo, 4 lopr, 8 ropr, 16 wfn0; 31 = all). */
int mpsxx_batch_sigma_twosite_run(mpsxx_batch b, uint32_t convert_mask, int use_graph, void* stream);
/* mean device time (HIP events on `stream`) of one run over `reps` runs; sgv0 accumulates every run */
int mpsxx_batch_time(mpsxx_batch b, int32_t reps, uint32_t convert_mask, int use_graph, double* ms, void* stream);
/* Times every kernel launch separately (as mpsxx_sigma_twosite_profile; outputs redirected to scratch): per launch the
 * device ms, algorithmic flops and bytes, kernel id and launch-group index (FP32 input conversion excluded). */
int mpsxx_batch_profile(mpsxx_batch b, int32_t reps, int32_t max_entries, double* ms, int64_t* flops, int64_t* bytes,
                        int32_t* kernel, int32_t* step, int32_t* nentries, void* stream);
/* Autotuner: times candidate tile policies per step on the batch's own task lists (outputs redirected to scratch) and
 * keeps the fastest; policy[4*k..4*k+3] = {small-block limit, split threshold, grid cap, 64x64 split target} of
 * launch k (mpsxx_batch_set_policy takes the same layout). */
int mpsxx_batch_tune(mpsxx_batch b, int32_t reps, int32_t max_steps, int32_t* policy, void* stream);
int mpsxx_batch_set_policy(mpsxx_batch b, int32_t nsteps, const int32_t* policy);
int mpsxx_batch_info(mpsxx_batch b, int64_t* flops, int64_t* bytes, int32_t* nsteps, int64_t* work_items);
int mpsxx_batch_free(mpsxx_batch b);

/* Batched Davidson: Davidson::diagonalize (src/diagonalize.hpp:283-449, one operator group per instance, one root)
 * on n independent two-site problems in lockstep.  Each step applies every instance's sigma vector in one batched
 * launch sequence (as mpsxx_batch_sigma_twosite_run, replayed from a hipGraph when use_graph != 0) and runs the dot
 * products, Ritz rotations, residuals, preconditioner and Gram-Schmidt of all instances as single segmented launches.
 * Per instance the result is mpsxx_davidson_twosite's on the same inputs up to rounding: wfnc[i] (the start vector)
 * is overwritten by the normalized eigenvector, energies[i] receives the lowest eigenvalue, n_sigma[i] and
 * n_sigma_low[i] the sigma vectors applied to the instance while it iterated (of them in FP32 / split-BF16x3),
 * residuals[i] its last ||r||^2 (below tolerance iff it converged).
 * precision: MPSXX_FP64, or MPSXX_FP32 / MPSXX_BF16X3 (low-precision sigma vectors until ||r||^2 < 1e-9 / 1e-6,
 * then FP64 refinement, as mpsxx_davidson_twosite).  diag[i] may be empty.  Output arrays may be NULL. */
typedef struct {
  int32_t batch_sigmas;      /* batched sigma applications (every instance at once) */
  int32_t batch_sigmas_low;  /* of them in FP32 / split-BF16x3 */
  int32_t macro_iterations;  /* Davidson restarts over both phases */
  int32_t converged;         /* instances that reached ||r||^2 < tolerance in the FP64 phase */
  double sigma_ms;           /* device time of the batched sigma applications */
  double seconds;            /* wall time of the call */
} mpsxx_batch_davidson_stats;
int mpsxx_batch_davidson_twosite(int32_t n, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                                 const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, const mpsxx_tensor* diag,
                                 const mpsxx_tensor* wfnc, double tolerance, int precision, int use_graph,
                                 double* energies, int32_t* n_sigma, int32_t* n_sigma_low, double* residuals,
                                 mpsxx_batch_davidson_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPSXX_AMD_H */
