// Device task formats and kernel launchers (implemented in kernels.hip).
//
// All block-sparse work is expressed as flat task lists built on the host by the planner:
//   * GemmStep : C_blk(m x n, row-major) (+)= alpha * sum_p A'_p (m x k_p) * B'_p (k_p x n)
//                where A'_p(i,k) = A[a_off + i*sa_m + k*sa_k] and B'_p(k,j) = B[b_off + k*sb_k + j*sb_n]
//                (strided views replace the reference's explicit Permute copies, QSTCONTRACT.h:30-47);
//   * AxpyStep : C_blk[r] (+)= alpha * sum_p S[s_off_p] * V[v_off_p + r*v_stride_p]
//                (the contractions with a 1x1x1x1 MPO block, whose GEMM has k = 1 and one unit side);
//   * CopySegs : y[dst + r] = beta*y[dst + r] + alpha * x[src + r*stride]   (Tie gathers, up-casting copies, Axpy).
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>

namespace mpx {

struct GemmPair {
  int64_t a_off, b_off;
  int32_t sa_m, sa_k;
  int32_t sb_k, sb_n;
  int32_t k;
  int32_t pad;
};

struct GemmTask {
  int64_t c_off;
  int32_t m, n;
  int32_t pair_begin, pair_count;
  int32_t inst;    // instance of a batched launch (index into the operand pointer tables); 0 otherwise
  int32_t layout;  // kTaskContigA: every pair's A has unit stride along k; kTaskContigB: every pair's B
};
constexpr int32_t kTaskContigA = 1, kTaskContigB = 2;

// Operand bases of one launch: direct pointers (table == 0), or device tables of per-instance pointers indexed by the
// task's `inst` (table == 1), so one grouped launch covers many independent problems.
struct Operands {
  const void* a;
  const void* b;
  const void* b2;
  void* c;
  int table;
};

struct GemmTile {
  int32_t task;
  int32_t row0, col0;
  int32_t pad;
};

// A 64x64 tile of a large block over the pair range [pair_lo, pair_hi) of its task; slot >= 0: the reduction is split
// and this work-group writes its partial tile to slot `slot` of the step's partial buffer.
struct BigTile {
  int32_t task;
  int32_t row0, col0;
  int32_t pair_lo, pair_hi;
  int32_t slot;
};

// Tile whose partial sums occupy slots [slot0, slot0 + nslots) (summed in that order).
struct RedTile {
  int32_t task;
  int32_t row0, col0;
  int32_t slot0, nslots;
  int32_t pad;
};

// s2_off >= 0 marks a fused pair of two consecutive coefficient contractions: C[r] += S[s_off] * S2[s2_off] * V[..]
struct AxpyPair {
  int64_t v_off, s_off, s2_off;
  int32_t v_stride;
  int32_t pad;
};

struct AxpyTask {
  int64_t c_off;
  int32_t len;
  int32_t pair_begin, pair_count;
  int32_t inst;
};

struct Chunk {
  int32_t task;
  int32_t start;
};

struct CopySeg {
  int64_t dst_off, src_off;
  int32_t len, stride;
};

constexpr int kGemmTile = 64;     // C tile edge of the grouped GEMM
constexpr int kGemmTile2 = 128;   // C tile edge of the large-block grouped GEMM (gemm128.hpp)
// Tiles per XCD chunk of the 64x64 launches (MPSXX_XCD_CHUNK, default 16; <= 1: plain round-robin), see xcd_item.
int xcd_chunk();
constexpr int kAxpyChunk = 1024;  // elements per work-group in the AXPY / copy kernels

// Launchers.  T = storage/arithmetic type of A and B (double or float), TC = type of C.  max_groups > 0 caps the grid
// (persistent form: each work-group loops over the task list with a grid stride); 0 launches one per work item.
// 64x64 tiles, 4 waves, LDS-staged and software-pipelined (large blocks), grouped by their task's layout flags 3, 1,
// 2, 0 (ntiles[q] of group q, one launch each); split reductions write partial tiles (type T, 64*64 per slot) that a
// second launch sums in slot order.
template <typename T, typename TC>
void launch_gemm(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                 const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                 hipStream_t s, int max_groups = 0);
// 128x128 tiles of the largest blocks (gemm128.hpp), grouped by their task's layout flags 3, 1, 2, 0 (ntiles[q] of
// group q, one launch each); partial slots of 128x128
template <typename T, typename TC>
void launch_gemm128(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                    const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                    hipStream_t s, int max_groups = 0);

// Same 64x64 tiles with FP32 operands multiplied as split-BF16 pairs (three bf16 MFMAs per product, FP32
// accumulation; gemm_x3.hpp), grouped by layout as for launch_gemm.
template <typename TC>
void launch_gemm_x3(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                    const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                    hipStream_t s, int max_groups = 0);
// The 128x128 tiles of launch_gemm128 as split-BF16 products (gemm_x3.hpp); FP32 partial slots of 128x128.  The
// tiles come grouped by their task's layout flags (GemmTask::layout) 3, 1, 2, 0, ntiles[q] of group q, one launch each.
template <typename TC>
void launch_gemm_x3_128(const BigTile* tiles, const int (&ntiles)[4], const RedTile* red, int nred, void* partial,
                        const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                        hipStream_t s, int max_groups = 0);

// 32x32 tiles, one wave each, operands loaded straight into the MFMA registers (blocks up to ~64 on a side).
constexpr int kSmallTile = 32;
template <typename T, typename TC>
void launch_gemm_small(const GemmTile* tiles, int ntiles, const GemmTask* tasks, const GemmPair* pairs, Operands ops,
                       double alpha, int accumulate, hipStream_t s, int max_groups = 0);

// 32x32 tiles with the (pair, k-panel) reduction split over the 4 waves of a work-group, LDS sum at the end.  The
// tiles cover pair ranges; those with slot >= 0 write partial 32x32 tiles (T, row-major) that are then summed in slot
// order over `red` into C.
template <typename T, typename TC>
void launch_gemm_split(const BigTile* tiles, int ntiles, const RedTile* red, int nred, void* partial,
                       const GemmTask* tasks, const GemmPair* pairs, Operands ops, double alpha, int accumulate,
                       hipStream_t s, int max_groups = 0);

// ops.a = V (vector side), ops.b = S (coefficients), ops.b2 = second coefficients of fused pairs, ops.c = C
template <typename T, typename TC>
void launch_axpy(const Chunk* chunks, int nchunks, const AxpyTask* tasks, const AxpyPair* pairs, Operands ops,
                 double alpha, int accumulate, hipStream_t s, int max_groups = 0);

// One environment-sized sub-block of a merged block operator: its placement in the merged output (row length ld), the
// environment slice (bra, ket sector pair) it reads and the MPO column (site pair and outgoing operator) it sums over.
struct MergedOpTask {
  int64_t y_off;
  int32_t rows, cols, ld;
  int32_t slice, col;
  int32_t pad;
};
// out[sub-block] = sum over the column's entries (w, v) in list order of v * env[block of (slice, w)] (entries whose
// slice block is absent are skipped); slice_blk[slice * nw + w] = environment block or -1, env_off[block] = offset.
void launch_merged_operator(const Chunk* chunks, int nchunks, const MergedOpTask* tasks, const int32_t* slice_blk,
                            int nw, const int64_t* env_off, const int32_t* col_begin, const int32_t* ent_w,
                            const int64_t* ent_off, const double* env, const double* mpo, double* out, hipStream_t s);

// FP64 -> FP32 copies of many buffers in one launch (mixed-precision inputs of a batched launch); the chunks are
// (segment, start) pairs of kAxpyChunk elements.
struct ConvSeg {
  const double* src;
  float* dst;
  int64_t n;
};
void launch_convert_segments(const Chunk* chunks, int nchunks, const ConvSeg* segs, hipStream_t s, int max_groups = 0);

template <typename TX, typename TY>
void launch_copy(const Chunk* chunks, int nchunks, const CopySeg* segs, const TX* x, TY* y, double alpha,
                 double beta, hipStream_t s);

// Dot products over matched segments (x_off = dst_off, y_off = src_off, stride ignored).  Writes per-work-group
// partial sums into `partial` (size >= nchunks) and the total into out[0] with a deterministic second pass.
void launch_dot(const Chunk* chunks, int nchunks, const CopySeg* segs, const double* x, const double* y,
                double* partial, double* out, hipStream_t s);

// Dense helpers on contiguous buffers.
void launch_scale(double* x, int64_t n, double alpha, hipStream_t s);
template <typename TX, typename TY>
void launch_convert(const TX* x, TY* y, int64_t n, hipStream_t s);
void launch_fill(double* x, int64_t n, double v, hipStream_t s);
// y = a*x + b*y
void launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
// X[b*stride + i*cols + c] *= S[b*kmax + (by_col ? c : i)] for a batch of row-major (rows x cols) matrices
void launch_svd_scale(double* X, int batch, int rows, int cols, int64_t stride, const double* S, int kmax, int by_col,
                      hipStream_t s);

// Batched one-sided (Hestenes) Jacobi SVD of the quantum-number blocks of a merged wavefunction.  Block A (m x n,
// row-major) is decomposed through its kk = min(m, n) vectors of length len = max(m, n): the columns of A when
// n <= m, its rows otherwise.  X holds them column-major (len x kk) and is rotated in place until its columns are
// mutually orthogonal; V (kk x kk, column-major, starts as the identity) accumulates the rotations, so X = A V
// (or A^T V) holds U s (or V s) and V the other factor.
struct SvdBlk {
  int64_t x_off, v_off;
  int32_t m, n, len, kk;
};
// One work-group of a round: the columns [c0, c0 + n0) and [c1, c1 + n1) of block `blk` (n1 = 0: one group) are
// swept once through their Gram matrix, every pair of them (cross == 0) or only the n0 x n1 pairs across the two
// groups (cross == 1, n0 >= n1); the tasks of one round touch disjoint columns.
struct SvdTask {
  int32_t blk, c0, n0, c1, n1, cross;
};
// Largest dynamic LDS (bytes) a work-group of svd_round may use on the current device, and what a round with tasks
// of up to `cols` columns (8, 16 or 32) of length up to len_max needs.
size_t svd_round_lds_limit();
size_t svd_round_lds_bytes(int cols, int len_max);
void launch_svd_init(const SvdBlk* blks, int nb, double* V, hipStream_t s);
// cols = the most columns of any task (8, 16 or 32); len_max = the longest vector.
// rot[blk] is increased by the number of rotations applied; blocks with active[blk] == 0 are skipped.
void launch_svd_round(const SvdTask* tasks, int ntasks, const SvdBlk* blks, double* X, double* V,
                      const int32_t* active, int32_t* rot, int cols, int len_max, hipStream_t s);
// End of a sweep, on the device: active[k] = (rot[k] > 0), rot[k] = 0 for the nb blocks, and *any = 1 when any block
// rotated (0 when the sweep converged every block).
void launch_svd_sweep_end(int32_t* rot, int32_t* active, int nb, int32_t* any, hipStream_t s);
// QR preconditioning of the Jacobi sweeps (blocked Householder, one work-group per block, panels of kSvdQrPanel
// columns): X (len x kk) = Q R.  The sweeps then run on Y = R^T (kk x kk), whose columns are graded, and the left
// vectors of X are Z = Q [W; 0] with W the sweeps' accumulated rotations.  In place: R above and the Householder
// vectors below X's diagonal, tau (kk) and each panel's triangular factor T (kSvdQrPanel^2, block reflector
// I - V T V^T) per block.
constexpr int kSvdQrPanel = 16;
struct SvdQrBlk {
  int64_t x_off, y_off, w_off, z_off, tau_off, t_off;
  int32_t len, kk;
};
// LDS bytes of the QR / Q-application kernels for trailing chunks of cw (16 or 32) columns of length up to len_max.
size_t svd_qr_lds_bytes(int cw, int len_max);
void launch_svd_qr(const SvdQrBlk* blks, int nb, double* X, double* tau, double* T, int cw, int len_max,
                   hipStream_t s);
// The same factorization as 2 x npanels launches: per panel, its factorization (one work-group per block), then the
// update of the columns right of it in chunks of kSvdQrPanel, one work-group per item (block, first column) of
// `items` (every block's chunks from column kSvdQrPanel on).  Needs svd_qr_lds_bytes(kSvdQrPanel, len_max) of LDS.
void launch_svd_qr_split(const SvdQrBlk* blks, int nb, const int2* items, int nitems, int npanels, double* X,
                         double* tau, double* T, int len_max, hipStream_t s);
// Y = R^T (zero above the diagonal) and Z = [W; 0] from W (kk x kk), then Z <- Q Z.
void launch_svd_rt(const SvdQrBlk* blks, int nb, const double* X, double* Y, hipStream_t s);
// items: (block, first column) of every cw-wide column chunk of the blocks' Z, one work-group each.
void launch_svd_qr_apply(const int2* items, int nitems, const SvdQrBlk* blks, const double* X, const double* T,
                         const double* W, double* Z, int cw, int len_max, hipStream_t s);
// Column sorting before the QR, a cheap stand-in for column pivoting: each block's columns of X (len x kk) are copied
// to Xs (same offsets) in order of decreasing Euclidean norm, ties by index, and perm[tau_off + r] is the column
// placed at r.  R's rows are then graded, and the sweeps over R^T converge in fewer passes.  After the sweeps the rows
// of Y (kk x kk, at y_off) are put back in X's column order: Yu(perm[i], c) = Y(i, c).  Needs kk_max * 12 bytes of
// LDS.
void launch_svd_sort(const SvdQrBlk* blks, int nb, const double* X, double* Xs, int32_t* perm, int kk_max,
                     hipStream_t s);
void launch_svd_unsort_rows(const SvdQrBlk* blks, int nb, const int32_t* perm, const double* Y, double* Yu,
                            hipStream_t s);

// Output of a converged block: sigma_j = |xn_j| (xn: kk columns of length m when u_from_x, else n), sorted descending
// into S[blk * kmax + r].  U_A (m x kmax row-major, slot stride su_a) and Vt_A (kmax x nmax row-major, slot stride
// svt_a) are filled as a LAPACK-style gesvd of A returns them: one of them from xn / sigma (U_A when u_from_x), the
// other from va (kk columns of length n when u_from_x, else m).  Columns of zero singular value are left zero.
// zero_rank != 0: values below max(m, n) eps |A|_F of their block are reported as exact zeros.
struct SvdOut {
  int64_t xn_off, va_off;
  int32_t m, n, kk, u_from_x;
};
void launch_svd_finish(const SvdOut* outs, int nb, const double* Xn, const double* Va, double* S, int kmax, double* Ua,
                       int64_t su_a, double* Vta, int64_t svt_a, int nmax, int kk_max, int zero_rank, hipStream_t s);

// Davidson helpers on flat vectors of one layout (length n).
// y = sum_i c[i] * X_i  with X_i = base + i*ld  (FP64; c lives in device memory)
void launch_combine(const double* base, int64_t ld, int m, const double* c, double* y, int64_t n, hipStream_t s);
// r[i] = r[i] / d(i) with the reference's clamp; d_present[i] == 0 marks elements of blocks absent from diag
void launch_precondition(double* r, const double* diag, const unsigned char* present, double eval, int64_t n,
                         hipStream_t s);
// G[i*m + j] = <X_i, Y_j> for i < mx, j < my (FP64, deterministic tree reduction)
void launch_gram(const double* X, int64_t ldx, int mx, const double* Y, int64_t ldy, int my, int64_t n, double* G,
                 double* partial, int npart, hipStream_t s);
// y -= sum_i c[i] * X_i (c in device memory: Gram-Schmidt projections without a host round trip)
void launch_combine_sub(const double* base, int64_t ld, int m, const double* c, double* y, int64_t n, hipStream_t s);
// x *= 1 / sqrt(nrm2[0]) (0 when nrm2[0] == 0), nrm2 in device memory
void launch_scale_rnorm(double* x, int64_t n, const double* nrm2, hipStream_t s);
// the Ritz rotation of both vector sets in one pass: TO_j = sum_k c[j*cld + k] T_k, SO_j likewise from S (j, k < m)
void launch_rotate2(const double* T, const double* S, int64_t ld, int m, const double* c, int cld, double* TO,
                    double* SO, int64_t n, hipStream_t s);
// r = s0 - e t0 and G[0] = <r, r> (partial sums, then the tree reduction of launch_gram)
void launch_residual(const double* s0, const double* t0, double e, double* r, int64_t n, double* G, double* partial,
                     int npart, hipStream_t s);

// Davidson helpers over many independent vector spaces at once (batched Davidson): every instance owns one contiguous
// segment of a batch vector, cut into chunks of at most kSegChunk elements (one work-group per chunk).  Per-instance
// scalars (eigenvalues, norms, coefficient matrices) are read from device arrays indexed by the chunk's instance.
// Reductions are deterministic: one partial sum per (chunk, pair), then each instance sums its chunks in order.
constexpr int kSegChunk = 4096;
struct SegChunk {
  int64_t begin;  // first element (batch-vector offset)
  int32_t len;    // elements, 1..kSegChunk
  int32_t inst;
};
// G[i * gstride + a * my + b] = <X_a, Y_b> over instance i's segment, for a < mx, b < my; chunk0[i]..chunk0[i+1]-1 are
// the chunks of instance i (partial: nchunks * mx * my doubles)
void launch_seg_gram(const SegChunk* ch, int nch, const int32_t* chunk0, int ninst, const double* X, int64_t ldx,
                     int mx, const double* Y, int64_t ldy, int my, double* G, int gstride, double* partial,
                     hipStream_t s);
// r = s0 - e[i] t0 on instance i's segment and G[i * gstride] = <r, r>
void launch_seg_residual(const SegChunk* ch, int nch, const int32_t* chunk0, int ninst, const double* s0,
                         const double* t0, const double* e, double* r, double* G, int gstride, double* partial,
                         hipStream_t s);
// Ritz rotation of both vector sets with instance i's coefficients c + i * cstride (TO_j = sum_k c[j*cld + k] T_k)
void launch_seg_rotate2(const SegChunk* ch, int nch, const double* T, const double* S, int64_t ld, int m,
                        const double* c, int cld, int cstride, double* TO, double* SO, hipStream_t s);
// y -= sum_a c[i * cstride + a] X_a on instance i's segment
void launch_seg_combine_sub(const SegChunk* ch, int nch, const double* base, int64_t ld, int m, const double* c,
                            int cstride, double* y, hipStream_t s);
// x *= 1 / sqrt(nrm2[i * stride]) (0 when it is 0) on instance i's segment
void launch_seg_scale_rnorm(const SegChunk* ch, int nch, double* x, const double* nrm2, int stride, hipStream_t s);
// the reference preconditioner with instance i's eigenvalue eval[i]
void launch_seg_precondition(const SegChunk* ch, int nch, double* r, const double* diag, const unsigned char* present,
                             const double* eval, hipStream_t s);
// dst = src on the segments of the instances with mask[i] != 0 (mask == nullptr: every instance)
void launch_seg_copy(const SegChunk* ch, int nch, const double* src, double* dst, const int32_t* mask,
                     hipStream_t s);

}  // namespace mpx
