// Two-site canonicalization: quantum-number-blocked SVD with the reference's truncation rule, FP64 on the device.
//
// Reference: src/canonicalize.hpp:27-46 -> btas::Gesvd<.., LeftArrow|RightArrow> (include/btas/QSPARSE/QSTLAPACK.h:
// 122-288).  psi(l,n1,n2,r) is merged into a matrix over (l,n1) x (n2,r) sectors (QSTmerge), every sector block is
// decomposed (LAPACKE_dgesvd 'S','S' in the reference; here a batched one-sided Jacobi SVD over all blocks at once,
// or, for vectors longer than one work-group's LDS holds, one rocSOLVER strided-batched gesvd over all blocks
// zero-padded to the largest, which leaves the non-zero spectrum and its vectors unchanged), all singular
// values are sorted and cutoff = s[M-1] (M > 0 and fewer than the total, else 1e-16); every value >= cutoff is kept
// (so ties at the cutoff keep more than M states), and the factors are expanded back (QSTexpand).  LeftArrow gives
// U with total QN 0 and s*Vt carrying psi's QN (forward sweep); RightArrow gives U*s carrying psi's QN and Vt with
// total 0 (backward sweep).  Blocks of the row (column) sectors keep the merged-sector order of QSTmergeInfo.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <climits>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {

namespace {
thread_local int t_trunc_rule = MPSXX_TRUNCATION_STABLE;
}

int truncation_rule() { return t_trunc_rule; }

int set_truncation_rule(int rule) {
  if (rule != MPSXX_TRUNCATION_REFERENCE && rule != MPSXX_TRUNCATION_STABLE)
    throw Error(MPSXX_ERR_ARG, "canonicalize: unknown truncation rule");
  const int old = t_trunc_rule;
  t_trunc_rule = rule;
  return old;
}

namespace {

rocblas_handle solver_handle() {
  static rocblas_handle h = nullptr;
  if (!h) {
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(MPSXX_ERR_HIP, "rocblas_create_handle failed");
  }
  return h;
}

void rect_copy_segments(int64_t src, int64_t src_ld, int64_t R, int64_t C, int64_t dst, int64_t dst_ld,
                        std::vector<CopySeg>& segs) {
  for (int64_t r = 0; r < R; ++r) segs.push_back({dst + r * dst_ld, src + r * src_ld, (int32_t)C, 1});
}

void run_copy(const std::vector<CopySeg>& segs, const double* x, double* y, hipStream_t s) {
  if (segs.empty()) return;
  std::vector<Chunk> chunks;
  for (size_t i = 0; i < segs.size(); ++i)
    for (int c = 0; c < segs[i].len; c += kAxpyChunk) chunks.push_back({(int32_t)i, c});
  DevBuf ds = upload_vec(segs), dc = upload_vec(chunks);
  launch_copy<double, double>(dc.as<Chunk>(), (int)chunks.size(), ds.as<CopySeg>(), x, y, 1.0, 0.0, s);
  hip_check(hipStreamSynchronize(s), "copy");
}

}  // namespace

// make_noise (src/canonicalize.hpp:11-24).  The reference builds wfpt = QSTArray(q, qshape, check_net_dshape(),
// std::bind(uniform_real_distribution(-1, 1), mt19937())) and adds noise * wfpt.  Its generator constructor passes the
// generator by value to every block (STArray.h:243-262, TArray.h:779-782), so every block is filled, row-major, by a
// fresh copy of the default-seeded engine: element e of any block is the e-th draw.  The draws come from the same
// standard-library engine and distribution on the host, once, and one copy launch adds noise * draws[0, len) to every
// block of the net-dense-shape pattern (every allowed index tuple whose sectors all hold a stored block), inserting
// the missing ones as Axpy does (STBLAS.h:135-145).
void op_make_noise(Tensor& w, double noise, hipStream_t s) {
  if (noise < 1.0e-8 || w.empty) return;
  const Layout& X = w.L;
  const int N = X.rank;
  std::vector<std::vector<char>> present(N);
  for (int i = 0; i < N; ++i) present[i].assign(X.ds[i].size(), 0);
  for (int64_t b = 0; b < X.nblocks(); ++b)
    for (int i = 0; i < N; ++i) present[i][X.index(b)[i]] = 1;
  Layout U = X;
  int64_t combos = 1;
  for (int i = 0; i < N; ++i) combos *= std::max(1, X.nsec(i));
  if (combos > (int64_t)1 << 26) throw Error(MPSXX_ERR_ARG, "make_noise: too many sector combinations");
  std::vector<int32_t> idx(N, 0);
  for (int64_t t = 0; t < combos && N > 0; ++t) {
    int64_t r = t;
    bool ok = true;
    for (int i = N - 1; i >= 0; --i) {
      idx[i] = (int32_t)(r % X.nsec(i));
      r /= X.nsec(i);
      ok = ok && present[i][idx[i]] && X.ds[i][idx[i]] > 0;
    }
    if (ok && X.allowed(idx.data())) U.add_block(idx.data());
  }
  U.finalize();
  w.relayout(U, s);
  int64_t maxlen = 0;
  for (int64_t b = 0; b < U.nblocks(); ++b) maxlen = std::max(maxlen, U.bsize[b]);
  if (maxlen > INT32_MAX) throw Error(MPSXX_ERR_ARG, "make_noise: block too large");
  // the draw sequence of the default-seeded engine, extended on demand and kept on the device
  static std::vector<double> draws;
  static DevBuf ddraws;
  if ((int64_t)draws.size() < maxlen) {
    std::mt19937 rgen;
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    draws.resize((size_t)maxlen);
    for (double& v : draws) v = dist(rgen);
    ddraws = upload_vec(draws);
  }
  std::vector<CopySeg> segs;
  std::vector<Chunk> chunks;
  for (int64_t b = 0; b < U.nblocks(); ++b) {
    if (U.bsize[b] == 0) continue;
    segs.push_back({U.off[b], 0, (int32_t)U.bsize[b], 1});
    for (int c = 0; c < segs.back().len; c += kAxpyChunk) chunks.push_back({(int32_t)(segs.size() - 1), c});
  }
  if (chunks.empty()) return;
  DevBuf ds = upload_vec(segs), dc = upload_vec(chunks);
  launch_copy<double, double>(dc.as<Chunk>(), (int)chunks.size(), ds.as<CopySeg>(), ddraws.as<double>(), w.ptr(), noise,
                              1.0, s);
  hip_check(hipGetLastError(), "make_noise launch");
  hip_check(hipStreamSynchronize(s), "make_noise");
}

namespace {

struct Blk { int i, j, m, n; };  // merged (row sector, col sector) and its dims
// A dense source sub-rectangle (R x C, row length ld) placed at (ro, co) of the padded slot of block k.
struct Rect { int k; int64_t src, ld, R, C, ro, co; };

double canonicalize_core(bool forward, const Layout& X, int nr, const MergeInfo& MR, const MergeInfo& MC,
                         const std::vector<Blk>& blks, const std::vector<Rect>& rects, const double* src,
                         int64_t src_elems, Tensor& lmps, Tensor& rmps, int M, hipStream_t s);

}  // namespace

double op_canonicalize_twosite(bool forward, Tensor& wfnc, Tensor& lmps, Tensor& rmps, int M, hipStream_t s) {
  if (wfnc.empty || wfnc.L.rank != 4) throw Error(MPSXX_ERR_ARG, "canonicalize: expects a 4-index wavefunction");
  const Layout& X = wfnc.L;
  MergeInfo MR = merge_info({X.qs[0], X.qs[1]}, {X.ds[0], X.ds[1]});
  MergeInfo MC = merge_info({X.qs[2], X.qs[3]}, {X.ds[2], X.ds[3]});
  // merged matrix blocks (row sector, col sector) that hold at least one psi block
  std::vector<Blk> blks;
  {
    std::vector<std::pair<int, int>> seen;
    for (int64_t b = 0; b < X.nblocks(); ++b) {
      const int32_t* xi = X.index(b);
      const int i = MR.p2m[xi[0] * MR.radix[1] + xi[1]], j = MC.p2m[xi[2] * MC.radix[1] + xi[3]];
      seen.push_back({i, j});
    }
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    for (auto& p : seen) blks.push_back({p.first, p.second, MR.dm[p.first], MC.dm[p.second]});
  }
  std::vector<Rect> rects;
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    const int pr = xi[0] * MR.radix[1] + xi[1], pc = xi[2] * MC.radix[1] + xi[3];
    const Blk key{MR.p2m[pr], MC.p2m[pc], 0, 0};
    const int k = (int)(std::lower_bound(blks.begin(), blks.end(), key, [](const Blk& x, const Blk& y) {
                          return x.i != y.i ? x.i < y.i : x.j < y.j;
                        }) - blks.begin());
    // sub-block (pdim_r x pdim_c) at (poff_r, poff_c) of its merged block
    rects.push_back({k, X.off[b], MC.pdim[pc], MR.pdim[pr], MC.pdim[pc], MR.poff[pr], MC.poff[pc]});
  }
  return canonicalize_core(forward, X, 2, MR, MC, blks, rects, wfnc.ptr(), X.nelem, lmps, rmps, M, s);
}

double op_canonicalize_merged(bool forward, Tensor& wm, const Layout& shape4, Tensor& lmps, Tensor& rmps, int M,
                              hipStream_t s) {
  if (wm.empty || wm.L.rank != 2) throw Error(MPSXX_ERR_ARG, "canonicalize: expects a merged 2-index wavefunction");
  if (shape4.rank != 4) throw Error(MPSXX_ERR_ARG, "canonicalize: the expanded shape must have 4 indices");
  const Layout& X = shape4;
  MergeInfo MR = merge_info({X.qs[0], X.qs[1]}, {X.ds[0], X.ds[1]});
  MergeInfo MC = merge_info({X.qs[2], X.qs[3]}, {X.ds[2], X.ds[3]});
  const Layout& W = wm.L;
  if (W.qs[0] != MR.qm || W.qs[1] != MC.qm || W.ds[0] != MR.dm || W.ds[1] != MC.dm || W.q != X.q)
    throw Error(MPSXX_ERR_SHAPE, "canonicalize: the merged wavefunction is not the merge of the given shape");
  std::vector<Blk> blks;
  std::vector<Rect> rects;
  for (int64_t b = 0; b < W.nblocks(); ++b) {
    const int32_t* wi = W.index(b);
    blks.push_back({wi[0], wi[1], MR.dm[wi[0]], MC.dm[wi[1]]});
    rects.push_back({(int)b, W.off[b], MC.dm[wi[1]], MR.dm[wi[0]], MC.dm[wi[1]], 0, 0});
  }
  return canonicalize_core(forward, X, 2, MR, MC, blks, rects, wm.ptr(), W.nelem, lmps, rmps, M, s);
}

double op_canonicalize_onesite(bool forward, Tensor& wfnc, Tensor& lmps, Tensor& rmps, int M, hipStream_t s) {
  if (wfnc.empty || (wfnc.L.rank != 3 && wfnc.L.rank != 2))
    throw Error(MPSXX_ERR_ARG, "canonicalize: expects a 3-index (or 2-index merged) wavefunction");
  const Layout& X = wfnc.L;
  // LeftArrow: rows (l, n), columns r; RightArrow: rows l, columns (n, r)  (canonicalize.hpp:71-84); a 2-index
  // (merged) wavefunction is the matrix itself (canonicalize.hpp:88-106)
  const int nr = X.rank == 2 ? 1 : forward ? 2 : 1;
  auto side = [&](int lo, int hi) {
    std::vector<std::vector<QN>> qs(X.qs.begin() + lo, X.qs.begin() + hi);
    std::vector<std::vector<int>> ds(X.ds.begin() + lo, X.ds.begin() + hi);
    return merge_info(qs, ds);
  };
  MergeInfo MR = side(0, nr), MC = side(nr, X.rank);
  auto packed = [&](const int32_t* xi, int lo, int hi, const MergeInfo& Mi) {
    return hi - lo == 1 ? xi[lo] : xi[lo] * Mi.radix[1] + xi[lo + 1];
  };
  std::vector<Blk> blks;
  {
    std::vector<std::pair<int, int>> seen;
    for (int64_t b = 0; b < X.nblocks(); ++b) {
      const int32_t* xi = X.index(b);
      seen.push_back({MR.p2m[packed(xi, 0, nr, MR)], MC.p2m[packed(xi, nr, X.rank, MC)]});
    }
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    for (auto& p : seen) blks.push_back({p.first, p.second, MR.dm[p.first], MC.dm[p.second]});
  }
  std::vector<Rect> rects;
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    const int pr = packed(xi, 0, nr, MR), pc = packed(xi, nr, X.rank, MC);
    const Blk key{MR.p2m[pr], MC.p2m[pc], 0, 0};
    const int k = (int)(std::lower_bound(blks.begin(), blks.end(), key, [](const Blk& x, const Blk& y) {
                          return x.i != y.i ? x.i < y.i : x.j < y.j;
                        }) - blks.begin());
    rects.push_back({k, X.off[b], MC.pdim[pc], MR.pdim[pr], MC.pdim[pc], MR.poff[pr], MC.poff[pc]});
  }
  return canonicalize_core(forward, X, nr, MR, MC, blks, rects, wfnc.ptr(), X.nelem, lmps, rmps, M, s);
}

namespace {

constexpr int kJacobiMaxSweeps = 60;
constexpr int kJacobiFirstCheck = 4;

// Batched one-sided Jacobi SVD of every block.  Each block's kk vectors X (columns or rows of A, whichever are
// fewer) are sorted by decreasing norm (X P, svd_sort_kernel; MPSXX_SVD_SORT=0 skips it) and QR-factored, X P = Q R
// (svd_qr_kernel), and the Jacobi sweeps (svd_round_kernel) orthogonalize the columns of Y = R^T, which are graded,
// so the sweeps converge in a few passes: R^T W = U_R s, hence X = (Q W) s (P U_R)^T; Q W comes from applying the
// Householder reflectors to W (svd_qr_apply_kernel) and P U_R from putting Y's rows back in order.  The sort is a
// cheap stand-in for the column pivoting of the Drmac-Veselic preconditioner: on the L = 32, M = 512 Hubbard sweeps
// it cuts the sweeps of a centre-bond SVD from ~16 to ~9.  Blocks too long for the QR
// kernels' LDS are swept directly (X V = U_X s).  Writes the descending singular values to Sv (kmax per block), U_A
// row-major (mmax x kmax) to Ua and Vt_A row-major (kmax x nmax) to Vta, the slots LAPACK-style gesvd fills.  Returns
// false, having written none of them, when a vector is longer than one work-group's LDS holds or the sweeps do not
// converge.
bool jacobi_svd(const std::vector<Blk>& blks, const std::vector<Rect>& rects, const double* src, int kmax, int nmax,
                int64_t su, int64_t sv, double* Sv, double* Ua, double* Vta, int* sweeps_out, hipStream_t s) {
  const int nb = (int)blks.size();
  const bool trace = std::getenv("MPSXX_SVD_TRACE") != nullptr;
  std::vector<double> tph;  // phase end times (ms since entry, stream synchronized) when tracing
  const auto t_in = std::chrono::steady_clock::now();
  auto phase = [&]() {
    if (!trace) return;
    hip_check(hipStreamSynchronize(s), "svd trace");
    tph.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_in).count());
  };
  std::vector<SvdBlk> sb(nb);
  int64_t xo = 0, vo = 0;
  int len_max = 1, kk_max = 1;
  for (int k = 0; k < nb; ++k) {
    const int m = blks[k].m, n = blks[k].n;
    sb[k] = {xo, vo, m, n, std::max(m, n), std::min(m, n)};
    xo += (int64_t)sb[k].len * sb[k].kk;
    vo += (int64_t)sb[k].kk * sb[k].kk;
    len_max = std::max(len_max, sb[k].len);
    kk_max = std::max(kk_max, sb[k].kk);
  }
  const size_t lim = svd_round_lds_limit();
  const char* mode = std::getenv("MPSXX_SVD");
  const bool allow_qr = !(mode && std::strcmp(mode, "jacobi_noqr") == 0);
  const int qcw = !allow_qr ? 0 : svd_qr_lds_bytes(32, len_max) <= lim ? 32 : svd_qr_lds_bytes(16, len_max) <= lim ? 16 : 0;
  const bool qr = qcw > 0;
  const int jl_max = qr ? kk_max : len_max;  // length of the swept vectors
  int cols = 32;
  while (cols > 8 && svd_round_lds_bytes(cols, jl_max) > lim) cols /= 2;
  if (svd_round_lds_bytes(cols, jl_max) > lim || (size_t)kk_max * (sizeof(double) + sizeof(int)) > lim) return false;
  const int g = cols / 2;  // columns per group of a paired task

  // X: the kk vectors of each block, column-major (zero where no psi block lies)
  DevBuf X(std::max<int64_t>(xo, 1) * sizeof(double)), V(std::max<int64_t>(vo, 1) * sizeof(double));
  hip_check(hipMemsetAsync(X.p, 0, std::max<int64_t>(xo, 1) * sizeof(double), s), "memset");
  {
    std::vector<CopySeg> segs;
    for (const Rect& r : rects) {
      const SvdBlk& b = sb[r.k];
      if (b.n <= b.m) {  // columns of A: column co + c, rows ro ..
        for (int64_t c = 0; c < r.C; ++c)
          segs.push_back({b.x_off + (r.co + c) * b.m + r.ro, r.src + c, (int32_t)r.R, (int32_t)r.ld});
      } else {  // rows of A
        for (int64_t rr = 0; rr < r.R; ++rr)
          segs.push_back({b.x_off + (r.ro + rr) * b.n + r.co, r.src + rr * r.ld, (int32_t)r.C, 1});
      }
    }
    run_copy(segs, src, X.as<double>(), s);
  }
  phase();  // 0: gather
  // the swept matrices: Y = R^T (kk x kk) with QR, else X itself; V accumulates the rotations
  std::vector<SvdBlk> jb = sb;
  std::vector<SvdQrBlk> qb;
  DevBuf Y, Z, tau, T, dqb, Xs, perm;
  const char* sort_env = std::getenv("MPSXX_SVD_SORT");
  const bool sort_cols = qr && !(sort_env && std::strcmp(sort_env, "0") == 0);
  double* Xq = X.as<double>();  // the matrix the QR factors (X, or X with its columns sorted)
  int64_t y_elems = 0;
  if (qr) {
    int64_t yo = 0, zo = 0, to = 0, po = 0;
    for (int k = 0; k < nb; ++k) {
      const SvdBlk& b = sb[k];
      qb.push_back({b.x_off, yo, b.v_off, zo, to, po, b.len, b.kk});
      jb[k].x_off = yo;
      jb[k].len = b.kk;
      yo += (int64_t)b.kk * b.kk;
      zo += (int64_t)b.len * b.kk;
      to += b.kk;
      po += (int64_t)((b.kk + kSvdQrPanel - 1) / kSvdQrPanel) * kSvdQrPanel * kSvdQrPanel;
    }
    Y = DevBuf(std::max<int64_t>(yo, 1) * sizeof(double));
    Z = DevBuf(std::max<int64_t>(zo, 1) * sizeof(double));
    tau = DevBuf(std::max<int64_t>(to, 1) * sizeof(double));
    T = DevBuf(std::max<int64_t>(po, 1) * sizeof(double));
    dqb = upload_vec(qb);
    y_elems = yo;
    if (sort_cols) {  // the QR factors the norm-sorted columns; Y's rows are put back in order after the sweeps
      Xs = DevBuf(std::max<int64_t>(xo, 1) * sizeof(double));
      perm = DevBuf(std::max<int64_t>(to, 1) * sizeof(int32_t));
      launch_svd_sort(dqb.as<SvdQrBlk>(), nb, X.as<double>(), Xs.as<double>(), perm.as<int32_t>(), kk_max, s);
      Xq = Xs.as<double>();
    }
    // per panel: its factorization, then the columns right of it in 16-column chunks on work-groups of their own
    // (MPSXX_SVD_QR_SPLIT=0: the whole factorization of a block on one work-group)
    const char* split_env = std::getenv("MPSXX_SVD_QR_SPLIT");
    if (split_env && std::strcmp(split_env, "0") == 0) {
      launch_svd_qr(dqb.as<SvdQrBlk>(), nb, Xq, tau.as<double>(), T.as<double>(), qcw, len_max, s);
    } else {
      std::vector<int2> trail;
      for (int k = 0; k < nb; ++k)
        for (int c0 = kSvdQrPanel; c0 < qb[k].kk; c0 += kSvdQrPanel) trail.push_back(make_int2(k, c0));
      DevBuf dtrail = upload_vec(trail);
      launch_svd_qr_split(dqb.as<SvdQrBlk>(), nb, dtrail.as<int2>(), (int)trail.size(),
                          (kk_max + kSvdQrPanel - 1) / kSvdQrPanel, Xq, tau.as<double>(), T.as<double>(), len_max, s);
    }
    launch_svd_rt(dqb.as<SvdQrBlk>(), nb, Xq, Y.as<double>(), s);
    hip_check(hipGetLastError(), "svd qr launch");
  }
  phase();  // 1: QR
  double* Xj = qr ? Y.as<double>() : X.as<double>();
  DevBuf dblk = upload_vec(jb);
  launch_svd_init(dblk.as<SvdBlk>(), nb, V.as<double>(), s);

  // tasks of one sweep, round by round: blocks of at most `cols` vectors are one task of round 0 (all their pairs);
  // larger ones are cut into groups of g columns paired by the circle method, one task per pair of groups sweeping
  // all pairs of its 2g columns (the inner pairs of a group are swept again with each partner, which converges in
  // fewer sweeps than sweeping only the cross pairs, MPSXX_SVD_ORDER=cross, with the inner pairs in a round of their
  // own)
  const char* order = std::getenv("MPSXX_SVD_ORDER");
  const bool cross = order && std::strcmp(order, "cross") == 0;
  std::vector<std::vector<SvdTask>> rounds(1);
  for (int k = 0; k < nb; ++k) {
    const int kk = jb[k].kk;
    if (kk < 2) continue;
    if (kk <= cols) {
      rounds[0].push_back({k, 0, kk, 0, 0, 0});
      continue;
    }
    const int ng = (kk + g - 1) / g, P = ng + (ng & 1), r0 = cross ? 1 : 0;
    if ((int)rounds.size() < P - 1 + r0) rounds.resize(P - 1 + r0);
    auto gsize = [&](int x) { return std::min(g, kk - x * g); };
    if (cross)
      for (int x = 0; x < ng; ++x)
        if (gsize(x) > 1) rounds[0].push_back({k, x * g, gsize(x), 0, 0, 0});
    for (int r = 0; r < P - 1; ++r)
      for (int w = 0; w < P / 2; ++w) {
        int p = w == 0 ? P - 1 : (r + w) % (P - 1), q = w == 0 ? r : (r - w + P - 1) % (P - 1);
        if (p >= ng || q >= ng) continue;
        if (gsize(p) < gsize(q)) std::swap(p, q);  // the larger group first
        rounds[r + r0].push_back({k, p * g, gsize(p), q * g, gsize(q), cross ? 1 : 0});
      }
  }
  std::vector<SvdTask> flat;
  std::vector<int> roff{0};
  for (const auto& rt : rounds) {
    for (const SvdTask& t : rt) {
      const SvdBlk& b = jb[t.blk];
      if (t.n0 + t.n1 > cols || t.c0 + t.n0 > b.kk || t.c1 + t.n1 > b.kk || t.n0 < 1 || t.n1 < 0 ||
          (t.cross && (t.n1 < 1 || t.n1 > t.n0)))
        throw Error(MPSXX_ERR_ARG, "svd: bad Jacobi task");
      flat.push_back(t);
    }
    roff.push_back((int)flat.size());
  }
  DevBuf dtask = upload_vec(flat);
  phase();  // 2: schedule
  // a block takes part in a sweep while its previous sweep rotated (the device keeps the flags between sweeps); the
  // host reads whether any block rotated only from sweep kJacobiFirstCheck on (every sweep when tracing), since no
  // step converges sooner, and launches the earlier sweeps back to back
  std::vector<int32_t> act(nb, 1), cnt(nb, 0);
  DevBuf dact = upload_vec(act), drot = upload_vec(cnt), dany(sizeof(int32_t));
  bool converged = false;
  int sweep = 0;
  const int first_check = trace ? 1 : kJacobiFirstCheck;
  std::vector<int> last_sweep(nb, 0);  // per block, the last sweep that rotated (tracing)
  while (sweep < kJacobiMaxSweeps) {
    ++sweep;
    for (size_t r = 0; r + 1 < roff.size(); ++r)
      launch_svd_round(dtask.as<SvdTask>() + roff[r], roff[r + 1] - roff[r], dblk.as<SvdBlk>(), Xj, V.as<double>(),
                       dact.as<int32_t>(), drot.as<int32_t>(), cols, jl_max, s);
    launch_svd_sweep_end(drot.as<int32_t>(), dact.as<int32_t>(), nb, dany.as<int32_t>(), s);
    hip_check(hipGetLastError(), "svd round launch");
    if (sweep < first_check && sweep < kJacobiMaxSweeps) continue;
    int32_t any = 0;
    hip_check(hipMemcpyAsync(&any, dany.p, sizeof(int32_t), hipMemcpyDeviceToHost, s), "svd count");
    if (trace) hip_check(hipMemcpyAsync(act.data(), dact.p, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s), "svd act");
    hip_check(hipStreamSynchronize(s), "svd sweep");
    if (trace)
      for (int k = 0; k < nb; ++k)
        if (act[k]) last_sweep[k] = sweep;
    if (!any) {
      converged = true;
      break;
    }
  }
  if (sweeps_out) *sweeps_out = sweep;
  phase();  // 3: sweeps
  if (!converged) return false;
  std::vector<SvdOut> outs(nb);
  DevBuf Yu;
  if (qr) {
    std::vector<int2> items;  // the column chunks of every block's Z, one work-group each
    for (int k = 0; k < nb; ++k)
      for (int c0 = 0; c0 < qb[k].kk; c0 += qcw) items.push_back(make_int2(k, c0));
    DevBuf ditems = upload_vec(items);
    launch_svd_qr_apply(ditems.as<int2>(), (int)items.size(), dqb.as<SvdQrBlk>(), Xq, T.as<double>(), V.as<double>(),
                        Z.as<double>(), qcw, len_max, s);
    if (sort_cols) {
      Yu = DevBuf(std::max<int64_t>(y_elems, 1) * sizeof(double));
      launch_svd_unsort_rows(dqb.as<SvdQrBlk>(), nb, perm.as<int32_t>(), Y.as<double>(), Yu.as<double>(), s);
      Xj = Yu.as<double>();
    }
    for (int k = 0; k < nb; ++k) outs[k] = {qb[k].y_off, qb[k].z_off, sb[k].m, sb[k].n, sb[k].kk, sb[k].n > sb[k].m};
  } else {
    for (int k = 0; k < nb; ++k) outs[k] = {sb[k].x_off, sb[k].v_off, sb[k].m, sb[k].n, sb[k].kk, sb[k].n <= sb[k].m};
  }
  DevBuf douts = upload_vec(outs);
  hip_check(hipMemsetAsync(Ua, 0, nb * sv * sizeof(double), s), "memset");
  hip_check(hipMemsetAsync(Vta, 0, nb * su * sizeof(double), s), "memset");
  launch_svd_finish(douts.as<SvdOut>(), nb, Xj, qr ? Z.as<double>() : V.as<double>(), Sv, kmax, Ua, sv, Vta, su, nmax,
                    kk_max, truncation_rule() == MPSXX_TRUNCATION_STABLE ? 1 : 0, s);
  hip_check(hipGetLastError(), "svd finish launch");
  hip_check(hipStreamSynchronize(s), "svd finish");
  phase();  // 4: Q application and output
  if (trace) {  // per-call statistics of the batched Jacobi SVD on stderr
    fprintf(stderr, "svdtrace nb %d len_max %d kk_max %d qr %d cols %d rounds %d tasks %d sweeps %d ms", nb, len_max,
            kk_max, qcw, cols, (int)roff.size() - 1, (int)flat.size(), sweep);
    for (double t : tph) fprintf(stderr, " %.3f", t);
    fprintf(stderr, " | block kk:last-rotating-sweep");
    for (int k = 0; k < nb; ++k)
      if (jb[k].kk > 32) fprintf(stderr, " %d:%d", jb[k].kk, last_sweep[k]);
    fprintf(stderr, "\n");
  }
  return true;
}

// The argument rules of rocsolver_dgesvd_strided_batched / rocsolver_dgesvdj_strided_batched with singular vectors
// (rocSOLVER's documented requirements for an m x n column-major A, k = min(m, n): lda >= m, strideA >= lda n,
// strideS >= k, ldu >= m, strideU >= ldu k, ldv >= k, strideV >= ldv n, strideE >= k - 1), and that each device
// buffer holds the `batch` strided problems the call reads or writes, checked on the host before every call.
void check_gesvd_extents(int m, int n, int lda, int64_t sA, int64_t sS, int ldu, int64_t sU, int ldv, int64_t sV,
                         int64_t sE, int batch, int64_t bytesA, int64_t bytesS, int64_t bytesU, int64_t bytesV,
                         int64_t bytesE) {
  const int64_t k = std::min(m, n);
  auto fail = [&](const char* what) {
    throw Error(MPSXX_ERR_HIP, std::string("svd: rocSOLVER argument check failed (") + what + ") for m " +
                                   std::to_string(m) + " n " + std::to_string(n) + " batch " + std::to_string(batch));
  };
  if (m < 1 || n < 1 || batch < 1) fail("empty problem");
  if (lda < m || sA < (int64_t)lda * n) fail("lda / strideA");
  if (sS < k) fail("strideS");
  if (ldu < m || sU < (int64_t)ldu * k) fail("ldu / strideU");
  if (ldv < k || sV < (int64_t)ldv * n) fail("ldv / strideV");
  if (bytesE >= 0 && sE < k - 1) fail("strideE");
  auto holds = [&](int64_t bytes, int64_t stride, int64_t extent) {
    return bytes >= ((int64_t)(batch - 1) * stride + extent) * (int64_t)sizeof(double);
  };
  if (!holds(bytesA, sA, (int64_t)lda * n)) fail("A buffer");
  if (!holds(bytesS, sS, k)) fail("S buffer");
  if (!holds(bytesU, sU, (int64_t)ldu * k)) fail("U buffer");
  if (!holds(bytesV, sV, (int64_t)ldv * n)) fail("V buffer");
  if (bytesE >= 0 && !holds(bytesE, sE, std::max<int64_t>(k - 1, 1))) fail("E buffer");
}

void check_gesvd_args(int m, int n, int lda, int64_t sA, int64_t sS, int ldu, int64_t sU, int ldv, int64_t sV,
                      int64_t sE, int batch, const DevBuf& A, const DevBuf& S, const DevBuf& U, const DevBuf& V,
                      const DevBuf* E) {
  check_gesvd_extents(m, n, lda, sA, sS, ldu, sU, ldv, sV, sE, batch, (int64_t)A.bytes, (int64_t)S.bytes,
                      (int64_t)U.bytes, (int64_t)V.bytes, E ? (int64_t)E->bytes : -1);
}

// The rocSOLVER path: every block zero-padded to (mmax x nmax), one strided-batched gesvd (QR iteration) over all.
void rocsolver_svd(const std::vector<Blk>& blks, const std::vector<Rect>& rects, const double* src, int64_t src_elems,
                   int mmax, int nmax, int kmax, int64_t sa, int64_t su, int64_t sv, DevBuf& Sv, DevBuf& U, DevBuf& V,
                   std::vector<double>& sv_h, hipStream_t s) {
  const int nb = (int)blks.size();
  DevBuf A(nb * sa * sizeof(double)), E(nb * (int64_t)std::max(kmax, 2) * sizeof(double)),
      info(nb * sizeof(rocblas_int));
  hip_check(hipMemsetAsync(A.p, 0, nb * sa * sizeof(double), s), "memset");
  {
    // each source rectangle into the zero-padded row-major (mmax x nmax) slot of its block
    std::vector<CopySeg> segs;
    for (const Rect& r : rects)
      rect_copy_segments(r.src, r.ld, r.R, r.C, r.k * sa + r.ro * nmax + r.co, nmax, segs);
    run_copy(segs, src, A.as<double>(), s);
  }
  // row-major A (m x n) is column-major A^T (n x m): its left vectors are the rows of Vt, its right ones give U
  rocblas_handle h = solver_handle();
  rocblas_set_stream(h, s);
  // gesvd destroys A; keep the input for the (rare) blocks whose bidiagonal QR iteration does not converge
  DevBuf A0(nb * sa * sizeof(double));
  hip_check(hipMemcpyAsync(A0.p, A.p, nb * sa * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
  check_gesvd_args(nmax, mmax, nmax, sa, kmax, nmax, su, kmax, sv, std::max(kmax, 2), nb, A, Sv, U, V, &E);
  rocblas_status st = rocsolver_dgesvd_strided_batched(
      h, rocblas_svect_singular, rocblas_svect_singular, nmax, mmax, A.as<double>(), nmax, sa, Sv.as<double>(), kmax,
      U.as<double>(), nmax, su, V.as<double>(), kmax, sv, E.as<double>(), std::max(kmax, 2), rocblas_outofplace,
      info.as<rocblas_int>(), nb);
  if (st != rocblas_status_success) throw Error(MPSXX_ERR_HIP, "rocsolver_dgesvd_strided_batched failed");
  std::vector<rocblas_int> info_h(nb);
  hip_check(hipMemcpyAsync(sv_h.data(), Sv.p, sv_h.size() * sizeof(double), hipMemcpyDeviceToHost, s), "S");
  hip_check(hipMemcpyAsync(info_h.data(), info.p, nb * sizeof(rocblas_int), hipMemcpyDeviceToHost, s), "info");
  hip_check(hipStreamSynchronize(s), "svd sync");
  // blocks gesvd did not converge on (seen on small rank-deficient blocks of unconverged early-sweep states) are
  // decomposed again with the one-sided Jacobi driver, in place in their batch slots
  for (int k = 0; k < nb; ++k) {
    if (info_h[k] == 0) continue;
    DevBuf res(sizeof(double)), nsw(sizeof(rocblas_int)), inf(sizeof(rocblas_int));
    // block k alone, in its batch slot: the slot lies inside the buffers when slots 0..k do
    check_gesvd_args(nmax, mmax, nmax, sa, kmax, nmax, su, kmax, sv, 0, k + 1, A0, Sv, U, V, nullptr);
    st = rocsolver_dgesvdj_strided_batched(h, rocblas_svect_singular, rocblas_svect_singular, nmax, mmax,
                                           A0.as<double>() + k * sa, nmax, sa, 0.0, res.as<double>(), 100,
                                           nsw.as<rocblas_int>(), Sv.as<double>() + (int64_t)k * kmax, kmax,
                                           U.as<double>() + k * su, nmax, su, V.as<double>() + k * sv, kmax, sv,
                                           inf.as<rocblas_int>(), 1);
    if (st != rocblas_status_success) throw Error(MPSXX_ERR_HIP, "rocsolver_dgesvdj_strided_batched failed");
    rocblas_int info_j = 0;
    hip_check(hipMemcpyAsync(&info_j, inf.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, s), "info");
    hip_check(hipMemcpyAsync(sv_h.data() + (size_t)k * kmax, Sv.as<double>() + (int64_t)k * kmax,
                             kmax * sizeof(double), hipMemcpyDeviceToHost, s), "S");
    hip_check(hipStreamSynchronize(s), "svd sync");
    if (info_j == 0) info_h[k] = 0;
  }
  for (int k = 0; k < nb; ++k)
    if (info_h[k] != 0) {
      std::vector<double> x(src_elems);
      hip_check(hipMemcpy(x.data(), src, src_elems * sizeof(double), hipMemcpyDeviceToHost), "download");
      int64_t bad = 0;
      double amax = 0.0;
      for (double v : x) {
        if (!std::isfinite(v)) ++bad;
        else amax = std::max(amax, std::fabs(v));
      }
      throw Error(MPSXX_ERR_HIP, "gesvd did not converge (block " + std::to_string(k) + " of " + std::to_string(nb) +
                                     ", " + std::to_string(blks[k].m) + "x" + std::to_string(blks[k].n) +
                                     " padded to " + std::to_string(mmax) + "x" + std::to_string(nmax) + ", info " +
                                     std::to_string(info_h[k]) + "; input has " + std::to_string(bad) +
                                     " non-finite values, max |x| " + std::to_string(amax) + ")");
    }
}

double canonicalize_core(bool forward, const Layout& X, int nr, const MergeInfo& MR, const MergeInfo& MC,
                         const std::vector<Blk>& blks, const std::vector<Rect>& rects, const double* src,
                         int64_t src_elems, Tensor& lmps, Tensor& rmps, int M, hipStream_t s) {
  const int nb = (int)blks.size();
  if (nb == 0) throw Error(MPSXX_ERR_ARG, "canonicalize: empty wavefunction");
  const auto t_core = std::chrono::steady_clock::now();
  int mmax = 1, nmax = 1;
  for (const Blk& b : blks) { mmax = std::max(mmax, b.m); nmax = std::max(nmax, b.n); }
  const int kmax = std::min(mmax, nmax);
  const int64_t sa = (int64_t)mmax * nmax, su = (int64_t)nmax * kmax, sv = (int64_t)kmax * mmax;
  DevBuf Sv(nb * (int64_t)kmax * sizeof(double)), U(nb * su * sizeof(double)), V(nb * sv * sizeof(double));
  std::vector<double> sv_h((size_t)nb * kmax);
  const char* mode = std::getenv("MPSXX_SVD");
  const bool use_jacobi = !(mode && std::strcmp(mode, "rocsolver") == 0);
  if (use_jacobi && jacobi_svd(blks, rects, src, kmax, nmax, su, sv, Sv.as<double>(), V.as<double>(), U.as<double>(),
                               nullptr, s)) {
    hip_check(hipMemcpyAsync(sv_h.data(), Sv.p, sv_h.size() * sizeof(double), hipMemcpyDeviceToHost, s), "S");
    hip_check(hipStreamSynchronize(s), "svd sync");
    for (int k = 0; k < nb; ++k)
      for (int r = 0; r < std::min(blks[k].m, blks[k].n); ++r)
        if (!std::isfinite(sv_h[(size_t)k * kmax + r]))
          throw Error(MPSXX_ERR_HIP, "svd: non-finite singular value in block " + std::to_string(k) +
                                         " (the wavefunction has non-finite entries)");
  } else {
    rocsolver_svd(blks, rects, src, src_elems, mmax, nmax, kmax, sa, su, sv, Sv, U, V, sv_h, s);
    // numerical rank as on the Jacobi path: values below max(m, n) eps |A|_F count as zero
    for (int k = 0; k < nb && truncation_rule() == MPSXX_TRUNCATION_STABLE; ++k) {
      const int kk = std::min(blks[k].m, blks[k].n);
      double fro = 0.0;
      for (int r = 0; r < kk; ++r) fro += sv_h[(size_t)k * kmax + r] * sv_h[(size_t)k * kmax + r];
      const double thr = std::max(blks[k].m, blks[k].n) * 2.220446049250313e-16 * std::sqrt(fro);
      for (int r = 0; r < kk; ++r)
        if (sv_h[(size_t)k * kmax + r] < thr) sv_h[(size_t)k * kmax + r] = 0.0;
    }
  }
  // truncation (QSTLAPACK.h:190-232)
  std::vector<double> all;
  for (int k = 0; k < nb; ++k)
    for (int r = 0; r < std::min(blks[k].m, blks[k].n); ++r) all.push_back(sv_h[(size_t)k * kmax + r]);
  std::sort(all.rbegin(), all.rend());
  // MPSXX_TRUNCATION_REFERENCE: cutoff = the M-th largest value (0 < M < count), else 1e-16 (QSTLAPACK.h:209-213).
  // MPSXX_TRUNCATION_STABLE: values within a relative 1e-12 of the M-th largest count as equal to it, so exact
  // degeneracies (e.g. the +Sz/-Sz sector pairs) that rounding splits are kept together on every backend, and the
  // cutoff never drops below 1e-16 (when fewer than M values exceed it, zeros and rounding-level values are not kept
  // just because the M-th largest is one of them)
  double cutoff = 1.0e-16;
  if (M > 0 && M < (int)all.size())
    cutoff = truncation_rule() == MPSXX_TRUNCATION_REFERENCE ? all[M - 1] : std::max(cutoff, all[M - 1] * (1.0 - 1e-12));
  double dnorm = 0.0;
  std::vector<int> kept(nb, 0);
  for (int k = 0; k < nb; ++k) {
    const int kk = std::min(blks[k].m, blks[k].n);
    int r = 0;
    for (; r < kk; ++r)
      if (sv_h[(size_t)k * kmax + r] < cutoff) break;
    kept[k] = r;
    for (; r < kk; ++r) dnorm += sv_h[(size_t)k * kmax + r] * sv_h[(size_t)k * kmax + r];
  }
  // scale the factor that absorbs s (canonicalize.hpp:40-45: Dimm(s, Vt) forward, Dimm(U, s) backward)
  launch_svd_scale(forward ? U.as<double>() : V.as<double>(), nb, forward ? kmax : mmax, forward ? nmax : kmax,
                   forward ? su : sv, Sv.as<double>(), kmax, forward ? 0 : 1, s);
  // singular-value sectors: LeftArrow q_sval = q_rows, RightArrow q_sval = -q_cols; a sector is kept when its block
  // keeps at least one value
  std::vector<int> sec_of_blk(nb, -1);
  std::vector<QN> qnz;
  std::vector<int> dnz;
  {
    const int nsec = forward ? (int)MR.qm.size() : (int)MC.qm.size();
    for (int q = 0; q < nsec; ++q)
      for (int k = 0; k < nb; ++k)
        if ((forward ? blks[k].i : blks[k].j) == q && kept[k] > 0) {
          sec_of_blk[k] = (int)qnz.size();
          qnz.push_back(forward ? MR.qm[q] : -MC.qm[q]);
          dnz.push_back(kept[k]);
        }
  }
  // U(row indices..., s): quantum shape (q_rows..., -q_s); V(s, column indices...)
  const int nc = X.rank - nr;
  Layout LU;
  {
    std::vector<QN> nq;
    for (const QN& q : qnz) nq.push_back(-q);
    std::vector<std::vector<QN>> qs(X.qs.begin(), X.qs.begin() + nr);
    std::vector<std::vector<int>> ds(X.ds.begin(), X.ds.begin() + nr);
    qs.push_back(nq);
    ds.push_back(dnz);
    LU.set_shape(nr + 1, true, forward ? QN() : X.q, qs, ds);
  }
  Layout LV;
  {
    std::vector<std::vector<QN>> qs{qnz};
    std::vector<std::vector<int>> ds{dnz};
    qs.insert(qs.end(), X.qs.begin() + nr, X.qs.end());
    ds.insert(ds.end(), X.ds.begin() + nr, X.ds.end());
    LV.set_shape(nc + 1, true, forward ? X.q : QN(), qs, ds);
  }
  // packed index of a merged side -> its index tuple (first index slowest)
  auto unpack = [](const MergeInfo& Mi, int p, int n, int32_t* out) {
    if (n == 1) {
      out[0] = p;
    } else {
      out[0] = p / Mi.radix[1];
      out[1] = p % Mi.radix[1];
    }
  };
  std::vector<CopySeg> useg, vseg;
  struct Sub { int32_t idx[3]; int64_t src, ld; int R, C; };
  std::vector<Sub> us, vs;
  for (int k = 0; k < nb; ++k) {
    if (sec_of_blk[k] < 0) continue;
    const int d = kept[k];
    // rows of U: packed row indices of row sector i
    for (size_t p = 0; p < MR.p2m.size(); ++p) {
      if (MR.p2m[p] != blks[k].i || MR.pdim[p] == 0) continue;
      Sub sb;
      unpack(MR, (int)p, nr, sb.idx);
      sb.idx[nr] = sec_of_blk[k];
      if (!LU.allowed(sb.idx)) continue;
      sb.src = k * sv + (int64_t)MR.poff[p] * kmax;  // U_A row-major (mmax x kmax) in V
      sb.ld = kmax;
      sb.R = MR.pdim[p];
      sb.C = d;
      LU.add_block(sb.idx);
      us.push_back(sb);
    }
    for (size_t p = 0; p < MC.p2m.size(); ++p) {
      if (MC.p2m[p] != blks[k].j || MC.pdim[p] == 0) continue;
      Sub sb;
      sb.idx[0] = sec_of_blk[k];
      unpack(MC, (int)p, nc, sb.idx + 1);
      if (!LV.allowed(sb.idx)) continue;
      sb.src = k * su + MC.poff[p];  // Vt_A row-major (kmax x nmax) in U
      sb.ld = nmax;
      sb.R = d;
      sb.C = MC.pdim[p];
      LV.add_block(sb.idx);
      vs.push_back(sb);
    }
  }
  LU.finalize();
  LV.finalize();
  for (const Sub& sb : us) {
    const int64_t b = LU.find(LU.compute_tag(sb.idx));
    rect_copy_segments(sb.src, sb.ld, sb.R, sb.C, LU.off[b], sb.C, useg);
  }
  for (const Sub& sb : vs) {
    const int64_t b = LV.find(LV.compute_tag(sb.idx));
    rect_copy_segments(sb.src, sb.ld, sb.R, sb.C, LV.off[b], sb.C, vseg);
  }
  lmps.set_layout(LU, s);
  rmps.set_layout(LV, s);
  run_copy(useg, V.as<double>(), lmps.ptr(), s);
  run_copy(vseg, U.as<double>(), rmps.ptr(), s);
  if (std::getenv("MPSXX_SVD_TRACE")) {
    hip_check(hipStreamSynchronize(s), "svd trace");
    fprintf(stderr, "canontrace total %.3f ms\n",
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_core).count());
  }
  return dnorm;
}

}  // namespace

void gesvd_argument_check(int m, int n, int lda, int64_t sA, int64_t sS, int ldu, int64_t sU, int ldv, int64_t sV,
                          int64_t sE, int batch, int64_t bytesA, int64_t bytesS, int64_t bytesU, int64_t bytesV,
                          int64_t bytesE) {
  check_gesvd_extents(m, n, lda, sA, sS, ldu, sU, ldv, sV, sE, batch, bytesA, bytesS, bytesU, bytesV, bytesE);
}

}  // namespace mpx
