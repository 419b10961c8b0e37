// Two-site canonicalization: quantum-number-blocked SVD with the reference's truncation rule, FP64 on the device.
//
// Reference: src/canonicalize.hpp:27-46 -> btas::Gesvd<.., LeftArrow|RightArrow> (include/btas/QSPARSE/QSTLAPACK.h:
// 122-288).  psi(l,n1,n2,r) is merged into a matrix over (l,n1) x (n2,r) sectors (QSTmerge), every sector block is
// decomposed (LAPACKE_dgesvd 'S','S' in the reference; here one rocSOLVER strided-batched gesvd over all blocks,
// zero-padded to the largest block, which leaves the non-zero spectrum and its vectors unchanged), all singular
// values are sorted and cutoff = s[M-1] (M > 0 and fewer than the total, else 1e-16); every value >= cutoff is kept
// (so ties at the cutoff keep more than M states), and the factors are expanded back (QSTexpand).  LeftArrow gives
// U with total QN 0 and s*Vt carrying psi's QN (forward sweep); RightArrow gives U*s carrying psi's QN and Vt with
// total 0 (backward sweep).  Blocks of the row (column) sectors keep the merged-sector order of QSTmergeInfo.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {


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
  if (wfnc.empty || wfnc.L.rank != 3) throw Error(MPSXX_ERR_ARG, "canonicalize: expects a 3-index wavefunction");
  const Layout& X = wfnc.L;
  // LeftArrow: rows (l, n), columns r; RightArrow: rows l, columns (n, r)  (canonicalize.hpp:71-84)
  const int nr = forward ? 2 : 1;
  auto side = [&](int lo, int hi) {
    std::vector<std::vector<QN>> qs(X.qs.begin() + lo, X.qs.begin() + hi);
    std::vector<std::vector<int>> ds(X.ds.begin() + lo, X.ds.begin() + hi);
    return merge_info(qs, ds);
  };
  MergeInfo MR = side(0, nr), MC = side(nr, 3);
  auto packed = [&](const int32_t* xi, int lo, int hi, const MergeInfo& Mi) {
    return hi - lo == 1 ? xi[lo] : xi[lo] * Mi.radix[1] + xi[lo + 1];
  };
  std::vector<Blk> blks;
  {
    std::vector<std::pair<int, int>> seen;
    for (int64_t b = 0; b < X.nblocks(); ++b) {
      const int32_t* xi = X.index(b);
      seen.push_back({MR.p2m[packed(xi, 0, nr, MR)], MC.p2m[packed(xi, nr, 3, MC)]});
    }
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    for (auto& p : seen) blks.push_back({p.first, p.second, MR.dm[p.first], MC.dm[p.second]});
  }
  std::vector<Rect> rects;
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    const int pr = packed(xi, 0, nr, MR), pc = packed(xi, nr, 3, MC);
    const Blk key{MR.p2m[pr], MC.p2m[pc], 0, 0};
    const int k = (int)(std::lower_bound(blks.begin(), blks.end(), key, [](const Blk& x, const Blk& y) {
                          return x.i != y.i ? x.i < y.i : x.j < y.j;
                        }) - blks.begin());
    rects.push_back({k, X.off[b], MC.pdim[pc], MR.pdim[pr], MC.pdim[pc], MR.poff[pr], MC.poff[pc]});
  }
  return canonicalize_core(forward, X, nr, MR, MC, blks, rects, wfnc.ptr(), X.nelem, lmps, rmps, M, s);
}

namespace {

double canonicalize_core(bool forward, const Layout& X, int nr, const MergeInfo& MR, const MergeInfo& MC,
                         const std::vector<Blk>& blks, const std::vector<Rect>& rects, const double* src,
                         int64_t src_elems, Tensor& lmps, Tensor& rmps, int M, hipStream_t s) {
  const int nb = (int)blks.size();
  if (nb == 0) throw Error(MPSXX_ERR_ARG, "canonicalize: empty wavefunction");
  int mmax = 1, nmax = 1;
  for (const Blk& b : blks) { mmax = std::max(mmax, b.m); nmax = std::max(nmax, b.n); }
  const int kmax = std::min(mmax, nmax);
  const int64_t sa = (int64_t)mmax * nmax, su = (int64_t)nmax * kmax, sv = (int64_t)kmax * mmax;
  DevBuf A(nb * sa * sizeof(double)), Sv(nb * (int64_t)kmax * sizeof(double)), U(nb * su * sizeof(double)),
      V(nb * sv * sizeof(double)), E(nb * (int64_t)std::max(kmax, 2) * sizeof(double)), info(nb * sizeof(rocblas_int));
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
  rocblas_status st = rocsolver_dgesvd_strided_batched(
      h, rocblas_svect_singular, rocblas_svect_singular, nmax, mmax, A.as<double>(), nmax, sa, Sv.as<double>(), kmax,
      U.as<double>(), nmax, su, V.as<double>(), kmax, sv, E.as<double>(), std::max(kmax, 2), rocblas_outofplace,
      info.as<rocblas_int>(), nb);
  if (st != rocblas_status_success) throw Error(MPSXX_ERR_HIP, "rocsolver_dgesvd_strided_batched failed");
  std::vector<double> sv_h((size_t)nb * kmax);
  std::vector<rocblas_int> info_h(nb);
  hip_check(hipMemcpyAsync(sv_h.data(), Sv.p, sv_h.size() * sizeof(double), hipMemcpyDeviceToHost, s), "S");
  hip_check(hipMemcpyAsync(info_h.data(), info.p, nb * sizeof(rocblas_int), hipMemcpyDeviceToHost, s), "info");
  hip_check(hipStreamSynchronize(s), "svd sync");
  // blocks gesvd did not converge on (seen on small rank-deficient blocks of unconverged early-sweep states) are
  // decomposed again with the one-sided Jacobi driver, in place in their batch slots
  for (int k = 0; k < nb; ++k) {
    if (info_h[k] == 0) continue;
    DevBuf res(sizeof(double)), nsw(sizeof(rocblas_int)), inf(sizeof(rocblas_int));
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
  // truncation (QSTLAPACK.h:190-232)
  std::vector<double> all;
  for (int k = 0; k < nb; ++k)
    for (int r = 0; r < std::min(blks[k].m, blks[k].n); ++r) all.push_back(sv_h[(size_t)k * kmax + r]);
  std::sort(all.rbegin(), all.rend());
  // tie rule: values within a relative 1e-12 of the M-th largest count as equal to it, so exact degeneracies (e.g.
  // the +Sz/-Sz sector pairs) that rounding splits are kept together on every backend
  double cutoff = 1.0e-16;
  if (M > 0 && M < (int)all.size()) cutoff = all[M - 1] * (1.0 - 1e-12);
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
  return dnorm;
}

}  // namespace

}  // namespace mpx
