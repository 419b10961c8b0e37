// Device Davidson eigensolver for the two-site effective Hamiltonian (reference: src/diagonalize.hpp:283-449).
//
// All Davidson vectors share one block layout U: the closure of psi's block set under the sigma-vector pattern map
// (sigma of a vector on U has its blocks in U).  Blocks the reference would not have allocated hold exact zeros, so
// every dot, axpy, scaling and the preconditioner give the reference's values; the vectors are then flat device
// arrays and all BLAS-1 work is contiguous.  Subspace algebra (<= 10 x 10) is FP64 on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {

namespace {

// Cyclic Jacobi for a symmetric n x n matrix (row-major).  Eigenvalues ascending; eigenvectors in the columns of v
// (the layout LAPACKE_dsyev returns for row-major storage, TLAPACK.h:23-47).
void sym_eig(int n, std::vector<double> a, std::vector<double>& w, std::vector<double>& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += a[i * n + j] * a[i * n + j];
        if (i != j) off += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-32 * tot || off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x * n + x] < a[y * n + y]; });
  w.resize(n);
  std::vector<double> v2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = a[ord[j] * n + ord[j]];
    for (int i = 0; i < n; ++i) v2[i * n + j] = v[i * n + ord[j]];
  }
  v.swap(v2);
}

}  // namespace

namespace {

// copy: which of the cached plans of this structure (see plan_sigma_twosite)
Plan& group_plan(const DavGroup& g, const Layout& U, int prec, int copy = 0) {
  if (g.merged()) return plan_sigma_merged(g.lopr->L, g.ropr->L, U, &U, prec, copy);
  if (g.onesite()) return plan_sigma_onesite(g.lmpo->L, g.lopr->L, g.ropr->L, U, &U, prec, copy);
  return plan_sigma_twosite(g.lmpo->L, g.rmpo->L, g.lopr->L, g.ropr->L, U, &U, prec, copy);
}

// Structure key of a group's operands: groups with equal keys would share one cached plan.
std::vector<uint64_t> group_structure(const DavGroup& g) {
  std::vector<uint64_t> k{(uint64_t)g.wfn_rank()};
  for (const Tensor* t : {g.lmpo, g.rmpo, g.lopr, g.ropr}) k.push_back(t ? t->L.hash : 0ull);
  return k;
}

}  // namespace

// Closure of a block layout under the sigma pattern maps of all groups.
Layout davidson_layout(const std::vector<DavGroup>& groups, const Layout& psi) {
  Layout U = psi;
  for (int iter = 0; iter < 16; ++iter) {
    Layout next = U;
    for (const DavGroup& g : groups) {
      Plan& P = group_plan(g, U, MPSXX_FP64);
      for (int64_t b = 0; b < P.out.nblocks(); ++b)
        if (next.find(P.out.tag[b]) < 0) next.add_block(P.out.index(b));
    }
    next.finalize();
    if (next.same_structure(U)) return U;
    U = next;
  }
  return U;
}

double davidson_twosite(const std::vector<DavGroup>& groups, Tensor& diag, Tensor& wfnc, double tol, int prec,
                        DavidsonStats& st, hipStream_t s, const std::vector<DavLowerRoot>& lower) {
  if (wfnc.empty) throw Error(MPSXX_ERR_ARG, "Davidson: empty initial wavefunction");
  if (groups.empty()) throw Error(MPSXX_ERR_ARG, "Davidson: no operator groups");
  if (groups.size() > kPlanCacheKeep) throw Error(MPSXX_ERR_ARG, "Davidson: more than 16 operator groups");
  const int kind = groups[0].wfn_rank();  // 4 two-site, 3 one-site, 2 merged
  for (const DavGroup& g : groups)
    if (g.wfn_rank() != kind) throw Error(MPSXX_ERR_ARG, "Davidson: operator groups of different forms mixed");
  if (wfnc.L.rank != kind) throw Error(MPSXX_ERR_ARG, "Davidson: wavefunction rank does not fit the groups");
  const int max_ritz = 10;
  // the vector space also holds every block of the lower roots' vectors: the projector's Axpy adds them to v
  // (STBLAS.h:135-145 reserves missing blocks), so they belong to the closure
  Layout U0 = wfnc.L;
  for (const DavLowerRoot& r : lower) {
    for (const Tensor* t : {r.sigma, r.overlap}) {
      if (!t || t->empty) throw Error(MPSXX_ERR_ARG, "Davidson: empty lower-root vector");
      if (t->L.rank != U0.rank || !t->L.same_qshape(U0) || !t->L.same_dshape(U0) || t->L.q != U0.q)
        throw Error(MPSXX_ERR_QN, "Davidson: lower-root vectors must have the wavefunction's quantum shape");
      for (int64_t b = 0; b < t->L.nblocks(); ++b)
        if (U0.find(t->L.tag[b]) < 0) U0.add_block(t->L.index(b));
    }
  }
  U0.finalize();
  Layout U = davidson_layout(groups, U0);
  const int64_t n = U.nelem;
  if (n == 0) throw Error(MPSXX_ERR_ARG, "Davidson: empty vector space");
  // vector slots: T[0..9], S[0..9], T2[0..9], S2[0..9], errv
  const int nslot = 4 * max_ritz + 1;
  DevBuf V((size_t)nslot * n * sizeof(double));
  double* base = V.as<double>();
  // trial and sigma sets and their rotation targets; the Ritz rotation writes the targets and swaps the roles
  double* tb = base;
  double* sb = base + (int64_t)max_ritz * n;
  double* t2b = base + (int64_t)(2 * max_ritz) * n;
  double* s2b = base + (int64_t)(3 * max_ritz) * n;
  auto T = [&](int i) { return tb + (int64_t)i * n; };
  auto Sg = [&](int i) { return sb + (int64_t)i * n; };
  double* errv = base + (int64_t)(4 * max_ritz) * n;
  hip_check(hipMemsetAsync(base, 0, (size_t)nslot * n * sizeof(double), s), "memset");

  // trial[0] <- wfnc placed into U (one launch for all blocks)
  {
    std::vector<CopySeg> segs;
    for (int64_t b = 0; b < wfnc.L.nblocks(); ++b)
      if (wfnc.L.bsize[b] > 0)
        segs.push_back({U.off[U.find(wfnc.L.tag[b])], wfnc.L.off[b], (int32_t)wfnc.L.bsize[b], 1});
    copy_segments(segs, wfnc.ptr(), T(0), s);
  }
  // lower-root vectors aligned to U (H_k, S_k for every k)
  const int nlow = (int)lower.size();
  DevBuf lowbuf((size_t)std::max(2 * nlow, 1) * n * sizeof(double));
  if (nlow > 0) hip_check(hipMemsetAsync(lowbuf.p, 0, (size_t)2 * nlow * n * sizeof(double), s), "memset");
  auto Hk = [&](int k) { return lowbuf.as<double>() + (int64_t)(2 * k) * n; };
  auto Sk = [&](int k) { return lowbuf.as<double>() + (int64_t)(2 * k + 1) * n; };
  for (int k = 0; k < nlow; ++k) {
    for (int which = 0; which < 2; ++which) {
      const Tensor& t = which ? *lower[k].overlap : *lower[k].sigma;
      double* dst = which ? Sk(k) : Hk(k);
      std::vector<CopySeg> segs;
      for (int64_t b = 0; b < t.L.nblocks(); ++b)
        if (t.L.bsize[b] > 0) segs.push_back({U.off[U.find(t.L.tag[b])], t.L.off[b], (int32_t)t.L.bsize[b], 1});
      copy_segments(segs, t.ptr(), dst, s);
    }
  }

  // diagonal aligned to U, with the presence mask of diag's blocks (precondition, diagonalize.hpp:289-303)
  DevBuf dvals(n * sizeof(double)), dmask(n);
  {
    std::vector<unsigned char> mask(n, 0);
    hip_check(hipMemsetAsync(dvals.p, 0, n * sizeof(double), s), "memset");
    if (!diag.empty) {
      std::vector<CopySeg> segs;
      for (int64_t b = 0; b < U.nblocks(); ++b) {
        const int64_t j = diag.L.find(U.tag[b]);
        if (j < 0) continue;
        if (diag.L.bsize[j] != U.bsize[b]) throw Error(MPSXX_ERR_SHAPE, "Davidson: diag block shape mismatch");
        if (U.bsize[b] > 0) segs.push_back({U.off[b], diag.L.off[j], (int32_t)U.bsize[b], 1});
        std::memset(mask.data() + U.off[b], 1, U.bsize[b]);
      }
      copy_segments(segs, diag.ptr(), dvals.as<double>(), s);
    }
    hip_check(hipMemcpyAsync(dmask.p, mask.data(), n, hipMemcpyHostToDevice, s), "mask");
    hip_check(hipStreamSynchronize(s), "mask sync");
  }

  // device scalars: Gram results and combine coefficients
  const int npart = (int)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 512);
  DevBuf gram(2 * max_ritz * max_ritz * sizeof(double)), part((size_t)2 * max_ritz * max_ritz * npart * sizeof(double));
  DevBuf coef(max_ritz * max_ritz * sizeof(double));
  auto dot = [&](const double* x, const double* y) {
    launch_gram(x, n, 1, y, n, 1, n, gram.as<double>(), part.as<double>(), npart, s);
    double h;
    hip_check(hipMemcpyAsync(&h, gram.p, sizeof(double), hipMemcpyDeviceToHost, s), "dot");
    hip_check(hipStreamSynchronize(s), "dot sync");
    return h;
  };
  auto normalize = [&](double* x) { launch_scale(x, n, 1.0 / std::sqrt(dot(x, x)), s); };
  // the same on the device, with no host round trip (the norm stays in gram[kNrm])
  const int kNrm = 2 * max_ritz * max_ritz - 1, kProj = max_ritz * max_ritz;
  auto dnormalize = [&](double* x) {
    launch_gram(x, n, 1, x, n, 1, n, gram.as<double>() + kNrm, part.as<double>(), npart, s);
    launch_scale_rnorm(x, n, gram.as<double>() + kNrm, s);
  };
  auto axpy = [&](double a, const double* x, double* y) { launch_axpby(a, x, 1.0, y, n, s); };

  std::vector<Plan*> plans;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  // FP32 plans convert the environments and MPO tensors once per Davidson call (they do not change inside it) and
  // the trial vector on every application
  unsigned convert_mask = ~0u;
  // sigma plans replayed from their hipGraphs (the launch sequence of each group captured once per plan)
  const bool graphs = sigma_graphs_enabled();
  auto run = [&](Plan& P, const std::vector<const double*>& in, double* v) {
    if (graphs) run_plan_graph(P, in, v, s, convert_mask);
    else run_plan_ex(P, in, v, s, {}, convert_mask);
  };
  auto multH = [&](const double* c, double* v) {
    hip_check(hipEventRecord(e0, s), "event");
    hip_check(hipMemsetAsync(v, 0, n * sizeof(double), s), "memset");
    for (size_t g = 0; g < groups.size(); ++g) {
      const DavGroup& gr = groups[g];
      if (kind == 2)
        run(*plans[g], {gr.lopr->ptr(), gr.ropr->ptr(), c}, v);
      else if (kind == 3)
        run(*plans[g], {gr.lmpo->ptr(), gr.lopr->ptr(), gr.ropr->ptr(), c}, v);
      else
        run(*plans[g], {gr.lmpo->ptr(), gr.rmpo->ptr(), gr.lopr->ptr(), gr.ropr->ptr(), c}, v);
    }
    convert_mask = 1u << (kind == 2 ? 2 : kind == 3 ? 3 : 4);  // from now on only the trial vector changes
    // lower roots: v -= <H_k|c> S_k + <S_k|c> H_k - E_k <S_k|c> S_k (diagonalize.hpp:72-79)
    for (int k = 0; k < nlow; ++k) {
      const double hk = dot(Hk(k), c), sk = dot(Sk(k), c);
      axpy(-hk, Sk(k), v);
      axpy(-sk, Hk(k), v);
      axpy(lower[k].energy * sk, Sk(k), v);
    }
    hip_check(hipEventRecord(e1, s), "event");
    hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    st.sigma_ms += ms;
    ++st.nsigma;
  };

  double eval = 0.0;
  std::vector<double> G(max_ritz * max_ritz), R, coef_h(max_ritz * max_ritz, 0.0);
  // one Davidson solve from the current trial_0 with the sigma arithmetic of `p`
  // Reduced-precision plans keep FP32 copies of their group's operators for the whole solve, so a group whose
  // structure repeats an earlier group's takes a plan of its own; FP64 plans read the operators directly and are
  // shared.
  auto use_plans = [&](int p) {
    plans.clear();
    for (size_t g = 0; g < groups.size(); ++g) {
      int copy = 0;
      if (p != MPSXX_FP64)
        for (size_t h = 0; h < g; ++h) copy += group_structure(groups[h]) == group_structure(groups[g]);
      plans.push_back(&group_plan(groups[g], U, p, copy));
    }
    convert_mask = ~0u;
  };
  // have_sigma0: T(0) is normalized and Sg(0) = H T(0) was computed at this precision already
  auto solve = [&](int p, double tol_p, bool have_sigma0) {
    use_plans(p);
    int niter = 0, iconv = 0;
    while (iconv < 1 && niter < 4) {
    if (!have_sigma0) {
      dnormalize(T(0));
      multH(T(0), Sg(0));
    }
    have_sigma0 = false;
    for (int m = 1; m <= max_ritz; ++m) {
      // heff(i,j) = <trial_i|sigma_j>  (diagonalize.hpp:339-365; the overlap matrix is computed but unused there)
      launch_gram(T(0), n, m, Sg(0), n, m, n, gram.as<double>(), part.as<double>(), npart, s);
      hip_check(hipMemcpyAsync(G.data(), gram.p, m * m * sizeof(double), hipMemcpyDeviceToHost, s), "gram");
      hip_check(hipStreamSynchronize(s), "gram sync");
      std::vector<double> heff(m * m);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) heff[i * m + j] = heff[j * m + i] = G[i * m + j];  // H(i,j) = <t_i|s_j>, j<=i
      std::vector<double> rval;
      sym_eig(m, heff, rval, R);
      eval = rval[0];
      st.evals.push_back(eval);
      // rotate: trial_j <- sum_i rvec(i,j) trial_i (and sigma), diagonalize.hpp:378-394.  The host staging array
      // outlives the asynchronous copy (the next write to it follows several stream synchronisations).
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) coef_h[j * max_ritz + i] = R[i * m + j];
      hip_check(hipMemcpyAsync(coef.as<double>(), coef_h.data(), max_ritz * max_ritz * sizeof(double),
                               hipMemcpyHostToDevice, s),
                "coef");
      // both sets rotated in one pass into the other buffers, which then take over
      launch_rotate2(T(0), Sg(0), n, m, coef.as<double>(), max_ritz, t2b, s2b, n, s);
      std::swap(tb, t2b);
      std::swap(sb, s2b);
      // error vector r = sigma_0 - E trial_0 and ||r||^2 < tol (diagonalize.hpp:416-422)
      double rnorm = 0.0;
      launch_residual(Sg(0), T(0), eval, errv, n, gram.as<double>() + kNrm, part.as<double>(), npart, s);
      hip_check(hipMemcpyAsync(&rnorm, gram.as<double>() + kNrm, sizeof(double), hipMemcpyDeviceToHost, s), "rnorm");
      hip_check(hipStreamSynchronize(s), "rnorm sync");
      st.last_rnorm = rnorm;
      st.rnorms.push_back(rnorm);
      if (rnorm < tol_p) {
        ++iconv;
        break;
      }
      if (m < max_ritz) {
        // preconditioned residual orthogonalized against the trials and normalized (diagonalize.hpp:427-431
        // normalizes, then projects out one trial at a time).  Here: classical Gram-Schmidt applied twice, all m
        // projections of a pass in one Gram launch and one update, on the device -- the same direction up to
        // rounding, with no host round trip per trial
        launch_precondition(errv, dvals.as<double>(), dmask.as<unsigned char>(), eval, n, s);
        for (int pass = 0; pass < 2; ++pass) {
          launch_gram(T(0), n, m, errv, n, 1, n, gram.as<double>() + kProj, part.as<double>(), npart, s);
          launch_combine_sub(T(0), n, m, gram.as<double>() + kProj, errv, n, s);
        }
        dnormalize(errv);
        hip_check(hipMemcpyAsync(T(m), errv, n * sizeof(double), hipMemcpyDeviceToDevice, s), "trial");
        multH(T(m), Sg(m));
      }
    }
    ++niter;
    }
  };
  if (prec == MPSXX_FP32 || prec == MPSXX_BF16X3) {
    // mixed precision: FP32 (or split-BF16) sigma vectors with FP32 accumulation inside each block GEMM until their
    // noise floor, then FP64 refinement from that vector; Gram-Schmidt, the subspace problem and residuals stay FP64.
    // The first application is FP64: a start vector already within the low-precision noise floor (converged sweeps)
    // is refined in FP64 directly, without a low-precision phase that could not improve it.
    const double tol_lo = std::max(tol, prec == MPSXX_FP32 ? kMixedTolerance : kMixedToleranceX3);
    use_plans(MPSXX_FP64);
    dnormalize(T(0));
    multH(T(0), Sg(0));
    double r0 = 0.0;
    {
      const double e0 = dot(T(0), Sg(0));
      launch_residual(Sg(0), T(0), e0, errv, n, gram.as<double>() + kNrm, part.as<double>(), npart, s);
      hip_check(hipMemcpyAsync(&r0, gram.as<double>() + kNrm, sizeof(double), hipMemcpyDeviceToHost, s), "r0");
      hip_check(hipStreamSynchronize(s), "r0 sync");
    }
    if (r0 < tol_lo) {
      solve(MPSXX_FP64, tol, true);
    } else {
      const int64_t n64 = st.nsigma;
      solve(prec, tol_lo, false);
      st.nsigma_fp32 = st.nsigma - n64;
      solve(MPSXX_FP64, tol, false);
    }
  } else {
    solve(MPSXX_FP64, tol, false);
  }
  // wfnc <- trial_0 orthogonalized against the lower roots' S_k and normalized (Functor::orthonormalize)
  for (int k = 0; k < nlow; ++k) axpy(-dot(Sk(k), T(0)), Sk(k), T(0));
  normalize(T(0));
  wfnc.set_layout(U, s);
  hip_check(hipMemcpyAsync(wfnc.ptr(), T(0), n * sizeof(double), hipMemcpyDeviceToDevice, s), "result");
  hip_check(hipStreamSynchronize(s), "davidson end");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return eval;
}

}  // namespace mpx

