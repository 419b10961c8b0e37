// Batched Davidson: many independent two-site eigenproblems solved in lockstep (reference per instance:
// src/diagonalize.hpp:283-449, with the Functor of one operator group, :65-80).
//
// Every instance keeps the reference's algorithm (<= 10 Ritz vectors, <= 4 macro iterations, ||r||^2 < tol, the
// preconditioner's clamp, Gram-Schmidt of the preconditioned residual) on its own vector space U_i (the closure of its
// wavefunction's blocks under the sigma pattern map, as in davidson.cpp).  The spaces are laid end to end in one batch
// vector, so each Davidson step is one launch over all instances: the sigma vectors of all instances are one
// BatchSigma application (one grouped launch per contraction step, replayed from a hipGraph), and the dot products,
// Ritz rotation, residual, preconditioner and Gram-Schmidt are segmented kernels over the batch vector.  The host
// solves the instances' subspace problems (<= 10 x 10 each) from one read-back per step.
//
// Instances converge at different steps.  A converged instance's vector and Ritz value are saved at that step and its
// subspace is frozen (identity rotation) while the others go on; its sigma vectors are still computed with the batch
// but are not counted as its work.  Mixed precision follows davidson.cpp: a first FP64 sigma decides per instance
// whether the FP32 (or split-BF16x3) phase runs, then every instance is refined in FP64 from its phase-1 vector.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "batch.hpp"
#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {

namespace {

// Symmetric eigenproblem of the Davidson subspace (cyclic Jacobi; ascending eigenvalues, eigenvectors in columns,
// row-major n x n), as davidson.cpp solves it.
void subspace_eig(int n, std::vector<double> a, std::vector<double>& w, std::vector<double>& v) {
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
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x * n + x] < a[y * n + y]; });
  w.resize(n);
  std::vector<double> v2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = a[ord[j] * n + ord[j]];
    for (int i = 0; i < n; ++i) v2[i * n + j] = v[i * n + ord[j]];
  }
  v.swap(v2);
}

// Structure-only view of a layout for BatchSigma: the vector data lives in the batch vector.
std::unique_ptr<Tensor> structure_of(const Layout& L) {
  std::unique_ptr<Tensor> t(new Tensor());
  t->L = L;
  t->empty = false;
  return t;
}

}  // namespace

void batch_davidson_twosite(std::vector<BatchDavInstance>& inst, double tol, int prec, bool use_graph,
                            BatchDavidsonStats& st, hipStream_t s) {
  const auto t_start = std::chrono::steady_clock::now();
  const int n = (int)inst.size();
  if (n <= 0) throw Error(MPSXX_ERR_ARG, "batch Davidson: no instances");
  if (prec != MPSXX_FP64 && prec != MPSXX_FP32 && prec != MPSXX_BF16X3)
    throw Error(MPSXX_ERR_ARG, "batch Davidson: unknown precision");
  constexpr int max_ritz = 10, kMacro = 4;
  constexpr int gs = max_ritz * max_ritz;  // per-instance stride of the Gram results and coefficient matrices

  // vector spaces, laid end to end (each segment starts on a 64-element boundary)
  std::vector<Layout> U(n);
  std::vector<int64_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const BatchDavInstance& b = inst[i];
    for (const Tensor* t : {b.lmpo, b.rmpo, b.lopr, b.ropr, b.diag, b.wfnc})
      if (!t) throw Error(MPSXX_ERR_ARG, "batch Davidson: missing array");
    for (const Tensor* t : {b.lmpo, b.rmpo, b.lopr, b.ropr, b.wfnc})
      if (t->empty) throw Error(MPSXX_ERR_ARG, "batch Davidson: an input array is empty");
    if (b.wfnc->L.rank != 4) throw Error(MPSXX_ERR_ARG, "batch Davidson: two-site wavefunctions (rank 4) expected");
    U[i] = davidson_layout({DavGroup{b.lmpo, b.rmpo, b.lopr, b.ropr}}, b.wfnc->L);
    if (U[i].nelem == 0) throw Error(MPSXX_ERR_ARG, "batch Davidson: empty vector space");
    off[i + 1] = off[i] + ((U[i].nelem + 63) / 64) * 64;
  }
  const int64_t N = off[n];
  std::vector<SegChunk> chunks;
  std::vector<int32_t> chunk0(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    chunk0[i] = (int32_t)chunks.size();
    for (int64_t b = 0; b < U[i].nelem; b += kSegChunk)
      chunks.push_back({off[i] + b, (int32_t)std::min<int64_t>(kSegChunk, U[i].nelem - b), i});
  }
  chunk0[n] = (int32_t)chunks.size();
  const int nch = (int)chunks.size();
  DevBuf dch = upload_vec(chunks), dc0 = upload_vec(chunk0);
  const SegChunk* CH = dch.as<SegChunk>();
  const int32_t* C0 = dc0.as<int32_t>();

  // vector slots: T[0..9], S[0..9], T2[0..9], S2[0..9], errv, result
  const int nslot = 4 * max_ritz + 2;
  DevBuf V((size_t)nslot * N * sizeof(double));
  double* base = V.as<double>();
  double* tb = base;
  double* sb = base + (int64_t)max_ritz * N;
  double* t2b = base + (int64_t)(2 * max_ritz) * N;
  double* s2b = base + (int64_t)(3 * max_ritz) * N;
  double* errv = base + (int64_t)(4 * max_ritz) * N;
  double* res = errv + N;
  auto T = [&](int k) { return tb + (int64_t)k * N; };
  auto Sg = [&](int k) { return sb + (int64_t)k * N; };
  hip_check(hipMemsetAsync(base, 0, (size_t)nslot * N * sizeof(double), s), "memset");

  // start vectors and diagonals placed into their segments (one copy launch per instance)
  DevBuf dvals(N * sizeof(double)), dmask(N);
  hip_check(hipMemsetAsync(dvals.p, 0, N * sizeof(double), s), "memset");
  std::vector<unsigned char> mask_h(N, 0);
  for (int i = 0; i < n; ++i) {
    const Tensor& w = *inst[i].wfnc;
    std::vector<CopySeg> segs;
    for (int64_t b = 0; b < w.L.nblocks(); ++b)
      if (w.L.bsize[b] > 0)
        segs.push_back({off[i] + U[i].off[U[i].find(w.L.tag[b])], w.L.off[b], (int32_t)w.L.bsize[b], 1});
    copy_segments(segs, w.ptr(), T(0), s);
    const Tensor& d = *inst[i].diag;
    if (d.empty) continue;
    segs.clear();
    for (int64_t b = 0; b < U[i].nblocks(); ++b) {
      const int64_t j = d.L.find(U[i].tag[b]);
      if (j < 0) continue;
      if (d.L.bsize[j] != U[i].bsize[b]) throw Error(MPSXX_ERR_SHAPE, "batch Davidson: diag block shape mismatch");
      if (U[i].bsize[b] > 0) segs.push_back({off[i] + U[i].off[b], d.L.off[j], (int32_t)U[i].bsize[b], 1});
      std::memset(mask_h.data() + off[i] + U[i].off[b], 1, U[i].bsize[b]);
    }
    copy_segments(segs, d.ptr(), dvals.as<double>(), s);
  }
  hip_check(hipMemcpyAsync(dmask.p, mask_h.data(), N, hipMemcpyHostToDevice, s), "mask");
  hip_check(hipStreamSynchronize(s), "mask sync");

  // per-instance device scalars
  DevBuf G((size_t)n * gs * sizeof(double)), proj((size_t)n * max_ritz * sizeof(double)), nrm(n * sizeof(double));
  DevBuf coef((size_t)n * gs * sizeof(double)), evd(n * sizeof(double)), part((size_t)nch * gs * sizeof(double));
  DevBuf savemask(n * sizeof(int32_t));
  std::vector<double> G_h((size_t)n * gs), coef_h((size_t)n * gs, 0.0), ev_h(n, 0.0), nrm_h(n, 0.0);
  auto dnormalize = [&](double* x) {
    launch_seg_gram(CH, nch, C0, n, x, N, 1, x, N, 1, nrm.as<double>(), 1, part.as<double>(), s);
    launch_seg_scale_rnorm(CH, nch, x, nrm.as<double>(), 1, s);
  };
  auto read_norms = [&]() {
    hip_check(hipMemcpyAsync(nrm_h.data(), nrm.p, n * sizeof(double), hipMemcpyDeviceToHost, s), "norms");
    hip_check(hipStreamSynchronize(s), "norms sync");
  };
  // save the current trial_0 of the instances in `which` as their result
  std::vector<int32_t> save_h(n);
  auto save = [&](const std::vector<char>& which) {
    bool any = false;
    for (int i = 0; i < n; ++i) {
      save_h[i] = which[i] ? 1 : 0;
      any = any || which[i];
    }
    if (!any) return;
    hip_check(hipMemcpyAsync(savemask.p, save_h.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s), "mask");
    launch_seg_copy(CH, nch, T(0), res, savemask.as<int32_t>(), s);
    hip_check(hipStreamSynchronize(s), "save sync");  // save_h is reused by the next call
  };

  // the batched sigma operators: instance structures are described by structure-only tensors of U_i
  std::vector<std::unique_ptr<Tensor>> shapes;
  for (int i = 0; i < n; ++i) shapes.push_back(structure_of(U[i]));
  std::vector<BatchSigma::Operand6> ops;
  for (int i = 0; i < n; ++i)
    ops.push_back({inst[i].lmpo, inst[i].rmpo, inst[i].lopr, inst[i].ropr, shapes[i].get(), shapes[i].get()});
  std::unique_ptr<BatchSigma> b64(new BatchSigma(ops, MPSXX_FP64, s)), blo;
  if (prec != MPSXX_FP64) blo.reset(new BatchSigma(ops, prec, s));
  unsigned conv_lo = ~0u;  // the low-precision batch converts its operators on its first application only

  st.energy.assign(n, 0.0);
  st.rnorm.assign(n, 0.0);
  st.nsigma.assign(n, 0);
  st.nsigma_low.assign(n, 0);
  st.converged.assign(n, 0);
  // event pairs around every batched sigma application, read once the solve has ended (no host wait per sigma)
  std::vector<hipEvent_t> events;
  size_t nev = 0;
  auto event = [&]() {
    if (nev == events.size()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "event");
      events.push_back(e);
    }
    return events[nev++];
  };
  std::vector<const double*> win(n);
  std::vector<double*> sout(n);
  std::vector<char> active(n, 1);
  auto multH = [&](int k, bool low) {
    BatchSigma& B = low ? *blo : *b64;
    hip_check(hipEventRecord(event(), s), "event");
    hip_check(hipMemsetAsync(Sg(k), 0, N * sizeof(double), s), "memset");
    for (int i = 0; i < n; ++i) {
      win[i] = T(k) + off[i];
      sout[i] = Sg(k) + off[i];
    }
    B.set_vectors(win, sout);
    B.run(low ? conv_lo : 0u, use_graph, s, false);  // inputs checked when the batch was built, fixed since
    if (low) conv_lo = 1u << 4;  // from now on only the trial vectors change
    hip_check(hipEventRecord(event(), s), "event");
    ++st.batch_sigmas;
    if (low) ++st.batch_sigmas_low;
    for (int i = 0; i < n; ++i)
      if (active[i]) {
        ++st.nsigma[i];
        if (low) ++st.nsigma_low[i];
      }
  };

  // one lockstep Davidson over the instances flagged in `active` (cleared as they converge); every instance's result
  // (converged or not) is in `res` afterwards, and its last Ritz value in st.energy
  auto solve = [&](bool low, double tol_p) {
    std::vector<char> newly(n, 0);
    int niter = 0;
    auto any_active = [&] {
      for (int i = 0; i < n; ++i)
        if (active[i]) return true;
      return false;
    };
    while (any_active() && niter < kMacro) {
      dnormalize(T(0));
      multH(0, low);
      for (int m = 1; m <= max_ritz; ++m) {
        // heff(i,j) = <trial_i|sigma_j> for every instance at once (diagonalize.hpp:339-365)
        launch_seg_gram(CH, nch, C0, n, T(0), N, m, Sg(0), N, m, G.as<double>(), gs, part.as<double>(), s);
        hip_check(hipMemcpyAsync(G_h.data(), G.p, (size_t)n * gs * sizeof(double), hipMemcpyDeviceToHost, s), "gram");
        hip_check(hipStreamSynchronize(s), "gram sync");
        auto subspace = [&](int i) {
          double* c = coef_h.data() + (size_t)i * gs;
          std::fill(c, c + gs, 0.0);
          if (!active[i]) {
            for (int j = 0; j < m; ++j) c[j * max_ritz + j] = 1.0;  // frozen: identity rotation
            return;
          }
          std::vector<double> heff((size_t)m * m), w, R;
          const double* g = G_h.data() + (size_t)i * gs;
          for (int a = 0; a < m; ++a)
            for (int b = 0; b <= a; ++b) heff[a * m + b] = heff[b * m + a] = g[a * m + b];
          subspace_eig(m, heff, w, R);
          ev_h[i] = w[0];
          st.energy[i] = w[0];
          // trial_j <- sum_k rvec(k,j) trial_k (diagonalize.hpp:378-394)
          for (int j = 0; j < m; ++j)
            for (int k = 0; k < m; ++k) c[j * max_ritz + k] = R[k * m + j];
        };
        // the instances' subspace problems on several host threads (each writes only its own slots)
        const int nth = std::max(1, std::min(host_threads(), n / 8));
        parallel_for(nth, [&](int t) {
          for (int i = t; i < n; i += nth) subspace(i);
        });
        hip_check(hipMemcpyAsync(coef.p, coef_h.data(), (size_t)n * gs * sizeof(double), hipMemcpyHostToDevice, s),
                  "coef");
        hip_check(hipMemcpyAsync(evd.p, ev_h.data(), n * sizeof(double), hipMemcpyHostToDevice, s), "evals");
        launch_seg_rotate2(CH, nch, T(0), Sg(0), N, m, coef.as<double>(), max_ritz, gs, t2b, s2b, s);
        std::swap(tb, t2b);
        std::swap(sb, s2b);
        // r = sigma_0 - E trial_0 and ||r||^2 < tol (diagonalize.hpp:416-422)
        launch_seg_residual(CH, nch, C0, n, Sg(0), T(0), evd.as<double>(), errv, nrm.as<double>(), 1,
                            part.as<double>(), s);
        read_norms();  // also orders the host arrays' reuse after the copies above
        bool left = false;
        for (int i = 0; i < n; ++i) {
          newly[i] = 0;
          if (!active[i]) continue;
          st.rnorm[i] = nrm_h[i];
          if (nrm_h[i] < tol_p) {
            newly[i] = 1;
            active[i] = 0;
            st.converged[i] = 1;
          } else {
            left = true;
          }
        }
        save(newly);
        if (!left) break;
        if (m < max_ritz) {
          // preconditioned residual, classical Gram-Schmidt twice against the trials, normalized (diagonalize.hpp:
          // 427-431; the same direction up to rounding, as davidson.cpp)
          launch_seg_precondition(CH, nch, errv, dvals.as<double>(), dmask.as<unsigned char>(), evd.as<double>(), s);
          for (int pass = 0; pass < 2; ++pass) {
            launch_seg_gram(CH, nch, C0, n, T(0), N, m, errv, N, 1, proj.as<double>(), max_ritz, part.as<double>(),
                            s);
            launch_seg_combine_sub(CH, nch, T(0), N, m, proj.as<double>(), max_ritz, errv, s);
          }
          dnormalize(errv);
          hip_check(hipMemcpyAsync(T(m), errv, N * sizeof(double), hipMemcpyDeviceToDevice, s), "trial");
          multH(m, low);
        }
      }
      ++niter;
      ++st.macro_iterations;
    }
    // instances that ran out of iterations keep their last trial_0 (the reference returns it as well)
    std::vector<char> rest(n, 0);
    for (int i = 0; i < n; ++i) rest[i] = active[i];
    save(rest);
    for (int i = 0; i < n; ++i) active[i] = 0;
  };

  if (prec != MPSXX_FP64) {
    // first sigma in FP64; instances whose start vector is already within the low-precision noise floor skip the
    // low-precision phase (davidson.cpp)
    const double tol_lo = std::max(tol, prec == MPSXX_FP32 ? kMixedTolerance : kMixedToleranceX3);
    dnormalize(T(0));
    multH(0, false);
    launch_seg_gram(CH, nch, C0, n, T(0), N, 1, Sg(0), N, 1, evd.as<double>(), 1, part.as<double>(), s);
    launch_seg_residual(CH, nch, C0, n, Sg(0), T(0), evd.as<double>(), errv, nrm.as<double>(), 1, part.as<double>(),
                        s);
    read_norms();
    hip_check(hipMemcpy(ev_h.data(), evd.p, n * sizeof(double), hipMemcpyDeviceToHost), "energies");
    std::vector<char> skip(n, 0);
    for (int i = 0; i < n; ++i) {
      active[i] = nrm_h[i] >= tol_lo;
      skip[i] = !active[i];
    }
    save(skip);
    solve(true, tol_lo);
    // FP64 refinement of every instance from its low-precision result
    launch_seg_copy(CH, nch, res, T(0), nullptr, s);
    std::fill(active.begin(), active.end(), 1);
    std::fill(st.converged.begin(), st.converged.end(), 0);
    solve(false, tol);
  } else {
    solve(false, tol);
  }

  // results: normalized and placed into each wfnc on U_i
  dnormalize(res);
  for (int i = 0; i < n; ++i) {
    Tensor& w = *inst[i].wfnc;
    w.set_layout(U[i], s);
    hip_check(hipMemcpyAsync(w.ptr(), res + off[i], U[i].nelem * sizeof(double), hipMemcpyDeviceToDevice, s),
              "result");
  }
  hip_check(hipStreamSynchronize(s), "batch davidson end");
  for (size_t k = 0; k + 1 < nev; k += 2) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, events[k], events[k + 1]), "elapsed");
    st.sigma_ms += ms;
  }
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace mpx
