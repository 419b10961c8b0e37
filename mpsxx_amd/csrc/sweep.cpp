// Device-resident two-site DMRG sweeps (reference: dmrg/make_sweep.cpp:40-254 with src/optimize.hpp:164-265 and the
// random-MPS initialisation of src/make_random_mpss.cpp:97-272).
//
// Everything stays in HBM across the sweep: MPO site tensors, MPS site tensors, every left/right environment.  The
// reference instead writes each object to a per-site Boost archive and reloads it (src/fileio.h:9-25).  Per site:
// two-site wavefunction (Gemm), diagonal, Davidson (tolerance 0.1*T as make_sweep.cpp:112 passes), canonicalize with
// truncation to M, renormalize.  The sweep energy is the minimum of the site energies (make_sweep.cpp:115).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "fileio.hpp"
#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {

struct SweepStats {
  double seconds = 0.0, sigma_ms = 0.0, davidson_s = 0.0, svd_s = 0.0, renorm_s = 0.0, diag_s = 0.0;
  int nsigma = 0, max_bond = 0, nsigma_fp32 = 0;
  double dweight = 0.0;
};

// A converged lower root k for excited states (dmrg/make_sweep.cpp with iroot > 0, dmrg/optimize.hpp:127-230): its
// two-site wavefunction at every bond, its left- and right-canonical site tensors, and the cross environments of the
// current root against it, <cur|H|k> (lH, rH) and <cur|k> (lS, rS), renormalized with bra = current root and
// ket = root k as the sweep moves.
struct LowerRoot {
  double energy = 0.0;
  std::vector<std::unique_ptr<Tensor>> A, B, psi, C;  // C[i]: one-site wavefunction with the centre at site i
  std::vector<std::unique_ptr<Tensor>> lH, rH, lS, rS;
};

// One Davidson solve of a sweep: the site pair (site, site + 1) or one-site centre, its direction, the eigenvalue and
// the lowest Ritz value after every subspace solve (the reference prints them as micro iterations).
struct StepTrace {
  int site = 0, forward = 1;
  double energy = 0.0;
  std::vector<double> evals, rnorms;
};

struct Dmrg {
  int L = 0;
  std::vector<StepTrace> trace;  // the Davidson solves of the last sweep, in order
  std::vector<Tensor*> mpo;                 // borrowed
  std::vector<std::unique_ptr<Tensor>> mps; // site tensors (A left of the centre, wave at the centre, B right)
  std::vector<std::unique_ptr<Tensor>> lenv, renv;  // lenv[i]: left of site i; renv[i]: right of site i-1
  int center = 0;
  int form = MPSXX_SWEEP_FOURINDEX;  // two-site 4-index, merged (src/make_sweep.cpp:112) or one-site
  std::vector<LowerRoot> lower;
  double noise = 0.0;                   // DMRGInput::noise (src/input.h:43), passed to every canonicalize
  int trunc_rule = MPSXX_TRUNCATION_STABLE;
};

namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

std::unique_ptr<Tensor> boundary_env(hipStream_t s) {
  Layout L;
  L.set_shape(3, true, QN(), {{QN()}, {QN()}, {QN()}}, {{1}, {1}, {1}});
  int32_t z[3] = {0, 0, 0};
  L.add_block(z);
  L.finalize();
  auto t = std::make_unique<Tensor>();
  t->set_layout(L, s);
  launch_fill(t->ptr(), 1, 1.0, s);
  return t;
}

std::unique_ptr<Tensor> boundary_overlap(hipStream_t s) {
  Layout L;
  L.set_shape(2, true, QN(), {{QN()}, {QN()}}, {{1}, {1}});
  int32_t z[2] = {0, 0};
  L.add_block(z);
  L.finalize();
  auto t = std::make_unique<Tensor>();
  t->set_layout(L, s);
  launch_fill(t->ptr(), 1, 1.0, s);
  return t;
}

std::unique_ptr<Tensor> copy_of(Tensor& x, hipStream_t s) {
  auto t = std::make_unique<Tensor>();
  op_copy(x, *t, s);
  return t;
}

// two-site psi = Gemm(N, N, a(l,n,r), b(r,n',r')) (optimize.hpp:180-182)
void twosite(Tensor& a, Tensor& b, Tensor& psi, hipStream_t s) {
  psi = Tensor();
  op_contract(1.0, a, false, {2}, b, false, {0}, 1.0, psi, s);
}

Tensor shape_only(const Layout& L) {
  Layout e = L;
  e.tag.clear();
  e.bidx.clear();
  e.finalize();
  Tensor t;
  t.set_layout(e, nullptr);
  return t;
}

int bond_dim(const Layout& L, int idx) {
  int d = 0;
  for (int x : L.ds[idx]) d += x;
  return d;
}

}  // namespace

// right_canonical: init already has the centre at site 0 (B tensors right of it), as a saved sweep state has.
Dmrg* dmrg_create(int L, const std::vector<Tensor*>& mpo, const std::vector<Tensor*>& init, int M, hipStream_t s,
                  bool right_canonical = false) {
  if (L < 2 || (int)mpo.size() != L || (int)init.size() != L) throw Error(MPSXX_ERR_ARG, "dmrg: need L >= 2 sites");
  auto D = std::make_unique<Dmrg>();
  D->L = L;
  D->mpo = mpo;
  D->mps.resize(L);
  for (int i = 0; i < L; ++i) {
    D->mps[i] = std::make_unique<Tensor>();
    op_copy(*init[i], *D->mps[i], s);
  }
  // right-canonicalise from the right end: psi = T_{i-1} T_i, RightArrow SVD -> (U s, B_i)
  // (make_random_mpss.cpp:204-251 does the same with one-site canonicalize(0, ...))
  for (int i = right_canonical ? 0 : L - 1; i >= 1; --i) {
    Tensor psi, U, B;
    twosite(*D->mps[i - 1], *D->mps[i], psi, s);
    op_canonicalize_twosite(false, psi, U, B, M, s);
    const double nrm = op_nrm2(U, s);
    op_scal(1.0 / nrm, U, s);
    *D->mps[i - 1] = std::move(U);
    *D->mps[i] = std::move(B);
  }
  D->lenv.resize(L + 1);
  D->renv.resize(L + 1);
  D->lenv[0] = boundary_env(s);
  D->renv[L] = boundary_env(s);
  for (int i = L - 1; i >= 1; --i) {
    D->renv[i] = std::make_unique<Tensor>();
    op_renormalize(false, *mpo[i], *D->renv[i + 1], *D->mps[i], *D->mps[i], *D->renv[i], MPSXX_FP64, s);
  }
  D->center = 0;
  hip_check(hipStreamSynchronize(s), "dmrg init");
  return D.release();
}

constexpr char kDmrgRecord[9] = "MPSXXDMR";

// The state as site files: "lmps" left of the centre, "wave" at it, "rmps" right of it (src/make_sweep.cpp:124-247,
// src/make_random_mpss.cpp:155-217 name them so), then the record {nsites, centre}, written last.
void dmrg_save(const Dmrg& D, const std::string& prefix, int iroot, hipStream_t s) {
  for (int k = 0; k < D.L; ++k) {
    const char* name = k < D.center ? "lmps" : (k == D.center ? "wave" : "rmps");
    tensor_save(*D.mps[k], getfile(name, prefix, k, iroot, 0), s);
  }
  write_record(getfile("dmrg", prefix, 0, iroot, 0), kDmrgRecord, {D.L, D.center});
}

Dmrg* dmrg_load(int L, const std::vector<Tensor*>& mpo, const std::string& prefix, int iroot, int M, hipStream_t s) {
  const std::vector<int32_t> rec = read_record(getfile("dmrg", prefix, 0, iroot, 0), kDmrgRecord, 2);
  if (rec[0] != L)
    throw Error(MPSXX_ERR_ARG, "dmrg_load: the saved state has " + std::to_string(rec[0]) + " sites, not " +
                                   std::to_string(L));
  const int center = rec[1];
  if (center < 0 || center >= L) throw Error(MPSXX_ERR_IO, "dmrg_load: centre out of range in the state record");
  std::vector<Tensor> site(L);
  std::vector<Tensor*> p(L);
  for (int k = 0; k < L; ++k) {
    const char* name = k < center ? "lmps" : (k == center ? "wave" : "rmps");
    const std::string f = getfile(name, prefix, k, iroot, 0);
    tensor_load(f, site[k], s);
    if (site[k].empty || site[k].L.rank != 3) throw Error(MPSXX_ERR_IO, f + ": not a 3-index site tensor");
    p[k] = &site[k];
  }
  return dmrg_create(L, mpo, p, M, s, center == 0);
}

// One two-site step in the merged form (src/optimize.hpp:495-630, optimize_twosite_merged): the merged block
// operators L~ = merge(L W_i) and R~ = merge(W_{i+1} R) are built once (compute_merged_operator, :273-368), psi is
// merged to psi~((l,n1),(n2,r)), Davidson runs on the merged sigma vector (two GEMM families), the SVD splits psi~,
// and the environment of the next site is U~^+ L~ U~ (renormalize of the merged operator, renormalize.hpp:57-74).
// The merged-operator build is counted in renorm_s.
void optimize_merged(Dmrg& D, int i, bool forward, Tensor& psi, int M, double tol, int prec, SweepStats& st,
                     double& emin, hipStream_t s) {
  // MPSXX_SWEEP_TRACE: every phase is synchronized and timed on its own, one line per step on stderr
  static const bool trace = std::getenv("MPSXX_SWEEP_TRACE") != nullptr;
  double ph[8] = {0};
  auto lap = [&](clk::time_point& t, int k) {
    if (!trace) return;
    hip_check(hipStreamSynchronize(s), "trace");
    ph[k] = 1e3 * since(t);
    t = clk::now();
  };
  auto t = clk::now();
  auto tp = clk::now();
  Tensor Lm, Rm, wm;
  op_merged_operator(true, *D.mpo[i], *D.lenv[i], Lm, s);
  lap(tp, 0);
  op_merged_operator(false, *D.mpo[i + 1], *D.renv[i + 2], Rm, s);
  lap(tp, 1);
  op_qst_merge(psi, 2, 2, wm, s);
  st.renorm_s += since(t);
  t = clk::now();
  tp = clk::now();
  Tensor diag = shape_only(wm.L);
  op_diag_merged(Lm, Rm, diag, s);
  hip_check(hipStreamSynchronize(s), "diag");
  st.diag_s += since(t);
  ph[2] = 1e3 * since(tp);
  t = clk::now();
  DavidsonStats ds;
  std::vector<DavGroup> g{{nullptr, nullptr, &Lm, &Rm}};
  const double e = davidson_twosite(g, diag, wm, 0.1 * tol, prec, ds, s);
  D.trace.push_back({i, forward ? 1 : 0, e, ds.evals, ds.rnorms});
  st.davidson_s += since(t);
  ph[3] = 1e3 * since(t);
  st.nsigma += ds.nsigma;
  st.nsigma_fp32 += ds.nsigma_fp32;
  st.sigma_ms += ds.sigma_ms;
  if (e < emin) emin = e;
  t = clk::now();
  Tensor U, V;
  op_make_noise(wm, D.noise, s);  // canonicalize of the merged wavefunction (canonicalize.hpp:88-106)
  st.dweight += op_canonicalize_merged(forward, wm, psi.L, U, V, M, s);
  st.svd_s += since(t);
  ph[4] = 1e3 * since(t);
  st.max_bond = std::max(st.max_bond, bond_dim(U.L, 2));
  t = clk::now();
  tp = clk::now();
  Tensor Xm;
  if (forward) op_qst_merge(U, 2, 1, Xm, s);  // U~((l,n1), s)
  else op_qst_merge(V, 1, 2, Xm, s);          // V~(s, (n2,r))
  lap(tp, 5);
  std::unique_ptr<Tensor>& env = forward ? D.lenv[i + 1] : D.renv[i + 1];
  env = std::make_unique<Tensor>();
  op_renormalize_merged(forward, forward ? Lm : Rm, Xm, Xm, *env, MPSXX_FP64, s);
  const int64_t rflops = last_stats().flops;
  hip_check(hipStreamSynchronize(s), "renormalize");
  st.renorm_s += since(t);
  if (trace) {
    ph[6] = 1e3 * since(tp);
    std::fprintf(stderr,
                 "[sweep] site %d %s merge %.2f %.2f diag %.2f dav %.2f (%d sigma) svd %.2f qst %.2f renorm %.2f ms "
                 "(%.2f GFLOP, %.1f TFLOP/s) bond %d\n",
                 i, forward ? "fwd" : "bwd", ph[0], ph[1], ph[2], ph[3], ds.nsigma, ph[4], ph[5], ph[6],
                 rflops * 1e-9, ph[6] > 0 ? rflops * 1e-9 / ph[6] : 0.0, bond_dim(U.L, 2));
  }
  *D.mps[i] = std::move(U);
  *D.mps[i + 1] = std::move(V);
}

// One forward + one backward one-site sweep (dmrg/make_sweep.cpp with ONESITE, optimize_onesite
// dmrg/optimize.hpp:20-125, noise 0): the wave at site i is optimized with the one-site sigma vector and diagonal,
// split by the one-site SVD (LeftArrow forward, RightArrow backward, truncation to M) and the 2-index factor is
// multiplied into the next site, which becomes the wave.  Forward over sites 0..L-2, backward over L-1..1; the centre
// ends at site 0.
double dmrg_sweep_onesite(Dmrg& D, int M, double tol, int prec, SweepStats& st, hipStream_t s) {
  if (D.noise >= 1.0e-8)
    throw Error(MPSXX_ERR_UNSUPPORTED, "dmrg: noise is not supported in one-site sweeps (canonicalize.hpp:57-66 "
                                       "turns them into two-site decompositions)");
  TruncationScope trunc(D.trunc_rule);
  D.trace.clear();
  const int L = D.L;
  const auto t0 = clk::now();
  double emin = 1.0e8;
  auto optimize = [&](int i, bool forward) {
    Tensor psi;
    op_copy(*D.mps[i], psi, s);
    auto t = clk::now();
    Tensor diag = shape_only(psi.L);
    op_diag_onesite(*D.mpo[i], *D.lenv[i], *D.renv[i + 1], diag, s);
    hip_check(hipStreamSynchronize(s), "diag");
    st.diag_s += since(t);
    t = clk::now();
    DavidsonStats ds;
    std::vector<DavGroup> g{{D.mpo[i], nullptr, D.lenv[i].get(), D.renv[i + 1].get()}};
    const size_t nl = D.lower.size();
    std::vector<Tensor> Hk(nl), Sk(nl);
    std::vector<DavLowerRoot> low;
    for (size_t k = 0; k < nl; ++k) {
      LowerRoot& R = D.lower[k];
      op_sigma_onesite(*D.mpo[i], *R.lH[i], *R.rH[i + 1], *R.C[i], Hk[k], MPSXX_FP64, s);
      op_sigma_overlap(*R.lS[i], *R.rS[i + 1], *R.C[i], Sk[k], s);
      low.push_back({&Hk[k], &Sk[k], R.energy});
    }
    const double e = davidson_twosite(g, diag, psi, 0.1 * tol, prec, ds, s, low);
    D.trace.push_back({i, forward ? 1 : 0, e, ds.evals, ds.rnorms});
    st.davidson_s += since(t);
    st.nsigma += ds.nsigma;
    st.nsigma_fp32 += ds.nsigma_fp32;
    st.sigma_ms += ds.sigma_ms;
    if (e < emin) emin = e;
    t = clk::now();
    Tensor U, V;
    st.dweight += op_canonicalize_onesite(forward, psi, U, V, M, s);
    st.svd_s += since(t);
    t = clk::now();
    if (forward) {
      st.max_bond = std::max(st.max_bond, bond_dim(U.L, 2));
      Tensor next;
      op_contract(1.0, V, false, {1}, *D.mps[i + 1], false, {0}, 1.0, next, s);  // s*Vt times the next site
      *D.mps[i] = std::move(U);
      *D.mps[i + 1] = std::move(next);
      D.lenv[i + 1] = std::make_unique<Tensor>();
      op_renormalize(true, *D.mpo[i], *D.lenv[i], *D.mps[i], *D.mps[i], *D.lenv[i + 1], MPSXX_FP64, s);
      for (LowerRoot& R : D.lower) {
        R.lH[i + 1] = std::make_unique<Tensor>();
        op_renormalize(true, *D.mpo[i], *R.lH[i], *D.mps[i], *R.A[i], *R.lH[i + 1], MPSXX_FP64, s);
        R.lS[i + 1] = std::make_unique<Tensor>();
        op_renormalize_overlap(true, *R.lS[i], *D.mps[i], *R.A[i], *R.lS[i + 1], s);
      }
    } else {
      st.max_bond = std::max(st.max_bond, bond_dim(V.L, 0));
      Tensor prev;
      op_contract(1.0, *D.mps[i - 1], false, {2}, U, false, {0}, 1.0, prev, s);  // previous site times U*s
      *D.mps[i] = std::move(V);
      *D.mps[i - 1] = std::move(prev);
      D.renv[i] = std::make_unique<Tensor>();
      op_renormalize(false, *D.mpo[i], *D.renv[i + 1], *D.mps[i], *D.mps[i], *D.renv[i], MPSXX_FP64, s);
      for (LowerRoot& R : D.lower) {
        R.rH[i] = std::make_unique<Tensor>();
        op_renormalize(false, *D.mpo[i], *R.rH[i + 1], *D.mps[i], *R.B[i], *R.rH[i], MPSXX_FP64, s);
        R.rS[i] = std::make_unique<Tensor>();
        op_renormalize_overlap(false, *R.rS[i + 1], *D.mps[i], *R.B[i], *R.rS[i], s);
      }
    }
    hip_check(hipStreamSynchronize(s), "renormalize");
    st.renorm_s += since(t);
  };
  for (int i = 0; i < L - 1; ++i) optimize(i, true);
  for (int i = L - 1; i >= 1; --i) optimize(i, false);
  D.center = 0;
  st.seconds += since(t0);
  return emin;
}

// One forward + one backward two-site sweep; returns the minimum site energy.
double dmrg_sweep(Dmrg& D, int M, double tol, int prec, SweepStats& st, hipStream_t s) {
  TruncationScope trunc(D.trunc_rule);
  D.trace.clear();
  const int L = D.L;
  if (D.form == MPSXX_SWEEP_MERGED && !D.lower.empty())
    throw Error(MPSXX_ERR_UNSUPPORTED, "dmrg: lower roots are not supported with the merged sweep form");
  if (D.form == MPSXX_SWEEP_ONESITE) return dmrg_sweep_onesite(D, M, tol, prec, st, s);
  const auto t0 = clk::now();
  double emin = 1.0e8;
  auto optimize = [&](int i, bool forward) {
    // sites (i, i+1); forward: mps[i] is the wave, mps[i+1] right-canonical; backward: mps[i] left, mps[i+1] wave
    Tensor psi;
    twosite(*D.mps[i], *D.mps[i + 1], psi, s);
    if (D.form == MPSXX_SWEEP_MERGED) {
      optimize_merged(D, i, forward, psi, M, tol, prec, st, emin, s);
      return;
    }
    auto t = clk::now();
    Tensor diag = shape_only(psi.L);
    op_diag_twosite(*D.mpo[i], *D.mpo[i + 1], *D.lenv[i], *D.renv[i + 2], diag, s);
    hip_check(hipStreamSynchronize(s), "diag");
    st.diag_s += since(t);
    t = clk::now();
    DavidsonStats ds;
    std::vector<DavGroup> g{{D.mpo[i], D.mpo[i + 1], D.lenv[i].get(), D.renv[i + 2].get()}};
    // lower roots: H|k> and S|k> in the current root's basis (Davidson::Functor constructor, diagonalize.hpp:128-137)
    const size_t nl = D.lower.size();
    std::vector<Tensor> Hk(nl), Sk(nl);
    std::vector<DavLowerRoot> low;
    for (size_t k = 0; k < nl; ++k) {
      LowerRoot& R = D.lower[k];
      op_sigma_twosite(*D.mpo[i], *D.mpo[i + 1], *R.lH[i], *R.rH[i + 2], *R.psi[i], Hk[k], MPSXX_FP64, s);
      op_sigma_overlap(*R.lS[i], *R.rS[i + 2], *R.psi[i], Sk[k], s);
      low.push_back({&Hk[k], &Sk[k], R.energy});
    }
    const double e = davidson_twosite(g, diag, psi, 0.1 * tol, prec, ds, s, low);
    D.trace.push_back({i, forward ? 1 : 0, e, ds.evals, ds.rnorms});
    st.davidson_s += since(t);
    st.nsigma += ds.nsigma;
    st.nsigma_fp32 += ds.nsigma_fp32;
    st.sigma_ms += ds.sigma_ms;
    if (e < emin) emin = e;
    t = clk::now();
    Tensor U, V;
    op_make_noise(psi, D.noise, s);  // canonicalize.hpp:35
    st.dweight += op_canonicalize_twosite(forward, psi, U, V, M, s);
    st.svd_s += since(t);
    st.max_bond = std::max(st.max_bond, bond_dim(U.L, 2));
    *D.mps[i] = std::move(U);
    *D.mps[i + 1] = std::move(V);
    t = clk::now();
    if (forward) {
      D.lenv[i + 1] = std::make_unique<Tensor>();
      op_renormalize(true, *D.mpo[i], *D.lenv[i], *D.mps[i], *D.mps[i], *D.lenv[i + 1], MPSXX_FP64, s);
    } else {
      D.renv[i + 1] = std::make_unique<Tensor>();
      op_renormalize(false, *D.mpo[i + 1], *D.renv[i + 2], *D.mps[i + 1], *D.mps[i + 1], *D.renv[i + 1], MPSXX_FP64,
                     s);
    }
    // cross environments, bra = the current root's new site tensor, ket = root k's (dmrg/optimize.hpp:190-225)
    for (LowerRoot& R : D.lower) {
      if (forward) {
        R.lH[i + 1] = std::make_unique<Tensor>();
        op_renormalize(true, *D.mpo[i], *R.lH[i], *D.mps[i], *R.A[i], *R.lH[i + 1], MPSXX_FP64, s);
        R.lS[i + 1] = std::make_unique<Tensor>();
        op_renormalize_overlap(true, *R.lS[i], *D.mps[i], *R.A[i], *R.lS[i + 1], s);
      } else {
        R.rH[i + 1] = std::make_unique<Tensor>();
        op_renormalize(false, *D.mpo[i + 1], *R.rH[i + 2], *D.mps[i + 1], *R.B[i + 1], *R.rH[i + 1], MPSXX_FP64, s);
        R.rS[i + 1] = std::make_unique<Tensor>();
        op_renormalize_overlap(false, *R.rS[i + 2], *D.mps[i + 1], *R.B[i + 1], *R.rS[i + 1], s);
      }
    }
    hip_check(hipStreamSynchronize(s), "renormalize");
    st.renorm_s += since(t);
  };
  for (int i = 0; i < L - 1; ++i) optimize(i, true);
  for (int i = L - 2; i >= 0; --i) optimize(i, false);
  D.center = 0;
  st.seconds += since(t0);
  return emin;
}

namespace {

// Right cross environments of the current root (centre at site 0, right-canonical tensors) against lower root R.
void rebuild_cross_right(Dmrg& D, LowerRoot& R, hipStream_t s) {
  const int L = D.L;
  for (int i = L - 1; i >= 1; --i) {
    R.rH[i] = std::make_unique<Tensor>();
    op_renormalize(false, *D.mpo[i], *R.rH[i + 1], *D.mps[i], *R.B[i], *R.rH[i], MPSXX_FP64, s);
    R.rS[i] = std::make_unique<Tensor>();
    op_renormalize_overlap(false, *R.rS[i + 1], *D.mps[i], *R.B[i], *R.rS[i], s);
  }
  hip_check(hipStreamSynchronize(s), "lower root");
}

}  // namespace

// Register the converged state of `K` (centre at site 0, as a sweep leaves it) as a lower root of `D` (centre at site
// 0): a forward gauge sweep of K's copy without truncation gives its left-canonical tensors and its two-site
// wavefunction at every bond (right-canonical tensors are K's own), then the right cross environments of D against it
// are built from D's current right-canonical tensors, as make_random_mpss.cpp:176-236 builds them for iroot > 0.
void dmrg_add_lower_root(Dmrg& D, Dmrg& K, double energy, hipStream_t s) {
  const int L = D.L;
  if (K.L != L) throw Error(MPSXX_ERR_ARG, "dmrg: lower root has a different number of sites");
  if (D.center != 0 || K.center != 0) throw Error(MPSXX_ERR_ARG, "dmrg: both states must have the centre at site 0");
  if (D.form == MPSXX_SWEEP_MERGED)
    throw Error(MPSXX_ERR_UNSUPPORTED, "dmrg: lower roots are not supported with the merged sweep form");
  LowerRoot R;
  R.energy = energy;
  R.A.resize(L);
  R.B.resize(L);
  R.psi.resize(L - 1);
  R.C.resize(L);
  for (int i = 1; i < L; ++i) R.B[i] = copy_of(*K.mps[i], s);
  std::unique_ptr<Tensor> C = copy_of(*K.mps[0], s);
  for (int i = 0; i < L - 1; ++i) {
    R.C[i] = copy_of(*C, s);
    R.psi[i] = std::make_unique<Tensor>();
    twosite(*C, *R.B[i + 1], *R.psi[i], s);
    Tensor U, V;
    op_canonicalize_twosite(true, *R.psi[i], U, V, 0, s);
    R.A[i] = std::make_unique<Tensor>(std::move(U));
    C = std::make_unique<Tensor>(std::move(V));
  }
  R.C[L - 1] = std::move(C);
  R.lH.resize(L + 1);
  R.rH.resize(L + 1);
  R.lS.resize(L + 1);
  R.rS.resize(L + 1);
  R.lH[0] = boundary_env(s);
  R.lS[0] = boundary_overlap(s);
  R.rH[L] = boundary_env(s);
  R.rS[L] = boundary_overlap(s);
  D.lower.push_back(std::move(R));
  rebuild_cross_right(D, D.lower.back(), s);
}

// <psi| dH/dW_k |psi> for every site k (the adjoint of the sigma-vector contraction with respect to the MPO tensor),
// from the state a sweep leaves (centre at site 0, right-canonical sites, right environments valid):
//   sites 0..L-2, centre moved right site by site without truncation:
//     G_i(wl,n1b,n1k,w') = sum L_i(lb,wl,lk) psi*(lb,n1b,n2b,rb) psi(lk,n1k,n2k,rk) W_{i+1}(w',n2b,n2k,wr) R(rb,wr,rk)
//   last site, centre there:  G(wl,nb,nk,wr) = sum L(lb,wl,lk) c*(lb,nb,rb) c(lk,nk,rk) R(rb,wr,rk)
// then the centre is moved back to site 0 (environments recomputed, as renormalize does), so sweeps can continue.
// G_k has W_k's block indices with every quantum number negated (the dual space of W_k).
void dmrg_mpo_gradient(Dmrg& D, const std::vector<Tensor*>& G, hipStream_t s) {
  const int L = D.L;
  if ((int)G.size() != L) throw Error(MPSXX_ERR_ARG, "dmrg gradient: one output tensor per site expected");
  for (int i = 0; i < L - 1; ++i) {
    Tensor psi;
    twosite(*D.mps[i], *D.mps[i + 1], psi, s);
    Tensor X, Y, Z;
    op_contract(1.0, *D.mpo[i + 1], false, {3}, *D.renv[i + 2], false, {1}, 1.0, X, s);  // (w', n2b, n2k, rb, rk)
    op_contract(1.0, psi, false, {2, 3}, X, false, {2, 4}, 1.0, Y, s);                  // (lk, n1k, w', n2b, rb)
    op_contract(1.0, psi, true, {2, 3}, Y, false, {3, 4}, 1.0, Z, s);                   // (lb, n1b, lk, n1k, w')
    *G[i] = Tensor();
    op_contract(1.0, *D.lenv[i], false, {0, 2}, Z, false, {0, 2}, 1.0, *G[i], s);        // (wl, n1b, n1k, w')
    Tensor U, V;
    op_canonicalize_twosite(true, psi, U, V, 0, s);
    *D.mps[i] = std::move(U);
    *D.mps[i + 1] = std::move(V);
    D.lenv[i + 1] = std::make_unique<Tensor>();
    op_renormalize(true, *D.mpo[i], *D.lenv[i], *D.mps[i], *D.mps[i], *D.lenv[i + 1], MPSXX_FP64, s);
  }
  {
    Tensor& c = *D.mps[L - 1];
    Tensor X, Y;
    op_contract(1.0, *D.lenv[L - 1], false, {0}, c, true, {0}, 1.0, X, s);           // (wl, lk, nb, rb)
    op_contract(1.0, X, false, {1}, c, false, {0}, 1.0, Y, s);                       // (wl, nb, rb, nk, rk)
    *G[L - 1] = Tensor();
    op_contract(1.0, Y, false, {2, 4}, *D.renv[L], false, {0, 2}, 1.0, *G[L - 1], s);  // (wl, nb, nk, wr)
  }
  for (int i = L - 2; i >= 0; --i) {
    Tensor psi;
    twosite(*D.mps[i], *D.mps[i + 1], psi, s);
    Tensor U, V;
    op_canonicalize_twosite(false, psi, U, V, 0, s);
    *D.mps[i] = std::move(U);
    *D.mps[i + 1] = std::move(V);
    D.renv[i + 1] = std::make_unique<Tensor>();
    op_renormalize(false, *D.mpo[i + 1], *D.renv[i + 2], *D.mps[i + 1], *D.mps[i + 1], *D.renv[i + 1], MPSXX_FP64, s);
  }
  D.center = 0;
  // the gauge of the right-canonical tensors changed: rebuild the cross environments of any lower roots
  for (LowerRoot& R : D.lower) rebuild_cross_right(D, R, s);
  hip_check(hipStreamSynchronize(s), "gradient");
}

// <psi|O|psi> / <psi|psi> of the MPS psi = prod_i T_i by full left-to-right environment chains
// (src/expectation.hpp:55-114 runs the same renormalize chain).
double expectation(const std::vector<Tensor*>& mpo, const std::vector<Tensor*>& mps, hipStream_t s) {
  const int L = (int)mps.size();
  if (L < 1 || (int)mpo.size() != L) throw Error(MPSXX_ERR_ARG, "expectation: one MPO tensor per site expected");
  std::unique_ptr<Tensor> E = boundary_env(s);
  Tensor S;
  {
    Layout l;
    l.set_shape(2, true, QN(), {{QN()}, {QN()}}, {{1}, {1}});
    int32_t z[2] = {0, 0};
    l.add_block(z);
    l.finalize();
    S.set_layout(l, s);
    launch_fill(S.ptr(), 1, 1.0, s);
  }
  for (int i = 0; i < L; ++i) {
    auto E1 = std::make_unique<Tensor>();
    op_renormalize(true, *mpo[i], *E, *mps[i], *mps[i], *E1, MPSXX_FP64, s);
    E = std::move(E1);
    Tensor S1;
    op_renormalize_overlap(true, S, *mps[i], *mps[i], S1, s);
    S = std::move(S1);
  }
  if (E->L.nelem != 1 || S.L.nelem != 1)
    throw Error(MPSXX_ERR_SHAPE, "expectation: the chain does not close on one scalar (MPO total quantum number?)");
  double e = 0.0, n = 0.0;
  hip_check(hipMemcpyAsync(&e, E->ptr(), sizeof(double), hipMemcpyDeviceToHost, s), "download");
  hip_check(hipMemcpyAsync(&n, S.ptr(), sizeof(double), hipMemcpyDeviceToHost, s), "download");
  hip_check(hipStreamSynchronize(s), "expectation");
  return e / n;
}

void dmrg_free(Dmrg* D) { delete D; }

}  // namespace mpx

struct mpsxx_dmrg_s {
  mpx::Dmrg* d = nullptr;
};

namespace {
thread_local std::string g_dmrg_err;
template <typename F>
int dguard(F f) {
  try {
    f();
    return MPSXX_OK;
  } catch (const mpx::Error& e) {
    g_dmrg_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_dmrg_err = e.what();
    return MPSXX_ERR_ARG;
  }
}
}  // namespace

extern "C" {

const char* mpsxx_dmrg_last_error(void) { return g_dmrg_err.c_str(); }

int mpsxx_dmrg_create(int nsites, const mpsxx_tensor* mpo, const mpsxx_tensor* init_mps, int max_states,
                      mpsxx_dmrg* out, void* stream) {
  return dguard([&] {
    std::vector<mpx::Tensor*> m, i;
    for (int k = 0; k < nsites; ++k) {
      m.push_back(&mpo[k]->t);
      i.push_back(&init_mps[k]->t);
    }
    auto* h = new mpsxx_dmrg_s;
    h->d = mpx::dmrg_create(nsites, m, i, max_states, reinterpret_cast<hipStream_t>(stream));
    *out = h;
  });
}

int mpsxx_dmrg_save(mpsxx_dmrg d, const char* prefix, int iroot, void* stream) {
  return dguard([&] {
    if (!d || !d->d || !prefix) throw mpx::Error(MPSXX_ERR_ARG, "dmrg_save: null argument");
    mpx::dmrg_save(*d->d, prefix, iroot, reinterpret_cast<hipStream_t>(stream));
  });
}

int mpsxx_dmrg_load(int nsites, const mpsxx_tensor* mpo, const char* prefix, int iroot, int max_states,
                    mpsxx_dmrg* out, void* stream) {
  return dguard([&] {
    if (!mpo || !prefix || !out) throw mpx::Error(MPSXX_ERR_ARG, "dmrg_load: null argument");
    std::vector<mpx::Tensor*> m;
    for (int k = 0; k < nsites; ++k) {
      if (!mpo[k]) throw mpx::Error(MPSXX_ERR_ARG, "dmrg_load: null mpo tensor");
      m.push_back(&mpo[k]->t);
    }
    mpx::Dmrg* D = mpx::dmrg_load(nsites, m, prefix, iroot, max_states, reinterpret_cast<hipStream_t>(stream));
    auto* h = new mpsxx_dmrg_s;
    h->d = D;
    *out = h;
  });
}

int mpsxx_dmrg_sweep(mpsxx_dmrg d, int max_states, double tolerance, int precision, double* energy,
                     mpsxx_sweep_stats* stats, void* stream) {
  return dguard([&] {
    mpx::SweepStats st;
    const double e = mpx::dmrg_sweep(*d->d, max_states, tolerance, precision, st, reinterpret_cast<hipStream_t>(stream));
    if (energy) *energy = e;
    if (stats) {
      stats->seconds = st.seconds;
      stats->sigma_ms = st.sigma_ms;
      stats->davidson_s = st.davidson_s;
      stats->svd_s = st.svd_s;
      stats->renorm_s = st.renorm_s;
      stats->diag_s = st.diag_s;
      stats->nsigma = st.nsigma;
      stats->max_bond = st.max_bond;
      stats->discarded_weight = st.dweight;
      stats->nsigma_fp32 = st.nsigma_fp32;
      stats->pad = 0;
    }
  });
}

int mpsxx_dmrg_add_lower_root(mpsxx_dmrg d, mpsxx_dmrg root, double energy, void* stream) {
  return dguard([&] {
    if (!d || !d->d || !root || !root->d) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: null handle");
    if (d == root) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: a state cannot be its own lower root");
    mpx::dmrg_add_lower_root(*d->d, *root->d, energy, reinterpret_cast<hipStream_t>(stream));
  });
}

int mpsxx_dmrg_sweep_trace(mpsxx_dmrg d, int32_t max_steps, int32_t* nsteps, int32_t* step_site,
                           int32_t* step_forward, double* step_energy, int32_t* step_nevals, int32_t max_evals,
                           int32_t* nevals, double* evals, double* rnorms) {
  return dguard([&] {
    if (!d || !d->d) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: null handle");
    const auto& tr = d->d->trace;
    int64_t ne = 0;
    for (const auto& t : tr) ne += (int64_t)t.evals.size();
    if (nsteps) *nsteps = (int32_t)tr.size();
    if (nevals) *nevals = (int32_t)ne;
    if (!step_site && !step_forward && !step_energy && !step_nevals && !evals && !rnorms) return;  // counts only
    if ((int64_t)tr.size() > max_steps || ne > max_evals) throw mpx::Error(MPSXX_ERR_ARG, "dmrg trace: arrays too small");
    int64_t k = 0;
    for (size_t i = 0; i < tr.size(); ++i) {
      if (step_site) step_site[i] = tr[i].site;
      if (step_forward) step_forward[i] = tr[i].forward;
      if (step_energy) step_energy[i] = tr[i].energy;
      if (step_nevals) step_nevals[i] = (int32_t)tr[i].evals.size();
      for (size_t j = 0; j < tr[i].evals.size(); ++j, ++k) {
        if (evals) evals[k] = tr[i].evals[j];
        if (rnorms) rnorms[k] = j < tr[i].rnorms.size() ? tr[i].rnorms[j] : 0.0;
      }
    }
  });
}

int mpsxx_dmrg_set_truncation(mpsxx_dmrg d, double noise, int rule) {
  return dguard([&] {
    if (!d || !d->d) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: null handle");
    if (rule != MPSXX_TRUNCATION_REFERENCE && rule != MPSXX_TRUNCATION_STABLE)
      throw mpx::Error(MPSXX_ERR_ARG, "dmrg: unknown truncation rule");
    if (!(noise >= 0.0)) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: negative noise");
    d->d->noise = noise;
    d->d->trunc_rule = rule;
  });
}

int mpsxx_dmrg_set_form(mpsxx_dmrg d, int form) {
  return dguard([&] {
    if (!d || !d->d) throw mpx::Error(MPSXX_ERR_ARG, "dmrg: null handle");
    if (form != MPSXX_SWEEP_FOURINDEX && form != MPSXX_SWEEP_MERGED && form != MPSXX_SWEEP_ONESITE)
      throw mpx::Error(MPSXX_ERR_ARG, "dmrg: unknown form");
    if (form == MPSXX_SWEEP_MERGED && !d->d->lower.empty())
      throw mpx::Error(MPSXX_ERR_UNSUPPORTED, "dmrg: lower roots are not supported with the merged sweep form");
    d->d->form = form;
  });
}

int mpsxx_dmrg_mpo_gradient(mpsxx_dmrg d, const mpsxx_tensor* grads, void* stream) {
  return dguard([&] {
    if (!d || !grads) throw mpx::Error(MPSXX_ERR_ARG, "dmrg gradient: null argument");
    std::vector<mpx::Tensor*> g;
    for (int k = 0; k < d->d->L; ++k) g.push_back(&grads[k]->t);
    mpx::dmrg_mpo_gradient(*d->d, g, reinterpret_cast<hipStream_t>(stream));
  });
}

int mpsxx_expectation(int nsites, const mpsxx_tensor* mpo, const mpsxx_tensor* mps, double* value, void* stream) {
  return dguard([&] {
    if (!mpo || !mps || !value) throw mpx::Error(MPSXX_ERR_ARG, "expectation: null argument");
    std::vector<mpx::Tensor*> m, a;
    for (int k = 0; k < nsites; ++k) {
      m.push_back(&mpo[k]->t);
      a.push_back(&mps[k]->t);
    }
    *value = mpx::expectation(m, a, reinterpret_cast<hipStream_t>(stream));
  });
}

int mpsxx_dmrg_site(mpsxx_dmrg d, int site, mpsxx_tensor out, void* stream) {
  return dguard([&] {
    if (site < 0 || site >= d->d->L) throw mpx::Error(MPSXX_ERR_ARG, "site out of range");
    mpx::op_copy(*d->d->mps[site], out->t, reinterpret_cast<hipStream_t>(stream));
  });
}

int mpsxx_dmrg_free(mpsxx_dmrg d) {
  return dguard([&] {
    if (d) {
      mpx::dmrg_free(d->d);
      delete d;
    }
  });
}

}  // extern "C"
