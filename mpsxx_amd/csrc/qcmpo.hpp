// Quantum-chemistry MPO generator (host side; see qcmpo.cpp).
#pragma once

#include <array>
#include <vector>

#include "engine.hpp"

namespace mpx {

// One site tensor W(wl, n_bra, n_ket, wr): bond-channel quantum numbers (of the accumulated left operator) and the
// non-zero 4x4 site matrices of every (wl, wr) transition.
struct QcMpoSite {
  std::vector<QN> lq, rq;
  std::vector<int32_t> wl, wr;
  std::vector<std::array<double, 16>> mat;
};
// Where the coefficient of one Hamiltonian term enters the MPO: W_site[wl][wr] += coeff * mult * mats[mat]
// (each term's coefficient is applied at exactly one transition).
struct QcTermApp {
  int32_t site, wl, wr, mat;
  double mult;
  int32_t p, q;   // one-body term a+_ps a_qs: orbitals p, q; two-body: p = q = -1
  double coeff;   // the term's coefficient
};
struct QcMpo {
  int K = 0;
  std::vector<QcMpoSite> sites;
  std::vector<QcTermApp> apps;
  std::vector<std::array<double, 16>> mats;  // local factor matrices referenced by apps
  double ecore = 0.0;
};

// dE/dh and <V> from the MPO gradients G_k[wl, n_bra, n_ket, wr] = <psi| dH/dW_k[wl, n_bra, n_ket, wr] |psi>:
// per term d_t = <psi|term_t|psi> = mult * <G_site[wl, ., ., wr], mats[mat]>; dh[p*K+q] = sum over the one-body terms
// of (p, q) (both spins), dlambda = sum over two-body terms of coeff_t d_t (v -> lambda v at lambda = 1), energy =
// sum over all terms of coeff_t d_t + ecore.
void qc_mpo_gradient(const QcMpo& m, const std::vector<const Tensor*>& G, double* dh, double* dlambda, double* energy,
                     hipStream_t s);

// h[p*K+q] = h_pq; v[((p*K+q)*K+r)*K+s] = (pq|rs) (chemist notation, as the reference's twoint(p,r,q,s))
QcMpo build_qc_mpo(int K, const double* h, const double* v, double ecore, bool swap_sweep);
// Device tensor of one site (every non-zero element its own 1x1x1x1 block, as src/prime_operators.h:177-184).
void qc_mpo_tensor(const QcMpoSite& site, Tensor& out, hipStream_t s);

}  // namespace mpx

struct mpsxx_qcmpo_s {
  static constexpr uint64_t kMagic = 0x4d5053585143504full;
  uint64_t magic = kMagic;
  mpx::QcMpo m;
  ~mpsxx_qcmpo_s() { magic = 0; }
};
