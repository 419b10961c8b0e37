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
struct QcMpo {
  int K = 0;
  std::vector<QcMpoSite> sites;
};

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
