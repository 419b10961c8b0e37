// Drop-in C++ boundary: the reference's hot-path function templates (namespace mpsxx, same names, argument order and
// accumulate semantics) implemented over the C ABI in mpsxx_amd.h.
//
// Works with any block-sparse container offering the btas::QSTArray API subset the reference drivers use
// (include/btas/QSPARSE/QSTArray.h, include/btas/SPARSE/STArray.h): q(), qshape(i), dshape(i), begin()/end() over
// (tag, shared_ptr<dense block>) pairs in ascending tag order, index(tag), size(), resize(q, qshape, dshape, false) and
// reserve(tag); dense blocks provide data() and size().  Quantum numbers provide p(), s() and a (p, s) constructor
// (src/fermion.h).  Including this header in place of src/compute_sigma_vector.hpp, src/compute_diagonal_elements.hpp
// and src/renormalize.hpp routes Davidson (src/diagonalize.hpp) and the site optimizers (src/optimize.hpp) through the
// device path with unchanged call sites.
//
// Each call uploads its inputs, runs on the device and downloads the output (the output is uploaded first when it is
// non-empty, to keep the reference's accumulate semantics).  Callers that keep operands resident use the C ABI
// directly (mpsxx_tensor handles).  Errors from the C ABI are rethrown as std::runtime_error, as BTAS_THROW does.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mpsxx_amd.h"

namespace mpsxx_amd {

inline void check(int rc) {
  if (rc != MPSXX_OK) throw std::runtime_error(std::string("mpsxx_amd: ") + mpsxx_last_error());
}

// RAII handle
class Handle {
 public:
  Handle() { check(mpsxx_tensor_create_empty(&h_)); }
  explicit Handle(mpsxx_tensor h) : h_(h) {}
  ~Handle() {
    if (h_) mpsxx_tensor_free(h_);
  }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  Handle(Handle&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  mpsxx_tensor get() const { return h_; }

 private:
  mpsxx_tensor h_ = nullptr;
};

// host container -> device tensor.  A default-constructed container (no sectors on any index) is the empty tensor;
// one that carries q/qshape but no blocks (the reference's diag, built as QSTArray(q, qshape) from the wavefunction,
// optimize.hpp:67) keeps that structure.  with_blocks = false uploads the structure only.  Indices whose dense shape
// is not set (QSTArray(q, qshape) leaves it empty) get zero-sized sectors; the ops take dims from their inputs.
template <class Arr>
Handle upload(const Arr& a, bool with_blocks = true) {
  const int N = (int)a.qshape().size();
  bool shaped = false;
  for (int i = 0; i < N; ++i) shaped = shaped || !a.qshape(i).empty();
  if (!shaped) return Handle();
  std::vector<int32_t> nsec(N), secq, secd, bidx;
  for (int i = 0; i < N; ++i) {
    const auto& qs = a.qshape(i);
    const auto& ds = a.dshape(i);
    const bool has_dims = ds.size() == qs.size();
    nsec[i] = (int32_t)qs.size();
    for (size_t k = 0; k < qs.size(); ++k) {
      secq.push_back(qs[k].p());
      secq.push_back(qs[k].s());
      secd.push_back(has_dims ? (int32_t)ds[k] : 0);
    }
  }
  std::vector<double> data;
  if (with_blocks) {
    for (auto it = a.begin(); it != a.end(); ++it) {
      const auto idx = a.index(it->first);
      for (int i = 0; i < N; ++i) bidx.push_back((int32_t)idx[i]);
      const double* p = it->second->data();
      data.insert(data.end(), p, p + it->second->size());
    }
  }
  mpsxx_qst_desc d;
  d.rank = N;
  d.qsym = 1;
  d.q_total[0] = a.q().p();
  d.q_total[1] = a.q().s();
  d.nsec = nsec.data();
  d.sec_q = secq.data();
  d.sec_dim = secd.data();
  d.nblocks = (int64_t)(bidx.size() / (size_t)N);
  d.block_idx = bidx.data();
  mpsxx_tensor h;
  check(mpsxx_tensor_create(&d, &h));
  Handle out(h);
  check(mpsxx_tensor_upload(h, data.data(), (int64_t)data.size(), nullptr));
  return out;
}

// Target of Axpy(1.0, x, diag) (STBLAS.h:94-145): a container without blocks is shaped from the first x (its dense
// shapes may not be set yet), one with blocks is accumulated into.
template <class Arr>
Handle accumulator(const Arr& a) {
  if (a.begin() == a.end()) return Handle();
  return upload(a);
}

// device tensor -> host container (replaces its content)
template <class Arr>
void download(const Handle& t, Arr& a) {
  int32_t rank, qsym, q[2];
  int64_t nsec_total, nb, ne;
  check(mpsxx_tensor_info(t.get(), &rank, &qsym, q, &nsec_total, &nb, &ne));
  if (rank == 0) {
    a.clear();
    return;
  }
  std::vector<int32_t> nsec(rank), secq(2 * nsec_total), secd(nsec_total), bidx(nb * rank);
  std::vector<int64_t> boff(nb);
  check(mpsxx_tensor_layout(t.get(), nsec.data(), secq.data(), secd.data(), bidx.data(), boff.data()));
  std::vector<double> data(ne);
  check(mpsxx_tensor_download(t.get(), data.data(), ne, nullptr));
  typedef typename std::decay<decltype(a.q())>::type Q;
  auto qshape = a.qshape();  // TVector<Qshapes<Q>, N>
  auto dshape = a.dshape();  // TVector<Dshapes, N>
  int64_t k = 0;
  for (int i = 0; i < rank; ++i) {
    qshape[i].clear();
    dshape[i].clear();
    for (int j = 0; j < nsec[i]; ++j, ++k) {
      qshape[i].push_back(Q(secq[2 * k], secq[2 * k + 1]));
      dshape[i].push_back(secd[k]);
    }
  }
  a.resize(Q(q[0], q[1]), qshape, dshape, false);
  for (int64_t b = 0; b < nb; ++b) {
    int64_t tag = 0;
    for (int i = 0; i < rank; ++i) tag = tag * nsec[i] + bidx[b * rank + i];
    auto it = a.reserve((int)tag);
    if (it == a.end()) throw std::runtime_error("mpsxx_amd: downloaded block not allowed by the container");
    std::memcpy(it->second->data(), data.data() + boff[b], it->second->size() * sizeof(double));
  }
}

}  // namespace mpsxx_amd

namespace mpsxx {

//! Compute sigmavector for two-site algorithm: | phi > = H | psi >   (reference src/compute_sigma_vector.hpp:54-70)
template <template <typename, size_t, class> class Arr, class Q>
void compute_sigma_vector(const Arr<double, 4, Q>& lmpo, const Arr<double, 4, Q>& rmpo, const Arr<double, 3, Q>& lopr,
                          const Arr<double, 3, Q>& ropr, const Arr<double, 4, Q>& wfn0, Arr<double, 4, Q>& sgv0) {
  using namespace mpsxx_amd;
  Handle a = upload(lmpo), b = upload(rmpo), c = upload(lopr), d = upload(ropr), e = upload(wfn0), f = upload(sgv0);
  check(mpsxx_compute_sigma_vector_twosite(a.get(), b.get(), c.get(), d.get(), e.get(), f.get(), MPSXX_FP64, nullptr));
  download(f, sgv0);
}

//! Compute sigmavector: | phi > = H | psi >   (one-site, src/compute_sigma_vector.hpp:10-22)
template <template <typename, size_t, class> class Arr, class Q>
void compute_sigma_vector(const Arr<double, 4, Q>& mpo0, const Arr<double, 3, Q>& lopr, const Arr<double, 3, Q>& ropr,
                          const Arr<double, 3, Q>& wfn0, Arr<double, 3, Q>& sgv0) {
  using namespace mpsxx_amd;
  Handle a = upload(mpo0), c = upload(lopr), d = upload(ropr), e = upload(wfn0), f = upload(sgv0);
  check(mpsxx_compute_sigma_vector_onesite(a.get(), c.get(), d.get(), e.get(), f.get(), MPSXX_FP64, nullptr));
  download(f, sgv0);
}

//! Compute sigmavector for merged algorithm (src/compute_sigma_vector.hpp:86-96)
template <template <typename, size_t, class> class Arr, class Q>
void compute_sigma_vector(const Arr<double, 3, Q>& lopr, const Arr<double, 3, Q>& ropr, const Arr<double, 2, Q>& wfn0,
                          Arr<double, 2, Q>& sgv0) {
  using namespace mpsxx_amd;
  Handle c = upload(lopr), d = upload(ropr), e = upload(wfn0), f = upload(sgv0);
  check(mpsxx_compute_sigma_vector_merged(c.get(), d.get(), e.get(), f.get(), MPSXX_FP64, nullptr));
  download(f, sgv0);
}

//! Compute sigmavector: | phi > = S | psi >  (src/compute_sigma_vector.hpp:25-35, 73-83, 99-109)
template <template <typename, size_t, class> class Arr, class Q, size_t N>
void compute_sigma_vector(const Arr<double, 2, Q>& lopr, const Arr<double, 2, Q>& ropr, const Arr<double, N, Q>& wfn0,
                          Arr<double, N, Q>& sgv0) {
  using namespace mpsxx_amd;
  Handle c = upload(lopr), d = upload(ropr), e = upload(wfn0), f = upload(sgv0);
  check(mpsxx_compute_sigma_vector_overlap(c.get(), d.get(), e.get(), f.get(), nullptr));
  download(f, sgv0);
}

//! Compute sigmavector: < phi | = < psi | H^*   (src/compute_sigma_vector.hpp:38-51).  The reference contracts into
//! the temporary sgv0.conjugate(), so its result is discarded; here it lands in sgv0 (quantum numbers of the conjugate
//! flipped back), which is what a caller of this form wants.
template <template <typename, size_t, class> class Arr, class Q>
void compute_sigma_vector_conj(const Arr<double, 4, Q>& mpo0, const Arr<double, 3, Q>& lopr,
                               const Arr<double, 3, Q>& ropr, const Arr<double, 3, Q>& wfn0, Arr<double, 3, Q>& sgv0) {
  using namespace mpsxx_amd;
  Handle a = upload(mpo0), c = upload(lopr), d = upload(ropr), e = upload(wfn0), f = upload(sgv0);
  check(mpsxx_compute_sigma_vector_conj(a.get(), c.get(), d.get(), e.get(), f.get(), nullptr));
  download(f, sgv0);
}

//! Compute diagonal H elements for one-site algorithm (src/compute_diagonal_elements.hpp:11-32); diag carries the
//! wavefunction's q/qshape (blocks, if any, are replaced, as Copy(scr, diag, 1) does)
template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const Arr<double, 4, Q>& mpo0, const Arr<double, 3, Q>& lopr,
                               const Arr<double, 3, Q>& ropr, Arr<double, 3, Q>& diag) {
  using namespace mpsxx_amd;
  Handle a = upload(mpo0), c = upload(lopr), d = upload(ropr), f = upload(diag, false);
  check(mpsxx_compute_diagonal_elements_onesite(a.get(), c.get(), d.get(), f.get(), nullptr));
  download(f, diag);
}

//! Compute diagonal H elements for two-site algorithm (src/compute_diagonal_elements.hpp:35-61)
template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const Arr<double, 4, Q>& lmpo, const Arr<double, 4, Q>& rmpo,
                               const Arr<double, 3, Q>& lopr, const Arr<double, 3, Q>& ropr, Arr<double, 4, Q>& diag) {
  using namespace mpsxx_amd;
  Handle a = upload(lmpo), b = upload(rmpo), c = upload(lopr), d = upload(ropr), f = upload(diag, false);
  check(mpsxx_compute_diagonal_elements_twosite(a.get(), b.get(), c.get(), d.get(), f.get(), nullptr));
  download(f, diag);
}

//! Compute diagonal H elements for merged algorithm (src/compute_diagonal_elements.hpp:64-80)
template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const Arr<double, 3, Q>& lopr, const Arr<double, 3, Q>& ropr, Arr<double, 2, Q>& diag) {
  using namespace mpsxx_amd;
  Handle c = upload(lopr), d = upload(ropr), f = upload(diag, false);
  check(mpsxx_compute_diagonal_elements_merged(c.get(), d.get(), f.get(), nullptr));
  download(f, diag);
}

//! Group sums (src/compute_diagonal_elements.hpp:83-179): diag += sum over operator groups g of the group's diagonal.
//! One process holds every group here, so the reference's MPI reduce + broadcast has nothing left to combine.
template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const std::vector<Arr<double, 4, Q>>& mpo0, const std::vector<Arr<double, 3, Q>>& lopr,
                               const std::vector<Arr<double, 3, Q>>& ropr, Arr<double, 3, Q>& diag) {
  using namespace mpsxx_amd;
  if (mpo0.size() != lopr.size() || mpo0.size() != ropr.size())
    throw std::runtime_error("compute_diagonal_elements: operator group counts differ");
  if (mpo0.empty()) return;
  Handle out = accumulator(diag);
  for (size_t g = 0; g < mpo0.size(); ++g) {
    Handle a = upload(mpo0[g]), c = upload(lopr[g]), d = upload(ropr[g]), t = upload(diag, false);
    check(mpsxx_compute_diagonal_elements_onesite(a.get(), c.get(), d.get(), t.get(), nullptr));
    check(mpsxx_axpy(1.0, t.get(), out.get(), nullptr));
  }
  download(out, diag);
}

template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const std::vector<Arr<double, 4, Q>>& lmpo, const std::vector<Arr<double, 4, Q>>& rmpo,
                               const std::vector<Arr<double, 3, Q>>& lopr, const std::vector<Arr<double, 3, Q>>& ropr,
                               Arr<double, 4, Q>& diag) {
  using namespace mpsxx_amd;
  if (lmpo.size() != rmpo.size() || lmpo.size() != lopr.size() || rmpo.size() != ropr.size())
    throw std::runtime_error("compute_diagonal_elements: operator group counts differ");
  if (lmpo.empty()) return;
  Handle out = accumulator(diag);
  for (size_t g = 0; g < lmpo.size(); ++g) {
    Handle a = upload(lmpo[g]), b = upload(rmpo[g]), c = upload(lopr[g]), d = upload(ropr[g]), t = upload(diag, false);
    check(mpsxx_compute_diagonal_elements_twosite(a.get(), b.get(), c.get(), d.get(), t.get(), nullptr));
    check(mpsxx_axpy(1.0, t.get(), out.get(), nullptr));
  }
  download(out, diag);
}

template <template <typename, size_t, class> class Arr, class Q>
void compute_diagonal_elements(const std::vector<Arr<double, 3, Q>>& lopr, const std::vector<Arr<double, 3, Q>>& ropr,
                               Arr<double, 2, Q>& diag) {
  using namespace mpsxx_amd;
  if (lopr.size() != ropr.size()) throw std::runtime_error("compute_diagonal_elements: operator group counts differ");
  if (lopr.empty()) return;
  Handle out = accumulator(diag);
  for (size_t g = 0; g < lopr.size(); ++g) {
    Handle c = upload(lopr[g]), d = upload(ropr[g]), t = upload(diag, false);
    check(mpsxx_compute_diagonal_elements_merged(c.get(), d.get(), t.get(), nullptr));
    check(mpsxx_axpy(1.0, t.get(), out.get(), nullptr));
  }
  download(out, diag);
}

//! Renormalize operator tensor (src/renormalize.hpp:10-32)
template <template <typename, size_t, class> class Arr, class Q>
void renormalize(bool forward, const Arr<double, 4, Q>& mpo0, const Arr<double, 3, Q>& opr0,
                 const Arr<double, 3, Q>& bra0, const Arr<double, 3, Q>& ket0, Arr<double, 3, Q>& opr1) {
  using namespace mpsxx_amd;
  Handle a = upload(mpo0), b = upload(opr0), c = upload(bra0), d = upload(ket0), f = upload(opr1);
  check(mpsxx_renormalize(forward ? 1 : 0, a.get(), b.get(), c.get(), d.get(), f.get(), MPSXX_FP64, nullptr));
  download(f, opr1);
}

//! Renormalize overlap matrix (src/renormalize.hpp:36-53) and merged overlap matrix (:78-95)
template <template <typename, size_t, class> class Arr, class Q, size_t N>
void renormalize(bool forward, const Arr<double, 2, Q>& sov0, const Arr<double, N, Q>& bra0,
                 const Arr<double, N, Q>& ket0, Arr<double, 2, Q>& sov1) {
  static_assert(N == 2 || N == 3, "renormalize(overlap): bra/ket have 3 indices (MPS site) or 2 (merged)");
  using namespace mpsxx_amd;
  Handle a = upload(sov0), c = upload(bra0), d = upload(ket0), f = upload(sov1);
  check(mpsxx_renormalize_overlap(forward ? 1 : 0, a.get(), c.get(), d.get(), f.get(), nullptr));
  download(f, sov1);
}

//! Renormalize merged operator matrix (src/renormalize.hpp:57-74)
template <template <typename, size_t, class> class Arr, class Q>
void renormalize(bool forward, const Arr<double, 3, Q>& opr0, const Arr<double, 2, Q>& bra0,
                 const Arr<double, 2, Q>& ket0, Arr<double, 3, Q>& opr1) {
  using namespace mpsxx_amd;
  Handle a = upload(opr0), c = upload(bra0), d = upload(ket0), f = upload(opr1);
  check(mpsxx_renormalize_merged(forward ? 1 : 0, a.get(), c.get(), d.get(), f.get(), MPSXX_FP64, nullptr));
  download(f, opr1);
}

//! Compute merged block operators (src/optimize.hpp:273-368): for every root k and operator group g,
//! mHopr[k][g] = merge(xHopr[k][g] . xmpo[g]) (forward: rows (l, n), columns (l', n'); backward: rows (n, r), columns
//! (n', r')), and mSopr[k] = merge(xSopr[k] (x) I_site).  The outputs are resized as the reference resizes them.
template <template <typename, size_t, class> class Arr, class Q>
void compute_merged_operator(bool forward, const std::vector<Arr<double, 4, Q>>& xmpo,
                             const std::vector<std::vector<Arr<double, 3, Q>>>& xHopr,
                             std::vector<std::vector<Arr<double, 3, Q>>>& mHopr,
                             const std::vector<Arr<double, 2, Q>>& xSopr, std::vector<Arr<double, 2, Q>>& mSopr) {
  using namespace mpsxx_amd;
  if (xmpo.empty() || xHopr.empty()) throw std::runtime_error("compute_merged_operator: no groups or roots");
  const size_t nroot = xHopr.size();
  if (xSopr.size() != nroot) throw std::runtime_error("compute_merged_operator: one overlap per root");
  mHopr.resize(nroot);
  mSopr.resize(nroot);
  for (size_t k = 0; k < nroot; ++k) {
    if (xHopr[k].size() != xmpo.size()) throw std::runtime_error("compute_merged_operator: one operator per group");
    mHopr[k].resize(xmpo.size());
    for (size_t g = 0; g < xmpo.size(); ++g) {
      Handle w = upload(xmpo[g]), o = upload(xHopr[k][g]), out;
      check(mpsxx_compute_merged_operator(forward ? 1 : 0, w.get(), o.get(), out.get(), nullptr));
      download(out, mHopr[k][g]);
    }
    Handle w = upload(xmpo[0]), so = upload(xSopr[k]), out;
    check(mpsxx_compute_merged_overlap(forward ? 1 : 0, w.get(), so.get(), out.get(), nullptr));
    download(out, mSopr[k]);
  }
}

//! Canonicalize two-site MPS (src/canonicalize.hpp:27-46): make_noise on wfnc (in place), blocked SVD truncated by
//! the reference's rule (QSTLAPACK.h:209-222), lmps = U and rmps = s Vt (forward) or lmps = U s and rmps = Vt.
template <template <typename, size_t, class> class Arr, class Q>
void canonicalize(bool forward, Arr<double, 4, Q>& wfnc, Arr<double, 3, Q>& lmps, Arr<double, 3, Q>& rmps,
                  const int& M = 0, const double& noise = 0.0) {
  using namespace mpsxx_amd;
  Handle w = upload(wfnc), l, r;
  check(mpsxx_canonicalize_twosite_ex(forward ? 1 : 0, w.get(), l.get(), r.get(), M, noise,
                                      MPSXX_TRUNCATION_REFERENCE, nullptr, nullptr));
  download(l, lmps);
  download(r, rmps);
  if (noise >= 1.0e-8) download(w, wfnc);
}

//! Canonicalize one-site MPS (src/canonicalize.hpp:49-85): forward lmps = U and rmps <- s Vt rmps; backward
//! rmps = Vt and lmps <- lmps U s; with noise the two-site wavefunction with the neighbour goes through the two-site
//! form (:57-66).
template <template <typename, size_t, class> class Arr, class Q>
void canonicalize(bool forward, Arr<double, 3, Q>& wfnc, Arr<double, 3, Q>& lmps, Arr<double, 3, Q>& rmps,
                  const int& M = 0, const double& noise = 0.0) {
  using namespace mpsxx_amd;
  Handle w = upload(wfnc), l = upload(lmps), r = upload(rmps);
  check(mpsxx_canonicalize_onesite_ex(forward ? 1 : 0, w.get(), l.get(), r.get(), M, noise,
                                      MPSXX_TRUNCATION_REFERENCE, nullptr, nullptr));
  download(l, lmps);
  download(r, rmps);
}

//! Canonicalize merged MPS (src/canonicalize.hpp:88-106): make_noise on wfnc (in place), lmat = U and rmat = s Vt
//! (forward) or lmat = U s and rmat = Vt.
template <template <typename, size_t, class> class Arr, class Q>
void canonicalize(bool forward, Arr<double, 2, Q>& wfnc, Arr<double, 2, Q>& lmat, Arr<double, 2, Q>& rmat,
                  const int& M = 0, const double& noise = 0.0) {
  using namespace mpsxx_amd;
  Handle w = upload(wfnc), l, r;
  check(mpsxx_canonicalize_merged_ex(forward ? 1 : 0, w.get(), l.get(), r.get(), M, noise,
                                     MPSXX_TRUNCATION_REFERENCE, nullptr, nullptr));
  download(l, lmat);
  download(r, rmat);
  if (noise >= 1.0e-8) download(w, wfnc);
}

}  // namespace mpsxx
