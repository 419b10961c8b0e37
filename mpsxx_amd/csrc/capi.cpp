// extern "C" boundary: opaque tensor handles, status codes, thread-local error text.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "batch.hpp"
#include "fileio.hpp"
#include "mpsxx_amd.h"
#include "noalloc.hpp"
#include "ops.hpp"
#include "qcmpo.hpp"

namespace {

thread_local std::string g_err;
bool g_init = false;

void ensure_init() {
  if (!g_init) {
    mpx::hip_check(hipSetDevice(0), "hipSetDevice");
    mpx::noalloc_init(0);
    g_init = true;
  }
}

template <typename F>
int guard(F f) {
  try {
    ensure_init();
    f();
    return MPSXX_OK;
  } catch (const mpx::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return MPSXX_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MPSXX_ERR_ARG;
  }
}

// host-only entry points: no device initialisation
template <typename F>
int guard_host(F f) {
  try {
    f();
    return MPSXX_OK;
  } catch (const mpx::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return MPSXX_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MPSXX_ERR_ARG;
  }
}

mpx::QcMpo& QM(mpsxx_qcmpo h) {
  if (!h || h->magic != mpsxx_qcmpo_s::kMagic) throw mpx::Error(MPSXX_ERR_ARG, "invalid or freed qc mpo handle");
  return h->m;
}

mpx::Tensor& T(mpsxx_tensor h) {
  if (!h) throw mpx::Error(MPSXX_ERR_ARG, "null tensor handle");
  if (h->magic != mpsxx_tensor_s::kMagic) throw mpx::Error(MPSXX_ERR_ARG, "invalid or freed tensor handle");
  return h->t;
}

mpx::BatchSigma& B(mpsxx_batch h) {
  if (!h || h->magic != mpsxx_batch_s::kMagic || !h->b) throw mpx::Error(MPSXX_ERR_ARG, "invalid or freed batch handle");
  return *h->b;
}

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

mpx::Layout layout_from_desc(const mpsxx_qst_desc* d) {
  if (!d || d->rank < 1 || d->rank > mpx::kMaxRank || !d->nsec || !d->sec_dim)
    throw mpx::Error(MPSXX_ERR_ARG, "invalid descriptor");
  std::vector<std::vector<mpx::QN>> qs(d->rank);
  std::vector<std::vector<int>> ds(d->rank);
  int64_t k = 0;
  for (int i = 0; i < d->rank; ++i) {
    if (d->nsec[i] < 0) throw mpx::Error(MPSXX_ERR_ARG, "negative sector count");
    for (int j = 0; j < d->nsec[i]; ++j, ++k) {
      ds[i].push_back(d->sec_dim[k]);
      if (d->qsym) {
        if (!d->sec_q) throw mpx::Error(MPSXX_ERR_ARG, "sec_q required for qsym arrays");
        qs[i].push_back(mpx::QN(d->sec_q[2 * k], d->sec_q[2 * k + 1]));
      }
    }
  }
  mpx::Layout L;
  L.set_shape(d->rank, d->qsym != 0, mpx::QN(d->q_total[0], d->q_total[1]), qs, ds);
  for (int64_t b = 0; b < d->nblocks; ++b) {
    const int32_t* idx = d->block_idx + b * d->rank;
    for (int i = 0; i < d->rank; ++i)
      if (idx[i] < 0 || idx[i] >= d->nsec[i]) throw mpx::Error(MPSXX_ERR_ARG, "block index out of range");
    if (!L.allowed(idx)) throw mpx::Error(MPSXX_ERR_QN, "block is not allowed by the quantum numbers");
    L.add_block(idx);
  }
  L.finalize();
  return L;
}

}  // namespace

extern "C" {

int mpsxx_init(int device) {
  try {
    mpx::hip_check(hipSetDevice(device), "hipSetDevice");
    mpx::noalloc_init(device);  // the scratch array's address on this device
    g_init = true;
    return MPSXX_OK;
  } catch (const mpx::Error& e) {
    g_err = e.what();
    return e.code;
  }
}

const char* mpsxx_last_error(void) { return g_err.c_str(); }

int mpsxx_synchronize(void* stream) {
  return guard([&] { mpx::hip_check(hipStreamSynchronize(S(stream)), "hipStreamSynchronize"); });
}

int mpsxx_tensor_create(const mpsxx_qst_desc* desc, mpsxx_tensor* out) {
  return guard([&] {
    if (!out) throw mpx::Error(MPSXX_ERR_ARG, "null output handle");
    mpx::Layout L = layout_from_desc(desc);
    auto* h = new mpsxx_tensor_s;
    try {
      h->t.set_layout(L, nullptr);
      mpx::hip_check(hipStreamSynchronize(nullptr), "create");
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mpsxx_tensor_create_empty(mpsxx_tensor* out) {
  return guard([&] {
    if (!out) throw mpx::Error(MPSXX_ERR_ARG, "null output handle");
    *out = new mpsxx_tensor_s;
  });
}

int mpsxx_tensor_free(mpsxx_tensor t) {
  return guard([&] { delete t; });
}

int mpsxx_tensor_clear(mpsxx_tensor t) {
  return guard([&] {
    mpx::Tensor& x = T(t);
    x.data.release();
    x.L = mpx::Layout();
    x.empty = true;
  });
}

int mpsxx_tensor_upload(mpsxx_tensor t, const double* host, int64_t nelem, void* stream) {
  return guard([&] {
    mpx::Tensor& x = T(t);
    if (x.empty || nelem != x.L.nelem) throw mpx::Error(MPSXX_ERR_SHAPE, "upload: element count mismatch");
    if (nelem > 0)
      mpx::hip_check(hipMemcpyAsync(x.ptr(), host, nelem * sizeof(double), hipMemcpyHostToDevice, S(stream)), "upload");
    mpx::hip_check(hipStreamSynchronize(S(stream)), "upload sync");
  });
}

int mpsxx_tensor_download(mpsxx_tensor t, double* host, int64_t nelem, void* stream) {
  return guard([&] {
    mpx::Tensor& x = T(t);
    const int64_t n = x.empty ? 0 : x.L.nelem;
    if (nelem != n) throw mpx::Error(MPSXX_ERR_SHAPE, "download: element count mismatch");
    if (n > 0)
      mpx::hip_check(hipMemcpyAsync(host, x.ptr(), n * sizeof(double), hipMemcpyDeviceToHost, S(stream)), "download");
    mpx::hip_check(hipStreamSynchronize(S(stream)), "download sync");
  });
}

int mpsxx_tensor_info(mpsxx_tensor t, int32_t* rank, int32_t* qsym, int32_t q_total[2], int64_t* nsec_total,
                      int64_t* nblocks, int64_t* nelem) {
  return guard([&] {
    const mpx::Tensor& x = T(t);
    if (rank) *rank = x.empty ? 0 : x.L.rank;
    if (qsym) *qsym = x.L.qsym ? 1 : 0;
    if (q_total) { q_total[0] = x.L.q.p; q_total[1] = x.L.q.s; }
    int64_t ns = 0;
    for (int i = 0; i < x.L.rank; ++i) ns += x.L.nsec(i);
    if (nsec_total) *nsec_total = x.empty ? 0 : ns;
    if (nblocks) *nblocks = x.empty ? 0 : x.L.nblocks();
    if (nelem) *nelem = x.empty ? 0 : x.L.nelem;
  });
}

int mpsxx_tensor_layout(mpsxx_tensor t, int32_t* nsec, int32_t* sec_q, int32_t* sec_dim, int32_t* block_idx,
                        int64_t* block_off) {
  return guard([&] {
    const mpx::Tensor& x = T(t);
    if (x.empty) return;
    int64_t k = 0;
    for (int i = 0; i < x.L.rank; ++i) {
      if (nsec) nsec[i] = x.L.nsec(i);
      for (int j = 0; j < x.L.nsec(i); ++j, ++k) {
        if (sec_dim) sec_dim[k] = x.L.ds[i][j];
        if (sec_q) { sec_q[2 * k] = x.L.qs[i][j].p; sec_q[2 * k + 1] = x.L.qs[i][j].s; }
      }
    }
    if (block_idx && !x.L.bidx.empty()) std::memcpy(block_idx, x.L.bidx.data(), x.L.bidx.size() * sizeof(int32_t));
    if (block_off && !x.L.off.empty()) std::memcpy(block_off, x.L.off.data(), x.L.off.size() * sizeof(int64_t));
  });
}

int mpsxx_tensor_device_ptr(mpsxx_tensor t, double** dptr) {
  return guard([&] { *dptr = T(t).empty ? nullptr : T(t).ptr(); });
}

int mpsxx_tensor_copy(mpsxx_tensor x, mpsxx_tensor y, void* stream) {
  return guard([&] { mpx::op_copy(T(x), T(y), S(stream)); });
}

int mpsxx_qst_save(const char* path, const mpsxx_qst_desc* desc, const double* data, int64_t nelem) {
  return guard_host([&] {
    if (!path) throw mpx::Error(MPSXX_ERR_ARG, "qst_save: null path");
    if (desc && desc->rank == 0) {  // an array with no structure
      if (nelem != 0) throw mpx::Error(MPSXX_ERR_SHAPE, "qst_save: element count mismatch");
      mpx::write_qst(path, nullptr, nullptr);
      return;
    }
    const mpx::Layout L = layout_from_desc(desc);
    if (nelem != L.nelem) throw mpx::Error(MPSXX_ERR_SHAPE, "qst_save: element count mismatch");
    if (nelem > 0 && !data) throw mpx::Error(MPSXX_ERR_ARG, "qst_save: null data");
    mpx::write_qst(path, &L, data);
  });
}

int mpsxx_qst_file_info(const char* path, int32_t* rank, int32_t* qsym, int32_t q_total[2], int64_t* nsec_total,
                        int64_t* nblocks, int64_t* nelem) {
  return guard_host([&] {
    if (!path) throw mpx::Error(MPSXX_ERR_ARG, "qst_file_info: null path");
    const mpx::QstFile f = mpx::read_qst(path, false);
    const mpx::Layout& L = f.L;
    int64_t ns = 0;
    for (int i = 0; i < L.rank; ++i) ns += L.nsec(i);
    if (rank) *rank = f.empty ? 0 : L.rank;
    if (qsym) *qsym = f.empty ? 0 : (L.qsym ? 1 : 0);
    if (q_total) { q_total[0] = L.q.p; q_total[1] = L.q.s; }
    if (nsec_total) *nsec_total = ns;
    if (nblocks) *nblocks = L.nblocks();
    if (nelem) *nelem = L.nelem;
  });
}

int mpsxx_qst_file_read(const char* path, int32_t* nsec, int32_t* sec_q, int32_t* sec_dim, int32_t* block_idx,
                        double* data, int64_t nelem) {
  return guard_host([&] {
    if (!path) throw mpx::Error(MPSXX_ERR_ARG, "qst_file_read: null path");
    const mpx::QstFile f = mpx::read_qst(path, true);
    const mpx::Layout& L = f.L;
    if (nelem != L.nelem) throw mpx::Error(MPSXX_ERR_SHAPE, "qst_file_read: element count mismatch");
    int64_t k = 0;
    for (int i = 0; i < L.rank; ++i) {
      if (nsec) nsec[i] = L.nsec(i);
      for (int j = 0; j < L.nsec(i); ++j, ++k) {
        if (sec_q) {
          const mpx::QN q = L.qsym ? L.qs[i][j] : mpx::QN();
          sec_q[2 * k] = q.p;
          sec_q[2 * k + 1] = q.s;
        }
        if (sec_dim) sec_dim[k] = L.ds[i][j];
      }
    }
    if (block_idx && !L.bidx.empty()) std::memcpy(block_idx, L.bidx.data(), L.bidx.size() * sizeof(int32_t));
    if (L.nelem > 0) {
      if (!data) throw mpx::Error(MPSXX_ERR_ARG, "qst_file_read: null data");
      std::memcpy(data, f.data.data(), f.data.size() * sizeof(double));
    }
  });
}

int mpsxx_tensor_save(mpsxx_tensor t, const char* path, void* stream) {
  return guard([&] {
    if (!path) throw mpx::Error(MPSXX_ERR_ARG, "tensor_save: null path");
    mpx::tensor_save(T(t), path, S(stream));
  });
}

int mpsxx_tensor_load(const char* path, mpsxx_tensor t, void* stream) {
  return guard([&] {
    if (!path) throw mpx::Error(MPSXX_ERR_ARG, "tensor_load: null path");
    mpx::tensor_load(path, T(t), S(stream));
  });
}

int mpsxx_getfile(const char* name, const char* prefix, int site, int istate, int jstate, char* buf, int64_t cap,
                  int64_t* len) {
  return guard_host([&] {
    if (!name || !prefix) throw mpx::Error(MPSXX_ERR_ARG, "getfile: null name or prefix");
    const std::string f = mpx::getfile(name, prefix, site, istate, jstate);
    if (len) *len = (int64_t)f.size();
    if (!buf || cap < (int64_t)f.size() + 1) throw mpx::Error(MPSXX_ERR_ARG, "getfile: buffer too small");
    std::memcpy(buf, f.c_str(), f.size() + 1);
  });
}

int mpsxx_contract(double alpha, mpsxx_tensor a, int conj_a, const int32_t* contract_a, mpsxx_tensor b, int conj_b,
                   const int32_t* contract_b, int32_t K, double beta, mpsxx_tensor c, void* stream) {
  return guard([&] {
    if (K <= 0 || !contract_a || !contract_b) throw mpx::Error(MPSXX_ERR_ARG, "contract: empty index lists");
    std::vector<int> ca(contract_a, contract_a + K), cb(contract_b, contract_b + K);
    mpx::op_contract(alpha, T(a), conj_a != 0, ca, T(b), conj_b != 0, cb, beta, T(c), S(stream));
  });
}

int mpsxx_gemm(int trans_a, int trans_b, double alpha, mpsxx_tensor a, int conj_a, mpsxx_tensor b, int conj_b,
               double beta, mpsxx_tensor c, void* stream) {
  return guard([&] {
    mpx::op_gemm(trans_a, trans_b, alpha, T(a), conj_a != 0, T(b), conj_b != 0, beta, T(c), S(stream));
  });
}

int mpsxx_tie(mpsxx_tensor x, const int32_t* tie, int32_t K, mpsxx_tensor y, void* stream) {
  return guard([&] {
    if (K < 2 || !tie) throw mpx::Error(MPSXX_ERR_ARG, "tie: need at least two indices");
    mpx::op_tie(T(x), std::vector<int>(tie, tie + K), T(y), S(stream));
  });
}

int mpsxx_copy_upcast(mpsxx_tensor x, mpsxx_tensor y, void* stream) {
  return guard([&] { mpx::op_copy_upcast(T(x), T(y), S(stream)); });
}

int mpsxx_dotc(mpsxx_tensor x, mpsxx_tensor y, double* result, void* stream) {
  return guard([&] { *result = mpx::op_dotc(T(x), T(y), S(stream)); });
}

int mpsxx_dotc_noalloc(mpsxx_tensor x, mpsxx_tensor y, double* result, void* stream) {
  return guard([&] {
    if (!result) throw mpx::Error(MPSXX_ERR_ARG, "Dotc (no allocation): null result pointer");
    *result = mpx::op_dotc_noalloc(T(x), T(y), S(stream));
  });
}

int mpsxx_nrm2(mpsxx_tensor x, double* result, void* stream) {
  return guard([&] { *result = mpx::op_nrm2(T(x), S(stream)); });
}

int mpsxx_axpy(double alpha, mpsxx_tensor x, mpsxx_tensor y, void* stream) {
  return guard([&] { mpx::op_axpy(alpha, T(x), T(y), S(stream)); });
}

int mpsxx_scal(double alpha, mpsxx_tensor x, void* stream) {
  return guard([&] { mpx::op_scal(alpha, T(x), S(stream)); });
}

int mpsxx_compute_sigma_vector_twosite(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                       mpsxx_tensor wfn0, mpsxx_tensor sgv0, int precision, void* stream) {
  return guard([&] {
    mpx::op_sigma_twosite(T(lmpo), T(rmpo), T(lopr), T(ropr), T(wfn0), T(sgv0), precision, S(stream));
  });
}

int mpsxx_compute_sigma_vector_onesite(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0,
                                       mpsxx_tensor sgv0, int precision, void* stream) {
  return guard([&] { mpx::op_sigma_onesite(T(mpo0), T(lopr), T(ropr), T(wfn0), T(sgv0), precision, S(stream)); });
}

int mpsxx_compute_sigma_vector_twosite_ex(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                          mpsxx_tensor wfn0, mpsxx_tensor sgv0, int precision, uint32_t convert_mask,
                                          void* stream) {
  return guard([&] {
    mpx::op_sigma_twosite(T(lmpo), T(rmpo), T(lopr), T(ropr), T(wfn0), T(sgv0), precision, S(stream), convert_mask);
  });
}

int mpsxx_compute_sigma_vector_merged_ex(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                         int precision, uint32_t convert_mask, void* stream) {
  return guard([&] {
    mpx::op_sigma_merged(T(lopr), T(ropr), T(wfn0), T(sgv0), precision, S(stream), convert_mask);
  });
}

int mpsxx_compute_sigma_vector_merged(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                      int precision, void* stream) {
  return guard([&] { mpx::op_sigma_merged(T(lopr), T(ropr), T(wfn0), T(sgv0), precision, S(stream)); });
}

int mpsxx_compute_sigma_vector_overlap(mpsxx_tensor lsov, mpsxx_tensor rsov, mpsxx_tensor wfn0, mpsxx_tensor sgv0,
                                       void* stream) {
  return guard([&] { mpx::op_sigma_overlap(T(lsov), T(rsov), T(wfn0), T(sgv0), S(stream)); });
}

int mpsxx_compute_sigma_vector_conj(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor wfn0,
                                    mpsxx_tensor sgv0, void* stream) {
  return guard([&] { mpx::op_sigma_conj(T(mpo0), T(lopr), T(ropr), T(wfn0), T(sgv0), S(stream)); });
}

int mpsxx_compute_diagonal_elements_twosite(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr,
                                            mpsxx_tensor ropr, mpsxx_tensor diag, void* stream) {
  return guard([&] { mpx::op_diag_twosite(T(lmpo), T(rmpo), T(lopr), T(ropr), T(diag), S(stream)); });
}

int mpsxx_compute_diagonal_elements_onesite(mpsxx_tensor mpo0, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                            mpsxx_tensor diag, void* stream) {
  return guard([&] { mpx::op_diag_onesite(T(mpo0), T(lopr), T(ropr), T(diag), S(stream)); });
}

int mpsxx_compute_diagonal_elements_merged(mpsxx_tensor lopr, mpsxx_tensor ropr, mpsxx_tensor diag, void* stream) {
  return guard([&] { mpx::op_diag_merged(T(lopr), T(ropr), T(diag), S(stream)); });
}

int mpsxx_renormalize(int forward, mpsxx_tensor mpo0, mpsxx_tensor opr0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                      mpsxx_tensor opr1, int precision, void* stream) {
  return guard([&] {
    mpx::op_renormalize(forward != 0, T(mpo0), T(opr0), T(bra0), T(ket0), T(opr1), precision, S(stream));
  });
}

int mpsxx_renormalize_overlap(int forward, mpsxx_tensor sov0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                              mpsxx_tensor sov1, void* stream) {
  return guard([&] { mpx::op_renormalize_overlap(forward != 0, T(sov0), T(bra0), T(ket0), T(sov1), S(stream)); });
}

int mpsxx_renormalize_merged(int forward, mpsxx_tensor opr0, mpsxx_tensor bra0, mpsxx_tensor ket0,
                             mpsxx_tensor opr1, int precision, void* stream) {
  return guard([&] {
    mpx::op_renormalize_merged(forward != 0, T(opr0), T(bra0), T(ket0), T(opr1), precision, S(stream));
  });
}

int mpsxx_compute_merged_operator(int forward, mpsxx_tensor xmpo, mpsxx_tensor xopr, mpsxx_tensor mopr,
                                  void* stream) {
  return guard([&] { mpx::op_merged_operator(forward != 0, T(xmpo), T(xopr), T(mopr), S(stream)); });
}

int mpsxx_canonicalize_twosite(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps, int max_states,
                               double* discarded, void* stream) {
  return guard([&] {
    const double d = mpx::op_canonicalize_twosite(forward != 0, T(wfnc), T(lmps), T(rmps), max_states, S(stream));
    if (discarded) *discarded = d;
  });
}

int mpsxx_canonicalize_twosite_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps,
                                  int max_states, double noise, int rule, double* discarded, void* stream) {
  return guard([&] {
    mpx::TruncationScope ts(rule);
    mpx::op_make_noise(T(wfnc), noise, S(stream));
    const double d = mpx::op_canonicalize_twosite(forward != 0, T(wfnc), T(lmps), T(rmps), max_states, S(stream));
    if (discarded) *discarded = d;
  });
}

int mpsxx_canonicalize_onesite_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmps, mpsxx_tensor rmps,
                                  int max_states, double noise, int rule, double* discarded, void* stream) {
  return guard([&] {
    mpx::TruncationScope ts(rule);
    mpx::Tensor& w = T(wfnc);
    mpx::Tensor& l = T(lmps);
    mpx::Tensor& r = T(rmps);
    hipStream_t st = S(stream);
    double d = 0.0;
    if (noise >= 1.0e-8) {
      // canonicalize.hpp:57-66: the two-site wavefunction with the neighbour, then the two-site form with noise
      mpx::Tensor two;
      if (forward) mpx::op_contract(1.0, w, false, {2}, r, false, {0}, 0.0, two, st);
      else mpx::op_contract(1.0, l, false, {2}, w, false, {0}, 0.0, two, st);
      mpx::op_make_noise(two, noise, st);
      mpx::Tensor U, V;
      d = mpx::op_canonicalize_twosite(forward != 0, two, U, V, max_states, st);
      l = std::move(U);
      r = std::move(V);
    } else {
      // canonicalize.hpp:68-84: forward lmps = U, rmps <- (s Vt) rmps; backward rmps = Vt, lmps <- lmps (U s)
      mpx::Tensor U, V, next;
      d = mpx::op_canonicalize_onesite(forward != 0, w, U, V, max_states, st);
      if (forward) {
        mpx::op_contract(1.0, V, false, {1}, r, false, {0}, 0.0, next, st);
        l = std::move(U);
        r = std::move(next);
      } else {
        mpx::op_contract(1.0, l, false, {2}, U, false, {0}, 0.0, next, st);
        l = std::move(next);
        r = std::move(V);
      }
    }
    mpx::hip_check(hipStreamSynchronize(st), "canonicalize");
    if (discarded) *discarded = d;
  });
}

int mpsxx_canonicalize_merged_ex(int forward, mpsxx_tensor wfnc, mpsxx_tensor lmat, mpsxx_tensor rmat,
                                 int max_states, double noise, int rule, double* discarded, void* stream) {
  return guard([&] {
    mpx::TruncationScope ts(rule);
    mpx::Tensor& w = T(wfnc);
    if (w.empty || w.L.rank != 2) throw mpx::Error(MPSXX_ERR_ARG, "canonicalize: expects a 2-index wavefunction");
    mpx::op_make_noise(w, noise, S(stream));
    const double d = mpx::op_canonicalize_onesite(forward != 0, w, T(lmat), T(rmat), max_states, S(stream));
    if (discarded) *discarded = d;
  });
}

int mpsxx_make_noise(mpsxx_tensor x, double noise, void* stream) {
  return guard([&] { mpx::op_make_noise(T(x), noise, S(stream)); });
}

int mpsxx_compute_merged_overlap(int forward, mpsxx_tensor xmpo, mpsxx_tensor xsov, mpsxx_tensor msov,
                                 void* stream) {
  return guard([&] { mpx::op_merged_overlap(forward != 0, T(xmpo), T(xsov), T(msov), S(stream)); });
}

int mpsxx_qst_merge(mpsxx_tensor x, int32_t nrow, int32_t ncol, mpsxx_tensor y, void* stream) {
  return guard([&] { mpx::op_qst_merge(T(x), nrow, ncol, T(y), S(stream)); });
}

int mpsxx_davidson_twosite(int ngroups, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                           const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, mpsxx_tensor diag, mpsxx_tensor wfnc,
                           double tolerance, int precision, double* energy, int32_t* n_sigma, void* stream) {
  return guard([&] {
    if (ngroups <= 0) throw mpx::Error(MPSXX_ERR_ARG, "Davidson: need at least one operator group");
    std::vector<mpx::DavGroup> g;
    for (int i = 0; i < ngroups; ++i) g.push_back({&T(lmpo[i]), &T(rmpo[i]), &T(lopr[i]), &T(ropr[i])});
    mpx::DavidsonStats st;
    const double e = mpx::davidson_twosite(g, T(diag), T(wfnc), tolerance, precision, st, S(stream));
    if (energy) *energy = e;
    if (n_sigma) *n_sigma = st.nsigma;
  });
}

int mpsxx_davidson_twosite_ex(int ngroups, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                              const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, mpsxx_tensor diag, mpsxx_tensor wfnc,
                              int nlower, const mpsxx_tensor* lower_sigma, const mpsxx_tensor* lower_overlap,
                              const double* lower_energy, double tolerance, int precision, double* energy,
                              int32_t* n_sigma, void* stream) {
  return guard([&] {
    if (ngroups <= 0 || !lopr || !ropr) throw mpx::Error(MPSXX_ERR_ARG, "Davidson: need at least one operator group");
    if ((lmpo == nullptr) != (rmpo == nullptr)) throw mpx::Error(MPSXX_ERR_ARG, "Davidson: lmpo and rmpo both or neither");
    if (nlower < 0 || (nlower > 0 && (!lower_sigma || !lower_overlap || !lower_energy)))
      throw mpx::Error(MPSXX_ERR_ARG, "Davidson: lower roots need sigma, overlap and energy arrays");
    std::vector<mpx::DavGroup> g;
    for (int i = 0; i < ngroups; ++i)
      g.push_back({lmpo ? &T(lmpo[i]) : nullptr, rmpo ? &T(rmpo[i]) : nullptr, &T(lopr[i]), &T(ropr[i])});
    std::vector<mpx::DavLowerRoot> low;
    for (int k = 0; k < nlower; ++k) low.push_back({&T(lower_sigma[k]), &T(lower_overlap[k]), lower_energy[k]});
    mpx::DavidsonStats st;
    const double e = mpx::davidson_twosite(g, T(diag), T(wfnc), tolerance, precision, st, S(stream), low);
    if (energy) *energy = e;
    if (n_sigma) *n_sigma = st.nsigma;
  });
}

int mpsxx_sigma_graphs(int enable) { return mpx::set_sigma_graphs(enable); }

int mpsxx_check_gesvd_args(int m, int n, int lda, int64_t strideA, int64_t strideS, int ldu, int64_t strideU, int ldv,
                           int64_t strideV, int64_t strideE, int batch, int64_t bytesA, int64_t bytesS,
                           int64_t bytesU, int64_t bytesV, int64_t bytesE) {
  return guard_host([&] {
    mpx::gesvd_argument_check(m, n, lda, strideA, strideS, ldu, strideU, ldv, strideV, strideE, batch, bytesA, bytesS,
                              bytesU, bytesV, bytesE);
  });
}

int mpsxx_sigma_twosite_stats(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                              mpsxx_tensor wfn0, int precision, int64_t* flops, int64_t* bytes, int32_t* nsteps) {
  return guard([&] {
    mpx::Plan& P = mpx::plan_sigma_twosite(T(lmpo).L, T(rmpo).L, T(lopr).L, T(ropr).L, T(wfn0).L, nullptr, precision);
    if (flops) *flops = P.flops;
    if (bytes) *bytes = P.bytes;
    if (nsteps) *nsteps = (int32_t)P.steps.size();
  });
}

int mpsxx_sigma_twosite_profile(mpsxx_tensor lmpo, mpsxx_tensor rmpo, mpsxx_tensor lopr, mpsxx_tensor ropr,
                                mpsxx_tensor wfn0, int precision, int reps, int32_t max_entries, double* ms,
                                int64_t* flops, int64_t* bytes, int32_t* kernel, int32_t* step, int32_t* nentries,
                                void* stream) {
  return guard([&] {
    if (reps <= 0) throw mpx::Error(MPSXX_ERR_ARG, "profile: reps must be positive");
    mpx::Plan& P = mpx::plan_sigma_twosite(T(lmpo).L, T(rmpo).L, T(lopr).L, T(ropr).L, T(wfn0).L, nullptr, precision);
    mpx::DevBuf out(std::max<int64_t>(P.out.nelem, 1) * sizeof(double));
    std::vector<const double*> in{T(lmpo).ptr(), T(rmpo).ptr(), T(lopr).ptr(), T(ropr).ptr(), T(wfn0).ptr()};
    // one full run converts the FP32 inputs and fills every intermediate
    mpx::run_plan(P, in, out.as<double>(), S(stream));
    std::vector<mpx::PartTiming> t = mpx::profile_plan_parts(P, in, out.as<double>(), reps, S(stream));
    if (nentries) *nentries = (int32_t)t.size();
    if ((int)t.size() > max_entries) throw mpx::Error(MPSXX_ERR_ARG, "profile: max_entries too small");
    for (size_t i = 0; i < t.size(); ++i) {
      if (ms) ms[i] = t[i].ms;
      if (flops) flops[i] = t[i].part.flops;
      if (bytes) bytes[i] = t[i].part.bytes;
      if (kernel) kernel[i] = t[i].part.kind;
      if (step) step[i] = t[i].step;
    }
  });
}

int mpsxx_last_plan_profile(int reps, int32_t max_entries, double* ms, int64_t* flops, int64_t* bytes,
                            int32_t* kernel, int32_t* step, int32_t* nentries, int64_t* plan_flops, void* stream) {
  return guard([&] {
    if (reps <= 0) throw mpx::Error(MPSXX_ERR_ARG, "profile: reps must be positive");
    std::vector<const double*> in;
    mpx::Plan& P = mpx::last_plan(in);
    if (plan_flops) *plan_flops = P.flops;
    mpx::DevBuf out(std::max<int64_t>(P.out.nelem, 1) * sizeof(double));
    mpx::run_plan(P, in, out.as<double>(), S(stream));
    std::vector<mpx::PartTiming> t = mpx::profile_plan_parts(P, in, out.as<double>(), reps, S(stream));
    if (nentries) *nentries = (int32_t)t.size();
    if ((int)t.size() > max_entries) throw mpx::Error(MPSXX_ERR_ARG, "profile: max_entries too small");
    for (size_t i = 0; i < t.size(); ++i) {
      if (ms) ms[i] = t[i].ms;
      if (flops) flops[i] = t[i].part.flops;
      if (bytes) bytes[i] = t[i].part.bytes;
      if (kernel) kernel[i] = t[i].part.kind;
      if (step) step[i] = t[i].step;
    }
  });
}

int mpsxx_last_plan_tiles(int64_t tiles[8]) {
  return guard([&] {
    if (!tiles) throw mpx::Error(MPSXX_ERR_ARG, "plan tiles: null output");
    std::vector<const double*> in;
    const mpx::Plan& P = mpx::last_plan(in);
    for (int k = 0; k < 8; ++k) tiles[k] = 0;
    for (const mpx::Step& st : P.steps) {
      if (st.kind == mpx::Step::GEMM) {
        tiles[0] += st.nwork;
        tiles[1] += st.nwork_red;
        tiles[2] += st.nwork_small;
        tiles[3] += st.nwork_split;
        tiles[4] += st.nwork_split_red;
        tiles[6] += st.nwork2;
        tiles[7] += st.nwork_red2;
      } else if (st.kind == mpx::Step::AXPY) {
        tiles[5] += st.nwork;
      }
    }
  });
}

// ---------------------------------------------------------------------------------------------------------------------
// quantum-chemistry MPO

int mpsxx_qc_mpo_create(int32_t norb, const double* oneint, const double* twoint, double ecore, int swap_sweep,
                        mpsxx_qcmpo* out) {
  return guard_host([&] {
    if (!out || !oneint || !twoint) throw mpx::Error(MPSXX_ERR_ARG, "qc mpo: null argument");
    auto* m = new mpsxx_qcmpo_s;
    try {
      m->m = mpx::build_qc_mpo(norb, oneint, twoint, ecore, swap_sweep != 0);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int mpsxx_plan_cache_stats(int64_t* hits, int64_t* builds, int64_t* evictions, double* build_ms, int64_t* plans,
                           int64_t* device_bytes, int64_t* releases, int64_t* restores) {
  return guard_host([&] {
    const mpx::PlanCacheStats st = mpx::plan_cache_stats();
    if (hits) *hits = st.hits;
    if (builds) *builds = st.builds;
    if (evictions) *evictions = st.evictions;
    if (build_ms) *build_ms = st.build_ms;
    if (plans) *plans = st.plans;
    if (device_bytes) *device_bytes = st.device_bytes;
    if (releases) *releases = st.releases;
    if (restores) *restores = st.restores;
  });
}

int mpsxx_release_caches(void) {
  return guard([&] {
    mpx::plan_cache_clear();
    mpx::merged_plan_cache_clear();
    mpx::device_pool_trim();
  });
}

int mpsxx_plan_cache_collisions(int64_t* collisions) {
  return guard_host([&] {
    if (!collisions) throw mpx::Error(MPSXX_ERR_ARG, "plan cache collisions: null output");
    *collisions = mpx::plan_cache_stats().collisions;
  });
}

int mpsxx_set_structure_hash_bits(int32_t bits) {
  return guard_host([&] {
    if (bits < 1 || bits > 64) throw mpx::Error(MPSXX_ERR_ARG, "structure hash bits: expected 1..64");
    mpx::set_structure_hash_bits(bits);
  });
}

int mpsxx_device_memory_stats(int64_t* mallocs, int64_t* frees, double* malloc_ms, double* free_ms, int64_t* reuses,
                              int64_t* pool_syncs, int64_t* bytes_live, int64_t* cached_bytes) {
  return guard_host([&] {
    const mpx::DeviceMemoryStats st = mpx::device_memory_stats();
    if (mallocs) *mallocs = st.mallocs;
    if (frees) *frees = st.frees;
    if (malloc_ms) *malloc_ms = st.malloc_ms;
    if (free_ms) *free_ms = st.free_ms;
    if (reuses) *reuses = st.reuses;
    if (pool_syncs) *pool_syncs = st.pool_syncs;
    if (bytes_live) *bytes_live = st.bytes_live;
    if (cached_bytes) *cached_bytes = st.cached_bytes;
  });
}

int mpsxx_planner_threads(int threads, int64_t min_pairs_per_thread) {
  return guard_host([&] { mpx::set_planner_threads(threads, min_pairs_per_thread); });
}

int mpsxx_contract_structure(const mpsxx_qst_desc* a, int conj_a, const int32_t* ca, const mpsxx_qst_desc* b,
                             int conj_b, const int32_t* cb, int32_t k, int64_t* nblocks, int64_t* npairs,
                             int32_t* block_idx, int64_t* pair_begin, int64_t* pair_a, int64_t* pair_b) {
  return guard_host([&] {
    if (!ca || !cb || k < 1 || !nblocks || !npairs) throw mpx::Error(MPSXX_ERR_ARG, "contract_structure: bad arguments");
    const mpx::Layout A = layout_from_desc(a), B = layout_from_desc(b);
    std::vector<int> ia(ca, ca + k), ib(cb, cb + k);
    mpx::ContractSym sym = mpx::contract_symbolic(mpx::View{&A, conj_a != 0}, ia, mpx::View{&B, conj_b != 0}, ib, nullptr);
    const mpx::Layout& C = sym.C;
    *nblocks = C.nblocks();
    *npairs = (int64_t)sym.pairs.size();
    if (!block_idx) return;
    for (int64_t x = 0; x < C.nblocks(); ++x)
      for (int i = 0; i < C.rank; ++i) block_idx[x * C.rank + i] = C.index(x)[i];
    // without an existing output, the planner's tasks are the output blocks in order
    if ((int64_t)sym.tasks.size() != C.nblocks()) throw mpx::Error(MPSXX_ERR_ARG, "contract_structure: task mismatch");
    // the block holding element offset `off` (offsets ascend with the block number; size-0 blocks share the offset
    // of the next block and take part in no pair)
    auto block_at = [](const mpx::Layout& L, int64_t off) {
      int64_t x = std::lower_bound(L.off.begin(), L.off.end(), off) - L.off.begin();
      while (x < L.nblocks() && L.off[x] == off && L.bsize[x] == 0) ++x;
      return x < L.nblocks() && L.off[x] == off ? x : (int64_t)-1;
    };
    for (int64_t x = 0; x < C.nblocks(); ++x) {
      const mpx::GemmTask& t = sym.tasks[x];
      pair_begin[x] = t.pair_begin;
      for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p) {
        pair_a[p] = block_at(A, sym.pairs[p].a_off);
        pair_b[p] = block_at(B, sym.pairs[p].b_off);
      }
    }
    pair_begin[C.nblocks()] = (int64_t)sym.pairs.size();
  });
}

int mpsxx_qc_mpo_info(mpsxx_qcmpo m, int32_t* bond_dims, int64_t* nentries_total) {
  return guard_host([&] {
    const mpx::QcMpo& x = QM(m);
    int64_t tot = 0;
    for (int k = 0; k < x.K; ++k) {
      if (bond_dims) bond_dims[k] = (int32_t)x.sites[k].lq.size();
      tot += (int64_t)x.sites[k].mat.size();
    }
    if (bond_dims) bond_dims[x.K] = (int32_t)x.sites[x.K - 1].rq.size();
    if (nentries_total) *nentries_total = tot;
  });
}

int mpsxx_qc_mpo_site_entries(mpsxx_qcmpo m, int32_t site, int32_t* nl, int32_t* nr, int64_t* nentries, int32_t* lq,
                              int32_t* rq, int32_t* wl, int32_t* wr, double* mats) {
  return guard_host([&] {
    const mpx::QcMpo& x = QM(m);
    if (site < 0 || site >= x.K) throw mpx::Error(MPSXX_ERR_ARG, "qc mpo: site out of range");
    const mpx::QcMpoSite& w = x.sites[site];
    if (nl) *nl = (int32_t)w.lq.size();
    if (nr) *nr = (int32_t)w.rq.size();
    if (nentries) *nentries = (int64_t)w.mat.size();
    for (size_t i = 0; lq && i < w.lq.size(); ++i) { lq[2 * i] = w.lq[i].p; lq[2 * i + 1] = w.lq[i].s; }
    for (size_t i = 0; rq && i < w.rq.size(); ++i) { rq[2 * i] = w.rq[i].p; rq[2 * i + 1] = w.rq[i].s; }
    for (size_t e = 0; e < w.mat.size(); ++e) {
      if (wl) wl[e] = w.wl[e];
      if (wr) wr[e] = w.wr[e];
      if (mats) std::memcpy(mats + 16 * e, w.mat[e].data(), 16 * sizeof(double));
    }
  });
}

int mpsxx_qc_mpo_site(mpsxx_qcmpo m, int32_t site, mpsxx_tensor out, void* stream) {
  return guard([&] {
    const mpx::QcMpo& x = QM(m);
    if (site < 0 || site >= x.K) throw mpx::Error(MPSXX_ERR_ARG, "qc mpo: site out of range");
    mpx::qc_mpo_tensor(x.sites[site], T(out), S(stream));
  });
}

int mpsxx_qc_mpo_gradient(mpsxx_qcmpo m, int32_t nsites, const mpsxx_tensor* grads, double* dh, double* dlambda,
                          double* energy, void* stream) {
  return guard([&] {
    const mpx::QcMpo& x = QM(m);
    if (nsites != x.K || !grads || !dh) throw mpx::Error(MPSXX_ERR_ARG, "qc gradient: one tensor per site, dh required");
    std::vector<const mpx::Tensor*> g;
    for (int k = 0; k < nsites; ++k) g.push_back(&T(grads[k]));
    mpx::qc_mpo_gradient(x, g, dh, dlambda, energy, S(stream));
  });
}

int mpsxx_qc_mpo_free(mpsxx_qcmpo m) {
  return guard_host([&] {
    if (!m) return;
    QM(m);
    delete m;
  });
}

// ---------------------------------------------------------------------------------------------------------------------
// batched sigma vectors

int mpsxx_batch_sigma_twosite_create(int32_t n, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                                     const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, const mpsxx_tensor* wfn0,
                                     const mpsxx_tensor* sgv0, int precision, mpsxx_batch* out, void* stream) {
  return guard([&] {
    if (!out || n <= 0 || !lmpo || !rmpo || !lopr || !ropr || !wfn0 || !sgv0)
      throw mpx::Error(MPSXX_ERR_ARG, "batch create: bad arguments");
    std::vector<mpx::BatchSigma::Operand6> inst(n);
    for (int i = 0; i < n; ++i)
      inst[i] = {&T(lmpo[i]), &T(rmpo[i]), &T(lopr[i]), &T(ropr[i]), &T(wfn0[i]), &T(sgv0[i])};
    auto* b = new mpsxx_batch_s;
    try {
      b->b = std::make_unique<mpx::BatchSigma>(inst, precision, S(stream));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int mpsxx_batch_davidson_twosite(int32_t n, const mpsxx_tensor* lmpo, const mpsxx_tensor* rmpo,
                                 const mpsxx_tensor* lopr, const mpsxx_tensor* ropr, const mpsxx_tensor* diag,
                                 const mpsxx_tensor* wfnc, double tolerance, int precision, int use_graph,
                                 double* energies, int32_t* n_sigma, int32_t* n_sigma_low, double* residuals,
                                 mpsxx_batch_davidson_stats* stats, void* stream) {
  return guard([&] {
    if (n <= 0 || !lmpo || !rmpo || !lopr || !ropr || !diag || !wfnc)
      throw mpx::Error(MPSXX_ERR_ARG, "batch Davidson: bad arguments");
    std::vector<mpx::BatchDavInstance> inst(n);
    for (int i = 0; i < n; ++i)
      inst[i] = {&T(lmpo[i]), &T(rmpo[i]), &T(lopr[i]), &T(ropr[i]), &T(diag[i]), &T(wfnc[i])};
    mpx::BatchDavidsonStats st;
    mpx::batch_davidson_twosite(inst, tolerance, precision, use_graph != 0, st, S(stream));
    for (int i = 0; i < n; ++i) {
      if (energies) energies[i] = st.energy[i];
      if (n_sigma) n_sigma[i] = st.nsigma[i];
      if (n_sigma_low) n_sigma_low[i] = st.nsigma_low[i];
      if (residuals) residuals[i] = st.rnorm[i];
    }
    if (stats) {
      stats->batch_sigmas = st.batch_sigmas;
      stats->batch_sigmas_low = st.batch_sigmas_low;
      stats->macro_iterations = st.macro_iterations;
      stats->converged = 0;
      for (int c : st.converged) stats->converged += c;
      stats->sigma_ms = st.sigma_ms;
      stats->seconds = st.seconds;
    }
  });
}

int mpsxx_batch_sigma_twosite_run(mpsxx_batch b, uint32_t convert_mask, int use_graph, void* stream) {
  return guard([&] { B(b).run(convert_mask, use_graph != 0, S(stream)); });
}

int mpsxx_batch_time(mpsxx_batch b, int32_t reps, uint32_t convert_mask, int use_graph, double* ms, void* stream) {
  return guard([&] {
    const double t = B(b).time(reps, convert_mask, use_graph != 0, S(stream));
    if (ms) *ms = t;
  });
}

int mpsxx_batch_profile(mpsxx_batch b, int32_t reps, int32_t max_entries, double* ms, int64_t* flops, int64_t* bytes,
                        int32_t* kernel, int32_t* step, int32_t* nentries, void* stream) {
  return guard([&] {
    const std::vector<mpx::PartTiming> t = B(b).profile(reps, S(stream));
    if (nentries) *nentries = (int32_t)t.size();
    if ((int)t.size() > max_entries) throw mpx::Error(MPSXX_ERR_ARG, "batch profile: max_entries too small");
    for (size_t i = 0; i < t.size(); ++i) {
      if (ms) ms[i] = t[i].ms;
      if (flops) flops[i] = t[i].part.flops;
      if (bytes) bytes[i] = t[i].part.bytes;
      if (kernel) kernel[i] = t[i].part.kind;
      if (step) step[i] = t[i].step;
    }
  });
}

int mpsxx_batch_tune(mpsxx_batch b, int32_t reps, int32_t max_steps, int32_t* policy, void* stream) {
  return guard([&] {
    mpx::BatchSigma& x = B(b);
    if (x.nsteps() > max_steps) throw mpx::Error(MPSXX_ERR_ARG, "batch tune: max_steps too small");
    const std::vector<mpx::GemmPolicy> p = x.tune(reps, S(stream));
    for (size_t k = 0; policy && k < p.size(); ++k) {
      policy[4 * k] = p[k].small_limit;
      policy[4 * k + 1] = p[k].split_units;
      policy[4 * k + 2] = p[k].max_groups;
      policy[4 * k + 3] = p[k].split_big;
    }
  });
}

int mpsxx_batch_set_policy(mpsxx_batch b, int32_t nsteps, const int32_t* policy) {
  return guard([&] {
    mpx::BatchSigma& x = B(b);
    if (nsteps != x.nsteps() || !policy) throw mpx::Error(MPSXX_ERR_ARG, "batch set_policy: one policy per step");
    std::vector<mpx::GemmPolicy> p(nsteps);
    for (int k = 0; k < nsteps; ++k) {
      const int32_t* q = policy + 4 * k;
      if (q[0] < 0 || q[1] < 1 || q[2] < 0 || q[3] < 0)
        throw mpx::Error(MPSXX_ERR_ARG, "batch set_policy: negative limit, threshold, grid cap or split target");
      p[k] = mpx::default_policy(x.precision());  // the 32x32 split and 128x128 settings are not part of the layout
      p[k].small_limit = q[0];
      p[k].split_units = q[1];
      p[k].max_groups = q[2];
      p[k].split_big = q[3];
    }
    x.set_policies(p);
  });
}

int mpsxx_batch_info(mpsxx_batch b, int64_t* flops, int64_t* bytes, int32_t* nsteps, int64_t* work_items) {
  return guard([&] {
    mpx::BatchSigma& x = B(b);
    if (flops) *flops = x.flops();
    if (bytes) *bytes = x.bytes();
    if (nsteps) *nsteps = x.nsteps();
    if (work_items) *work_items = x.work_items();
  });
}

int mpsxx_batch_free(mpsxx_batch b) {
  return guard([&] {
    if (!b) return;
    B(b);
    delete b;
  });
}

}  // extern "C"
