// Block-sparse contraction planner (host, integer-only) and plan execution.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <mutex>
#include <map>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <atomic>
#include <climits>
#include <exception>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "mpsxx_amd.h"

namespace mpx {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory) throw Error(MPSXX_ERR_OOM, std::string(what) + ": out of device memory");
    throw Error(MPSXX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// Layout

void Layout::set_shape(int rank_, bool qsym_, QN q_, std::vector<std::vector<QN>> qs_,
                       std::vector<std::vector<int>> ds_) {
  if (rank_ < 1 || rank_ > kMaxRank) throw Error(MPSXX_ERR_ARG, "rank out of range");
  rank = rank_;
  qsym = qsym_;
  q = q_;
  qs = std::move(qs_);
  ds = std::move(ds_);
  if ((int)ds.size() != rank) throw Error(MPSXX_ERR_ARG, "dshape rank mismatch");
  if (qsym) {
    if ((int)qs.size() != rank) throw Error(MPSXX_ERR_ARG, "qshape rank mismatch");
    for (int i = 0; i < rank; ++i)
      if (qs[i].size() != ds[i].size()) throw Error(MPSXX_ERR_ARG, "qshape/dshape sector count mismatch");
  } else {
    qs.assign(rank, {});
    for (int i = 0; i < rank; ++i) qs[i].assign(ds[i].size(), QN());
  }
  for (int i = 0; i < rank; ++i)
    for (int d : ds[i])
      if (d < 0) throw Error(MPSXX_ERR_ARG, "negative sector dimension");
  tag.clear();
  bidx.clear();
  off.clear();
  bsize.clear();
  nelem = 0;
}

void Layout::finalize() {
  const int64_t nb = (int64_t)tag.size();
  bool ordered = true;  // blocks added in strictly ascending tag order (the planner's outputs) need no sort
  for (int64_t k = 1; k < nb && ordered; ++k) ordered = tag[k - 1] < tag[k];
  if (!ordered) {
    std::vector<int64_t> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return tag[x] < tag[y]; });
    std::vector<int64_t> t2;
    std::vector<int32_t> b2;
    t2.reserve(nb);
    b2.reserve(nb * rank);
    for (int64_t k = 0; k < nb; ++k) {
      const int64_t o = order[k];
      if (!t2.empty() && t2.back() == tag[o]) continue;
      t2.push_back(tag[o]);
      b2.insert(b2.end(), bidx.begin() + o * rank, bidx.begin() + (o + 1) * rank);
    }
    tag.swap(t2);
    bidx.swap(b2);
  }
  off.resize(tag.size());
  bsize.resize(tag.size());
  nelem = 0;
  for (size_t b = 0; b < tag.size(); ++b) {
    off[b] = nelem;
    bsize[b] = block_elems(index(b));
    nelem += bsize[b];
  }
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, (uint64_t)rank);
  h = fnv(h, (uint64_t)qsym);
  h = fnv(h, (uint64_t)(uint32_t)q.p);
  h = fnv(h, (uint64_t)(uint32_t)q.s);
  for (int i = 0; i < rank; ++i) {
    h = fnv(h, ds[i].size());
    for (size_t k = 0; k < ds[i].size(); ++k) {
      h = fnv(h, (uint64_t)(uint32_t)ds[i][k]);
      h = fnv(h, (uint64_t)(uint32_t)qs[i][k].p);
      h = fnv(h, (uint64_t)(uint32_t)qs[i][k].s);
    }
  }
  for (int64_t t : tag) h = fnv(h, (uint64_t)t);
  hash = h & structure_hash_mask();
}

namespace {
std::atomic<uint64_t> g_hash_mask{~0ull};
}
uint64_t structure_hash_mask() { return g_hash_mask.load(std::memory_order_relaxed); }
void set_structure_hash_bits(int bits) {
  g_hash_mask.store(bits >= 64 ? ~0ull : ((1ull << std::max(bits, 0)) - 1ull), std::memory_order_relaxed);
}

// ---------------------------------------------------------------------------------------------------------------------
// DevBuf / Tensor

DevBuf::~DevBuf() { release(); }

DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    release();
    p = o.p;
    bytes = o.bytes;
    cls = o.cls;
    o.p = nullptr;
    o.bytes = 0;
    o.cls = 0;
  }
  return *this;
}

// Caching allocator behind DevBuf.  hipMalloc of large buffers costs ~1 ms per GB (page mapping) and hipFree
// synchronizes the device, and a DMRG sweep allocates thousands of them (outputs, scratch, task lists).  Released
// buffers are kept in a size-class pool (up to MPSXX_DEVICE_CACHE_GB, default 31 % of the device's memory) and handed
// out again.  A buffer released since the last
// device-wide synchronization may still be read by queued work on any stream, so handing one out is preceded by
// hipDeviceSynchronize(), after which every pooled buffer is safe; allocations that find no pooled buffer call
// hipMalloc and, if the device is out of memory, return the pool's buffers to the device and retry.
size_t device_memory_fraction(double frac, double fallback_gb) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) {
    (void)hipGetLastError();
    return (size_t)(fallback_gb * 1073741824.0);
  }
  return (size_t)(frac * (double)total_b);
}

namespace {

struct DevicePool {
  std::mutex mu;
  std::multimap<size_t, void*> safe, pending;  // size class -> buffer
  size_t cached = 0;
  DeviceMemoryStats st;

  static size_t size_class(size_t n) {
    if (n <= ((size_t)1 << 20)) {
      size_t c = 256;
      while (c < n) c <<= 1;
      return c;
    }
    const size_t g = (size_t)2 << 20;
    return (n + g - 1) / g * g;
  }
  static size_t limit() {
    static const size_t b = [] {
      if (const char* e = std::getenv("MPSXX_DEVICE_CACHE_GB"))
        return (size_t)(std::max(0.0, std::atof(e)) * 1073741824.0);
      return device_memory_fraction(0.31, 32.0);
    }();
    return b;
  }
  // a pooled buffer of class c (at most 2c), nullptr if none
  static void* take(std::multimap<size_t, void*>& m, size_t c, size_t& got) {
    auto it = m.lower_bound(c);
    if (it == m.end() || it->first > 2 * c) return nullptr;
    void* p = it->second;
    got = it->first;
    m.erase(it);
    return p;
  }
  void drain(std::multimap<size_t, void*>& m) {
    for (auto& kv : m) {
      (void)hipFree(kv.second);
      cached -= kv.first;
    }
    m.clear();
  }
  void* alloc(size_t n, size_t& got) {
    std::lock_guard<std::mutex> lk(mu);
    const size_t c = size_class(n);
    void* p = take(safe, c, got);
    if (!p && !pending.empty()) {
      auto it = pending.lower_bound(c);
      if (it != pending.end() && it->first <= 2 * c) {
        hip_check(hipDeviceSynchronize(), "device pool sync");
        ++st.pool_syncs;
        safe.insert(pending.begin(), pending.end());
        pending.clear();
        p = take(safe, c, got);
      }
    }
    if (p) {
      cached -= got;
      ++st.reuses;
      st.bytes_live += (int64_t)got;
      return p;
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipMalloc(&p, c);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
      (void)hipGetLastError();
      hip_check(hipDeviceSynchronize(), "device pool sync");
      drain(safe);
      drain(pending);
      e = hipMalloc(&p, c);
    }
    hip_check(e, "hipMalloc");
    st.malloc_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ++st.mallocs;
    got = c;
    st.bytes_live += (int64_t)got;
    return p;
  }
  void release(void* p, size_t c) {
    std::lock_guard<std::mutex> lk(mu);
    st.bytes_live -= (int64_t)c;
    if (cached + c <= limit()) {
      pending.insert({c, p});
      cached += c;
      return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    (void)hipFree(p);
    st.free_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ++st.frees;
  }
};

DevicePool& pool() {
  static DevicePool* p = new DevicePool();  // never destroyed: buffers may be released during static destruction
  return *p;
}

}  // namespace

void device_pool_trim() {
  DevicePool& P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  hip_check(hipDeviceSynchronize(), "device pool trim");
  P.drain(P.safe);
  P.drain(P.pending);
}

DeviceMemoryStats device_memory_stats() {
  DevicePool& P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  DeviceMemoryStats st = P.st;
  st.cached_bytes = (int64_t)P.cached;
  return st;
}

void DevBuf::alloc(size_t n) {
  release();
  if (n == 0) return;
  size_t got = 0;
  p = pool().alloc(n, got);
  bytes = n;
  cls = got;
}

void DevBuf::release() {
  if (p) pool().release(p, cls);
  p = nullptr;
  bytes = 0;
  cls = 0;
}

namespace {
struct TensorRegistry {
  std::mutex mu;
  std::unordered_set<const Tensor*> live;
};
TensorRegistry& tensor_registry() {
  static TensorRegistry* r = new TensorRegistry;  // never destroyed: tensors may outlive static destruction order
  return *r;
}
void register_tensor(const Tensor* t) {
  TensorRegistry& r = tensor_registry();
  std::lock_guard<std::mutex> g(r.mu);
  r.live.insert(t);
}
}  // namespace

Tensor::Tensor() { register_tensor(this); }
Tensor::Tensor(Tensor&& o) noexcept : L(std::move(o.L)), data(std::move(o.data)), empty(o.empty) {
  o.empty = true;
  register_tensor(this);
}
Tensor::~Tensor() {
  TensorRegistry& r = tensor_registry();
  std::lock_guard<std::mutex> g(r.mu);
  r.live.erase(this);
}
bool Tensor::alive(const Tensor* t) {
  TensorRegistry& r = tensor_registry();
  std::lock_guard<std::mutex> g(r.mu);
  return r.live.count(t) != 0;
}

void Tensor::set_layout(const Layout& l, hipStream_t s) {
  L = l;
  data.alloc(std::max<int64_t>(L.nelem, 1) * sizeof(double));
  hip_check(hipMemsetAsync(data.p, 0, std::max<int64_t>(L.nelem, 1) * sizeof(double), s), "memset");
  empty = false;
}

void copy_segments(const std::vector<CopySeg>& segs, const double* x, double* y, hipStream_t s) {
  if (segs.empty()) return;
  std::vector<Chunk> chunks;
  for (size_t i = 0; i < segs.size(); ++i)
    for (int c = 0; c < segs[i].len; c += kAxpyChunk) chunks.push_back({(int32_t)i, c});
  // the uploaded lists go back to the device pool when this returns; the pool reuses them only after a device-wide
  // synchronization, so the launch below still reads them
  DevBuf ds = upload_vec(segs), dc = upload_vec(chunks);
  launch_copy<double, double>(dc.as<Chunk>(), (int)chunks.size(), ds.as<CopySeg>(), x, y, 1.0, 0.0, s);
}

void Tensor::relayout(const Layout& l, hipStream_t s) {
  if (empty) {
    set_layout(l, s);
    return;
  }
  if (l.same_structure(L)) return;
  DevBuf nb(std::max<int64_t>(l.nelem, 1) * sizeof(double));
  hip_check(hipMemsetAsync(nb.p, 0, std::max<int64_t>(l.nelem, 1) * sizeof(double), s), "memset");
  std::vector<CopySeg> segs;
  for (int64_t b = 0; b < L.nblocks(); ++b) {
    const int64_t j = l.find(L.tag[b]);
    if (j < 0 || L.bsize[b] == 0) continue;
    segs.push_back({l.off[j], L.off[b], (int32_t)L.bsize[b], 1});
  }
  copy_segments(segs, ptr(), nb.as<double>(), s);
  data = std::move(nb);
  L = l;
}

// ---------------------------------------------------------------------------------------------------------------------
// Symbolic contraction

namespace {

// Row-major strides of one block.
void block_strides(const Layout& L, const int32_t* idx, int64_t* dims, int64_t* st) {
  int64_t s = 1;
  for (int i = L.rank - 1; i >= 0; --i) {
    dims[i] = L.ds[i][idx[i]];
    st[i] = s;
    s *= dims[i];
  }
}

// A group of indices viewed as one flattened dimension of a permuted block: the non-unit dimensions must appear in
// the same relative order as stored and be packed (stride[g_i] == stride[g_{i+1}] * dim[g_{i+1}]); then the
// flattened index advances with the stride of the last non-unit member.  This is exactly the memory walk of the
// reference's Permute copy (QSTREINDEX.h:22-50) for the index orders that occur on the path.
bool group_view(const std::vector<int>& g, const int64_t* dims, const int64_t* st, int64_t& size, int64_t& stride) {
  size = 1;
  stride = 0;
  int last = -1;
  for (int i : g) {
    if (dims[i] == 1) continue;
    if (dims[i] == 0) { size = 0; return true; }
    if (last >= 0 && st[last] != st[i] * dims[i]) return false;
    size *= dims[i];
    last = i;
  }
  if (last >= 0) stride = st[last];
  return true;
}

}  // namespace

// Host threads of the planner: set_planner_threads, else MPSXX_HOST_THREADS, else the hardware concurrency (at most
// 16); a contraction uses at most one thread per min_pairs candidate pairs.
namespace {
std::atomic<int> g_threads{0};
std::atomic<int64_t> g_grain{200000};
}  // namespace

void set_planner_threads(int threads, int64_t min_pairs) {
  g_threads = std::max(0, threads);
  g_grain = std::max<int64_t>(1, min_pairs);
}

int host_threads() {
  if (g_threads > 0) return g_threads;
  static const int n = [] {
    if (const char* e = std::getenv("MPSXX_HOST_THREADS")) return std::max(1, std::atoi(e));
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::min(16u, std::max(1u, h));
  }();
  return n;
}

// Run f(0..n-1), f(0) on the calling thread; the first exception is rethrown after all have finished.
void parallel_for(int n, const std::function<void(int)>& f) {
  if (n <= 1) {
    if (n == 1) f(0);
    return;
  }
  std::vector<std::exception_ptr> err(n);
  std::vector<std::thread> th;
  th.reserve(n - 1);
  for (int c = 1; c < n; ++c)
    th.emplace_back([&, c] {
      try {
        f(c);
      } catch (...) {
        err[c] = std::current_exception();
      }
    });
  try {
    f(0);
  } catch (...) {
    err[0] = std::current_exception();
  }
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

void gemm_indices(int transa, int transb, int NA, int NB, int NC, std::vector<int>& ca, std::vector<int>& cb) {
  const int K = (NA + NB - NC) / 2;
  if (K <= 0 || 2 * K != NA + NB - NC) throw Error(MPSXX_ERR_ARG, "Gemm: inconsistent ranks");
  ca.clear();
  cb.clear();
  // btas::Gemm contracts the last K indices of A (first K if transposed) with the first K of B (last K if
  // transposed); QSTBLAS.h:213-225 with transposed_view (STArray.h:558-577)
  for (int k = 0; k < K; ++k) ca.push_back(transa ? k : NA - K + k);
  for (int k = 0; k < K; ++k) cb.push_back(transb ? NB - K + k : k);
}

ContractSym contract_symbolic(const View& Av, const std::vector<int>& ca, const View& Bv, const std::vector<int>& cb,
                              const Layout* existing, const Layout* mask) {
  const Layout& a = *Av.L;
  const Layout& b = *Bv.L;
  const int K = (int)ca.size();
  const int NA = a.rank, NB = b.rank;
  if ((int)cb.size() != K || K <= 0 || K >= NA || K >= NB)
    throw Error(MPSXX_ERR_UNSUPPORTED, "Contract: only 0 < K < rank(A), rank(B) is supported");
  std::vector<char> inA(NA, 0), inB(NB, 0);
  for (int k = 0; k < K; ++k) {
    if (ca[k] < 0 || ca[k] >= NA || inA[ca[k]]) throw Error(MPSXX_ERR_ARG, "Contract: bad A index list");
    if (cb[k] < 0 || cb[k] >= NB || inB[cb[k]]) throw Error(MPSXX_ERR_ARG, "Contract: bad B index list");
    inA[ca[k]] = 1;
    inB[cb[k]] = 1;
  }
  std::vector<int> ua, ub;
  for (int i = 0; i < NA; ++i)
    if (!inA[i]) ua.push_back(i);
  for (int i = 0; i < NB; ++i)
    if (!inB[i]) ub.push_back(i);

  const bool qsym = a.qsym && b.qsym;
  for (int k = 0; k < K; ++k) {
    const int ia = ca[k], ib = cb[k];
    if (a.ds[ia].size() != b.ds[ib].size()) throw Error(MPSXX_ERR_QN, "Contract: contracted sector counts differ");
    for (size_t s = 0; s < a.ds[ia].size(); ++s) {
      if (qsym && Av.qs(ia, (int)s) != -Bv.qs(ib, (int)s))
        throw Error(MPSXX_ERR_QN, "contraction of quantum numbers failed");  // btas_contract_qshape.h:87-107
      if (a.ds[ia][s] != b.ds[ib][s]) throw Error(MPSXX_ERR_SHAPE, "Contract: contracted dense shapes differ");
    }
  }

  ContractSym R;
  Layout& C = R.C;
  {
    std::vector<std::vector<QN>> qs;
    std::vector<std::vector<int>> ds;
    for (int i : ua) {
      std::vector<QN> v;
      for (size_t s = 0; s < a.ds[i].size(); ++s) v.push_back(Av.qs(i, (int)s));
      qs.push_back(v);
      ds.push_back(a.ds[i]);
    }
    for (int i : ub) {
      std::vector<QN> v;
      for (size_t s = 0; s < b.ds[i].size(); ++s) v.push_back(Bv.qs(i, (int)s));
      qs.push_back(v);
      ds.push_back(b.ds[i]);
    }
    C.set_shape((int)(ua.size() + ub.size()), qsym, Av.q() + Bv.q(), qs, ds);
  }
  bool use_mask = mask && mask->qsym && mask->rank == C.rank;
  for (int i = 0; use_mask && i < C.rank; ++i) use_mask = mask->nsec(i) == C.nsec(i);
  // whether the output block of tag t (row-major over the output's sector counts) is allowed by mask's quantum numbers
  auto mask_allows = [&](int64_t t) {
    int32_t idx[kMaxRank];
    for (int i = C.rank - 1; i >= 0; --i) {
      idx[i] = (int32_t)(t % C.nsec(i));
      t /= C.nsec(i);
    }
    return mask->allowed(idx);
  };
  if (existing && existing->rank > 0) {
    if (existing->rank != C.rank) throw Error(MPSXX_ERR_SHAPE, "output rank mismatch");
    if (qsym && existing->qsym) {
      if (existing->q != C.q) throw Error(MPSXX_ERR_QN, "Gemm(QSPARSE): quantum number of c must equal to a.q() + b.q().");
      if (!existing->same_qshape(C)) throw Error(MPSXX_ERR_QN, "Gemm(QSPARSE): c must have the same quantum shape of [ a * b ].");
    }
    if (!existing->same_dshape(C)) throw Error(MPSXX_ERR_SHAPE, "output dense shape mismatch");
  }

  // Every contributing (A block, B block) pair shares the contracted sector tuple ("key").  The output tag is
  // free_A * (sector count of B's free part) + free_B (row-major over C = A's free indices then B's), so walking the
  // A blocks grouped by their free tuple in ascending order, and each group's contributions sorted by (free_B, key),
  // yields the output blocks in ascending tag order and every block's pairs in ascending key order: the order of a
  // sort of all contributions by (output tag, key), without one.
  const int64_t nA = a.nblocks(), nB = b.nblocks();
  auto tag_over = [](const Layout& L, const std::vector<int>& ids, const int32_t* idx) {
    int64_t t = 0;
    for (int x : ids) t = t * L.nsec(x) + idx[x];
    return t;
  };
  // per-block strided 2-D views and quantum numbers of the free parts
  struct SideA { int64_t key, free; int64_t m, sm, k, sk; QN q; };
  struct SideB { int64_t key, free; int64_t k, sk, n, sn; QN q; };
  std::vector<SideA> A(nA);
  std::vector<SideB> Bs(nB);
  {
    // independent per block: contiguous block ranges on several host threads for large operands
    const int64_t nblk = nA + nB;
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)host_threads(), nblk / 16384));
    parallel_for(nth, [&](int t) {
      int64_t dims[kMaxRank], st[kMaxRank];
      for (int64_t i = nA * t / nth; i < nA * (t + 1) / nth; ++i) {
        const int32_t* ai = a.index(i);
        SideA& x = A[i];
        x.key = tag_over(a, ca, ai);
        x.free = tag_over(a, ua, ai);
        block_strides(a, ai, dims, st);
        if (!group_view(ua, dims, st, x.m, x.sm) || !group_view(ca, dims, st, x.k, x.sk))
          x.m = -1;  // only an error if the block takes part in a pair
        QN q;
        for (int u : ua) q = q + Av.qs(u, ai[u]);
        x.q = q;
      }
      for (int64_t j = nB * t / nth; j < nB * (t + 1) / nth; ++j) {
        const int32_t* bj = b.index(j);
        SideB& y = Bs[j];
        y.key = tag_over(b, cb, bj);
        y.free = tag_over(b, ub, bj);
        block_strides(b, bj, dims, st);
        if (!group_view(cb, dims, st, y.k, y.sk) || !group_view(ub, dims, st, y.n, y.sn)) y.n = -1;
        QN q;
        for (int u : ub) q = q + Bv.qs(u, bj[u]);
        y.q = q;
      }
    });
  }
  int64_t nfreeB = 1;
  for (int u : ub) nfreeB *= b.nsec(u);
  // B blocks by (key, quantum number of the free part), ascending free_B inside a list: an A block finds exactly
  // the B blocks whose pairing with it gives an allowed output block (STBLAS.h:331, c.allowed)
  struct KQ {
    int64_t key;
    int32_t p, s;
    bool operator==(const KQ& o) const { return key == o.key && p == o.p && s == o.s; }
  };
  // The lists are runs of one array (B blocks grouped by (key, QN), each group in ascending free_B: a stable
  // placement in free_B order), found through a flat open-addressing table of the (key, QN) groups.
  std::vector<int64_t> bord(nB);
  std::vector<int64_t> gstart;  // run of group g: bord[gstart[g], gstart[g + 1])
  size_t tcap = 16;
  while (tcap < 2 * (size_t)nB + 2) tcap <<= 1;
  struct Slot {
    KQ k;
    int32_t g;  // -1: empty
  };
  std::vector<Slot> table(tcap, Slot{KQ{0, 0, 0}, -1});
  auto slot_of = [&](const KQ& k) {
    uint64_t h = (uint64_t)k.key * 0x9e3779b97f4a7c15ull ^ ((uint64_t)(uint32_t)k.p << 32 | (uint32_t)k.s);
    h ^= h >> 31;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 29;
    size_t x = (size_t)h & (tcap - 1);
    while (table[x].g >= 0 && !(table[x].k == k)) x = (x + 1) & (tcap - 1);
    return x;
  };
  {
    std::vector<int32_t> bgrp(nB);
    std::vector<int64_t> cnt_g;
    for (int64_t j = 0; j < nB; ++j) {
      const QN q = qsym ? Bs[j].q : QN();
      const KQ k{Bs[j].key, q.p, q.s};
      const size_t x = slot_of(k);
      if (table[x].g < 0) {
        table[x] = Slot{k, (int32_t)cnt_g.size()};
        cnt_g.push_back(0);
      }
      bgrp[j] = table[x].g;
      ++cnt_g[bgrp[j]];
    }
    gstart.assign(cnt_g.size() + 1, 0);
    for (size_t g = 0; g < cnt_g.size(); ++g) gstart[g + 1] = gstart[g] + cnt_g[g];
    // B blocks in ascending free_B (counting sort when the range is moderate), placed into their groups in that order
    std::vector<int64_t> byfree(nB);
    if (nfreeB <= (int64_t(1) << 24)) {
      std::vector<int64_t> cf((size_t)nfreeB + 1, 0);
      for (int64_t j = 0; j < nB; ++j) ++cf[Bs[j].free + 1];
      for (int64_t f = 0; f < nfreeB; ++f) cf[f + 1] += cf[f];
      for (int64_t j = 0; j < nB; ++j) byfree[cf[Bs[j].free]++] = j;
    } else {
      std::iota(byfree.begin(), byfree.end(), 0);
      std::stable_sort(byfree.begin(), byfree.end(), [&](int64_t x, int64_t y) { return Bs[x].free < Bs[y].free; });
    }
    std::vector<int64_t> pos(gstart.begin(), gstart.end() - 1);
    for (int64_t j : byfree) bord[pos[bgrp[j]]++] = j;
  }
  struct Span {
    const int64_t* p;
    size_t n;
    const int64_t* begin() const { return p; }
    const int64_t* end() const { return p + n; }
    size_t size() const { return n; }
  };
  const QN cq = C.q;
  auto partners = [&](int64_t i) -> Span {
    const QN need = qsym ? cq - A[i].q : QN();
    const Slot& sl = table[slot_of(KQ{A[i].key, need.p, need.s})];
    if (sl.g < 0) return Span{nullptr, 0};
    return Span{bord.data() + gstart[sl.g], (size_t)(gstart[sl.g + 1] - gstart[sl.g])};
  };
  // A blocks by (free_A, key), then the boundaries of the free_A groups
  std::vector<int64_t> aord(nA);
  std::iota(aord.begin(), aord.end(), 0);
  auto a_before = [&](int64_t x, int64_t y) {
    return A[x].free != A[y].free ? A[x].free < A[y].free : A[x].key < A[y].key;
  };
  bool a_sorted = true;  // stored order already (free_A, key) when the contracted indices trail the free ones
  for (int64_t i = 1; i < nA && a_sorted; ++i) a_sorted = a_before(i - 1, i);
  if (!a_sorted) std::sort(aord.begin(), aord.end(), a_before);
  std::vector<size_t> gbeg;
  for (size_t r = 0; r < aord.size(); ++r)
    if (r == 0 || A[aord[r]].free != A[aord[r - 1]].free) gbeg.push_back(r);
  const size_t ngroups = gbeg.size();
  gbeg.push_back(aord.size());
  // Per chunk of consecutive groups: output tags (ascending), their first pair, and the pairs.  Chunks run on
  // several host threads when the contraction is large; their results are concatenated in order, so the plan is
  // the same for any thread count.
  struct Chunk {
    std::vector<int64_t> tag;
    std::vector<int64_t> first;  // chunk-local index of each output block's first pair
    std::vector<GemmPair> pairs;
    std::vector<int64_t> src;    // A block and B block of each output block's first pair
    int64_t flops = 0;
  };
  std::vector<int64_t> gwork(ngroups, 0);
  int64_t work = 0;
  for (size_t g = 0; g < ngroups; ++g) {
    for (size_t r = gbeg[g]; r < gbeg[g + 1]; ++r) {
      gwork[g] += (int64_t)partners(aord[r]).size();
    }
    work += gwork[g];
  }
  const int nthreads = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)host_threads(), work / g_grain.load(), (int64_t)ngroups}));
  std::vector<Chunk> chunks(nthreads);
  // nthreads ranges of consecutive groups with about equal candidate-pair counts
  std::vector<size_t> cut(nthreads + 1, ngroups);
  cut[0] = 0;
  {
    int64_t acc = 0;
    int c = 1;
    for (size_t g = 0; g < ngroups && c < nthreads; ++g) {
      acc += gwork[g];
      if (acc * nthreads >= work * c) cut[c++] = g + 1;
    }
  }
  // Order inside a group: the A blocks are walked in ascending key and each partner list is in ascending free_B, so a
  // stable counting sort by free_B gives (free_B, key) order; a comparison sort is the fallback for huge free_B ranges.
  const bool counting = nfreeB <= (int64_t(1) << 24);
  parallel_for(nthreads, [&](int c) {
    Chunk& out = chunks[c];
    struct Contrib { int64_t bfree, key; int32_t ia, ib; };
    std::vector<Contrib> loc, sorted;
    std::vector<int32_t> cnt(counting ? (size_t)nfreeB : 0, 0);
    std::vector<int64_t> seen;
    int64_t cand = 0;
    for (size_t g = cut[c]; g < cut[c + 1]; ++g) cand += gwork[g];
    out.pairs.reserve(cand);
    for (size_t g = cut[c]; g < cut[c + 1]; ++g) {
      const int64_t af = A[aord[gbeg[g]]].free;
      loc.clear();
      for (size_t r = gbeg[g]; r < gbeg[g + 1]; ++r) {
        const int64_t i = aord[r];
        for (int64_t j : partners(i)) {
          if (use_mask && !mask_allows(af * nfreeB + Bs[j].free)) continue;
          loc.push_back({Bs[j].free, A[i].key, (int32_t)i, (int32_t)j});
        }
      }
      if (counting) {
        seen.clear();
        for (const Contrib& x : loc)
          if (cnt[x.bfree]++ == 0) seen.push_back(x.bfree);
        std::sort(seen.begin(), seen.end());
        int32_t pos = 0;
        for (int64_t v : seen) {
          const int32_t n = cnt[v];
          cnt[v] = pos;
          pos += n;
        }
        sorted.resize(loc.size());
        for (const Contrib& x : loc) sorted[cnt[x.bfree]++] = x;
        for (int64_t v : seen) cnt[v] = 0;
        loc.swap(sorted);
      } else {
        std::sort(loc.begin(), loc.end(), [](const Contrib& x, const Contrib& y) {
          return x.bfree != y.bfree ? x.bfree < y.bfree : x.key < y.key;
        });
      }
      for (size_t r = 0; r < loc.size(); ++r) {
        const SideA& x = A[loc[r].ia];
        const SideB& y = Bs[loc[r].ib];
        if (r == 0 || loc[r].bfree != loc[r - 1].bfree) {
          out.tag.push_back(af * nfreeB + loc[r].bfree);
          out.first.push_back((int64_t)out.pairs.size());
          out.src.push_back(loc[r].ia);
          out.src.push_back(loc[r].ib);
        }
        if (x.m < 0 || y.n < 0)
          throw Error(MPSXX_ERR_UNSUPPORTED, "Contract: block permutation is not a strided 2-D view");
        if (x.k != y.k) throw Error(MPSXX_ERR_SHAPE, "Contract: contracted block dims differ");
        GemmPair gp;
        gp.a_off = a.off[loc[r].ia];
        gp.b_off = b.off[loc[r].ib];
        gp.sa_m = (int32_t)x.sm;
        gp.sa_k = (int32_t)x.sk;
        gp.sb_k = (int32_t)y.sk;
        gp.sb_n = (int32_t)y.sn;
        gp.k = (int32_t)x.k;
        gp.pad = 0;
        // a unit group carries stride 0; the kernel only ever touches index 0 of it, and the stride choice merely
        // selects which global->LDS thread mapping is used
        out.pairs.push_back(gp);
        out.flops += 2 * x.m * y.n * x.k;
      }
    }
  });

  // output structure: existing blocks U contributed blocks
  if (existing && existing->rank > 0) {
    for (int64_t k = 0; k < existing->nblocks(); ++k) C.add_block(existing->index(k));
  }
  std::vector<int64_t> out_base(nthreads + 1, 0), pair_base(nthreads + 1, 0);
  for (int c = 0; c < nthreads; ++c) {
    out_base[c + 1] = out_base[c] + (int64_t)chunks[c].tag.size();
    pair_base[c + 1] = pair_base[c] + (int64_t)chunks[c].pairs.size();
    R.flops += chunks[c].flops;
  }
  const int64_t nout = out_base[nthreads], npairs = pair_base[nthreads];
  if (npairs > INT32_MAX) throw Error(MPSXX_ERR_UNSUPPORTED, "Contract: more than 2^31 block pairs");
  {
    std::vector<int32_t> cidx(C.rank);
    for (int c = 0; c < nthreads; ++c)
      for (size_t t = 0; t < chunks[c].tag.size(); ++t) {
        const int32_t* ai = a.index(chunks[c].src[2 * t]);
        const int32_t* bj = b.index(chunks[c].src[2 * t + 1]);
        int n = 0;
        for (int y : ua) cidx[n++] = ai[y];
        for (int y : ub) cidx[n++] = bj[y];
        C.add_block(cidx.data());
      }
  }
  C.finalize();

  // tasks; without existing blocks the output blocks are exactly the contributed ones, in the same order
  R.tasks.resize(nout);
  R.pairs.resize(npairs);
  const bool only_new = !(existing && existing->rank > 0);
  parallel_for(nthreads, [&](int c) {
    Chunk& ch = chunks[c];
    std::copy(ch.pairs.begin(), ch.pairs.end(), R.pairs.begin() + pair_base[c]);
    std::vector<GemmPair>().swap(ch.pairs);
    for (size_t t = 0; t < ch.tag.size(); ++t) {
      const int64_t o = out_base[c] + (int64_t)t;
      const int64_t cb_ = only_new ? o : C.find(ch.tag[t]);
      const int64_t lo = ch.first[t];
      const int64_t hi = t + 1 < ch.tag.size() ? ch.first[t + 1] : pair_base[c + 1] - pair_base[c];
      GemmTask& task = R.tasks[o];
      task.c_off = C.off[cb_];
      task.pair_begin = (int32_t)(pair_base[c] + lo);
      task.pair_count = (int32_t)(hi - lo);
      task.inst = 0;
      task.m = (int32_t)A[ch.src[2 * t]].m;
      task.n = (int32_t)Bs[ch.src[2 * t + 1]].n;
      task.layout = kTaskContigA | kTaskContigB;
      for (int32_t q = task.pair_begin; q < task.pair_begin + task.pair_count; ++q) {
        const GemmPair& g = R.pairs[q];
        if (g.sa_k != 1) task.layout &= ~kTaskContigA;
        if (g.sb_k != 1) task.layout &= ~kTaskContigB;
      }
    }
  });
  return R;
}

Layout tie_layout(const Layout& X, const std::vector<int>& tie, std::vector<CopySeg>& segs) {
  const int N = X.rank, K = (int)tie.size();
  if (K < 2) throw Error(MPSXX_ERR_ARG, "Tie: need at least two indices");
  std::vector<char> tied(N, 0);
  for (int i : tie) {
    if (i < 0 || i >= N) throw Error(MPSXX_ERR_ARG, "Tie: index out of range");
    tied[i] = 1;
  }
  std::vector<int> uniq;
  for (int i = 0; i < N; ++i) {
    if (!tied[i] || i == tie[0]) uniq.push_back(i);
    else if (X.ds[i].size() != X.ds[tie[0]].size()) throw Error(MPSXX_ERR_SHAPE, "Tie(DENSE): index to be tied must be the same.");
  }
  Layout Y;
  {
    std::vector<std::vector<int>> ds;
    for (int i : uniq) ds.push_back(X.ds[i]);
    Y.set_shape((int)uniq.size(), false, QN(), {}, ds);
  }
  std::vector<int64_t> src;
  std::vector<int32_t> yi(Y.rank);
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    bool ok = true;
    for (int k = 1; k < K; ++k) ok &= (xi[tie[k]] == xi[tie[0]]);
    if (!ok) continue;
    for (int n = 0; n < Y.rank; ++n) yi[n] = xi[uniq[n]];
    Y.add_block(yi.data());
  }
  Y.finalize();
  segs.clear();
  int64_t dims[kMaxRank], st[kMaxRank];
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    bool ok = true;
    for (int k = 1; k < K; ++k) ok &= (xi[tie[k]] == xi[tie[0]]);
    if (!ok) continue;
    for (int n = 0; n < Y.rank; ++n) yi[n] = xi[uniq[n]];
    const int64_t yb = Y.find(Y.compute_tag(yi.data()));
    block_strides(X, xi, dims, st);
    for (int k = 1; k < K; ++k)
      if (dims[tie[k]] != dims[tie[0]]) throw Error(MPSXX_ERR_SHAPE, "Tie: tied dense dims differ");
    // output dims / source strides
    std::vector<int64_t> od, os;
    for (int n = 0; n < Y.rank; ++n) {
      const int i = uniq[n];
      od.push_back(dims[i]);
      int64_t s = st[i];
      if (i == tie[0]) {
        s = 0;
        for (int k = 0; k < K; ++k) s += st[tie[k]];
      }
      os.push_back(s);
    }
    // the innermost dimension of extent > 1 forms the segment (trailing unit extents leave the element order
    // unchanged: a diagonal of an (R, 1, R) block is one segment of R elements, not R of one), outer dims enumerate
    // segments
    int inner = Y.rank - 1;
    while (inner > 0 && od[inner] == 1) --inner;
    int64_t outer = 1;
    for (int n = 0; n < inner; ++n) outer *= od[n];
    if (od[inner] == 0) continue;
    for (int64_t o = 0; o < outer; ++o) {
      // o enumerates the outer output dims row-major; the destination of segment o starts at o * od[inner]
      int64_t rem = o, soff = 0;
      for (int n = inner - 1; n >= 0; --n) {
        const int64_t c = rem % od[n];
        rem /= od[n];
        soff += c * os[n];
      }
      segs.push_back({Y.off[yb] + o * od[inner], X.off[b] + soff, (int32_t)od[inner], (int32_t)os[inner]});
    }
  }
  return Y;
}

// ---------------------------------------------------------------------------------------------------------------------
// Steps

Step make_contract_step(const ContractSym& sym, int a, int b, int c, int accumulate, double alpha,
                        const GemmPolicy& pol) {
  Step st;
  st.a = a;
  st.b = b;
  st.c = c;
  st.accumulate = accumulate;
  st.alpha = alpha;
  st.flops = sym.flops;
  bool scalarB = true, scalarA = true;
  for (const GemmTask& t : sym.tasks) {
    for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p) {
      const GemmPair& g = sym.pairs[p];
      if (!(g.k == 1 && t.n == 1)) scalarB = false;
      if (!(g.k == 1 && t.m == 1)) scalarA = false;
    }
  }
  if (scalarB || scalarA) {
    // C[r] += S * V[r*stride]: the vector side is the one whose free extent is the block
    st.kind = Step::AXPY;
    if (!scalarB) std::swap(st.a, st.b);  // V must be slot a
    std::vector<AxpyTask> tasks;
    std::vector<AxpyPair> pairs;
    std::vector<Chunk> chunks;
    for (size_t ti = 0; ti < sym.tasks.size(); ++ti) {
      const GemmTask& t = sym.tasks[ti];
      AxpyTask at;
      at.c_off = t.c_off;
      at.len = scalarB ? t.m : t.n;
      at.pair_begin = (int32_t)pairs.size();
      at.pair_count = t.pair_count;
      at.inst = 0;
      for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p) {
        const GemmPair& g = sym.pairs[p];
        AxpyPair ap;
        if (scalarB) {
          ap.v_off = g.a_off;
          ap.v_stride = g.sa_m;
          ap.s_off = g.b_off;
        } else {
          ap.v_off = g.b_off;
          ap.v_stride = g.sb_n;
          ap.s_off = g.a_off;
        }
        ap.s2_off = -1;
        ap.pad = 0;
        pairs.push_back(ap);
      }
      for (int s = 0; s < at.len; s += kAxpyChunk) chunks.push_back({(int32_t)ti, s});
      tasks.push_back(at);
    }
    st.nwork = (int)chunks.size();
    st.work = upload_vec(chunks);
    st.tasks = upload_vec(tasks);
    st.pairs = upload_vec(pairs);
    st.host.atasks = std::move(tasks);
    st.host.apairs = std::move(pairs);
    st.host.chunks = std::move(chunks);
    return st;
  }
  st.kind = Step::GEMM;
  st.tasks = upload_vec(sym.tasks);
  st.pairs = upload_vec(sym.pairs);
  st.host.gtasks = sym.tasks;
  st.host.gpairs = sym.pairs;
  set_gemm_policy(st, pol);
  return st;
}

namespace {
DevBuf clone_buf(const DevBuf& b) {
  DevBuf c;
  if (!b.p) return c;
  c.alloc(b.bytes);
  hip_check(hipMemcpy(c.p, b.p, b.bytes, hipMemcpyDeviceToDevice), "clone step lists");
  return c;
}
}  // namespace

Step clone_step(const Step& st) {
  Step t;
  t.kind = st.kind;
  t.a = st.a;
  t.b = st.b;
  t.c = st.c;
  t.b2 = st.b2;
  t.accumulate = st.accumulate;
  t.alpha = st.alpha;
  t.beta = st.beta;
  t.nwork = st.nwork;
  t.work = clone_buf(st.work);
  t.tasks = clone_buf(st.tasks);
  t.pairs = clone_buf(st.pairs);
  t.nwork_red = st.nwork_red;
  t.work_red = clone_buf(st.work_red);
  t.partial = clone_buf(st.partial);  // scratch: only its size matters
  t.nwork_small = st.nwork_small;
  t.work_small = clone_buf(st.work_small);
  t.nwork_split = st.nwork_split;
  t.work_split = clone_buf(st.work_split);
  t.nwork_split_red = st.nwork_split_red;
  t.work_split_red = clone_buf(st.work_split_red);
  t.split_partial = clone_buf(st.split_partial);
  t.nwork2 = st.nwork2;
  for (int i = 0; i < 4; ++i) {
    t.nwork2_layout[i] = st.nwork2_layout[i];
    t.nwork_layout[i] = st.nwork_layout[i];
  }
  t.work2 = clone_buf(st.work2);
  t.nwork_red2 = st.nwork_red2;
  t.work_red2 = clone_buf(st.work_red2);
  t.partial2 = clone_buf(st.partial2);
  t.max_groups = st.max_groups;
  t.policy = st.policy;
  t.flops = st.flops;
  t.host = st.host;
  return t;
}

GemmPolicy default_policy() {
  // measured on the L=32 M=512 Hubbard sigma: single-pair blocks run best one wave per tile, multi-pair blocks best
  // split from 2 units on
  static const GemmPolicy p = [] {
    GemmPolicy q;
    if (const char* e = std::getenv("MPSXX_SMALL_GEMM_LIMIT")) q.small_limit = std::atoi(e);
    if (const char* e = std::getenv("MPSXX_SPLIT_UNITS")) q.split_units = std::atoi(e);
    if (const char* e = std::getenv("MPSXX_SPLIT_BIG")) q.split_big = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("MPSXX_SPLIT_SMALL")) q.split_small = std::atoi(e);
    if (const char* e = std::getenv("MPSXX_SPLIT_SEG")) q.split_seg = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("MPSXX_BIG_TILE_MIN")) q.big_tile_min = std::max(0, std::atoi(e));
    return q;
  }();
  return p;
}

GemmPolicy policy_for(GemmPolicy p, int precision) {
  (void)precision;  // every precision has 128x128 tiles
  return p;
}

GemmPolicy default_policy(int precision) {
  // measured on the norb=24 M=1024 merged sigma: 128x128 tiles take FP32 from 22.0 to 19.5 ms; FP64 with two
  // work-groups per CU (128 accumulator registers per lane) from 25.5 to 24.3 ms against 64x64 tiles throughout (with
  // one work-group per CU it had been slower).  MPSXX_BIG_TILE_MIN_F64 = 0 keeps FP64 on 64x64 tiles.
  // The threshold per precision, measured on that sigma and on the same bond of the converged M=1024 sweep state
  // (smaller sectors; `snapshot`), FP64 / FP32 / split-BF16x3 in ms:
  //   threshold    synthetic              converged state
  //   FP64  96 / 192     23.98 / 24.03          9.05 / 7.26   (64x64 throughout: 25.48 / 7.12)
  //   FP32  64 / 192     11.54 / 11.80          4.71 / 4.09
  //   BF16x3 64 / 192     6.98 / 7.31           3.22 / 3.35
  // so 192 for FP64 and FP32 (the QC sweeps: FP64 11.7 -> 11.0 s) and 64 for split-BF16x3.  MPSXX_BIG_TILE_MIN sets
  // the FP32 and split-BF16x3 thresholds, MPSXX_BIG_TILE_MIN_F64 the FP64 one.
  static const int f64_min = [] {
    const char* e = std::getenv("MPSXX_BIG_TILE_MIN_F64");
    return e ? std::max(0, std::atoi(e)) : GemmPolicy().big_tile_min;
  }();
  static const int x3_min = [] {
    const char* e = std::getenv("MPSXX_BIG_TILE_MIN");
    return e ? std::max(0, std::atoi(e)) : kBigTileMinX3;
  }();
  GemmPolicy p = default_policy();
  if (precision == MPSXX_FP64) p.big_tile_min = f64_min;
  if (precision == MPSXX_BF16X3) p.big_tile_min = x3_min;
  return policy_for(p, precision);
}

void gemm_tiles(const std::vector<GemmTask>& tasks, const std::vector<GemmPair>& pairs, const GemmPolicy& pol,
                GemmTiling& out) {
  out = GemmTiling();
  // heavy tasks first so the tail of each launch is made of small tiles
  std::vector<size_t> order(tasks.size());
  std::iota(order.begin(), order.end(), 0);
  std::vector<int64_t> cost(tasks.size(), 0);
  auto units_of = [&](int lo, int hi) {
    int64_t u = 0;
    for (int p = lo; p < hi; ++p) u += (pairs[p].k + 15) / 16;
    return u;
  };
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    const GemmTask& t = tasks[ti];
    int64_t ksum = 0;
    for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p) ksum += pairs[p].k;
    cost[ti] = ksum;
  }
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  // tile edge of a large block: 128 when both sides reach big_tile_min and every in-block offset of its pairs, in
  // elements and in bytes, fits the 32-bit offsets of gemm128_kernel and the buffer loads of gemm_x3_128_kernel
  // (below 2^28 elements), else 64
  auto edge_of = [&](const GemmTask& t) {
    if (pol.big_tile_min <= 0 || t.m < pol.big_tile_min || t.n < pol.big_tile_min) return kGemmTile;
    constexpr int64_t kMaxOff = int64_t(1) << 28;
    for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p) {
      const GemmPair& g = pairs[p];
      const int64_t pad_m = t.m + kGemmTile2, pad_n = t.n + kGemmTile2, pad_k = g.k + 32;
      if (g.sa_m < 0 || g.sa_k < 0 || g.sb_k < 0 || g.sb_n < 0) return kGemmTile;
      if (pad_m * g.sa_m + pad_k * g.sa_k >= kMaxOff || pad_k * g.sb_k + pad_n * g.sb_n >= kMaxOff) return kGemmTile;
    }
    return kGemmTile2;
  };
  std::vector<size_t> bigtasks, splittasks;
  int64_t nbig = 0, ubig = 0, usplit = 0;
  for (size_t ti : order) {
    const GemmTask& t = tasks[ti];
    const int64_t units = units_of(t.pair_begin, t.pair_begin + t.pair_count);
    if (t.m <= pol.small_limit && t.n <= pol.small_limit) {
      if (t.pair_count >= 2 && units >= pol.split_units) {
        splittasks.push_back(ti);
        usplit += (int64_t)((t.m + kSmallTile - 1) / kSmallTile) * ((t.n + kSmallTile - 1) / kSmallTile) * units;
        continue;
      }
      for (int r = 0; r < t.m; r += kSmallTile)
        for (int c = 0; c < t.n; c += kSmallTile) out.small.push_back({(int32_t)ti, r, c, 0});
      continue;
    }
    bigtasks.push_back(ti);
    // tile count in 64x64 equivalents (a 128x128 tile carries four)
    const int E = edge_of(t);
    const int64_t nt = (int64_t)((t.m + E - 1) / E) * ((t.n + E - 1) / E) * (E / kGemmTile) * (E / kGemmTile);
    nbig += nt;
    ubig += nt * units;
  }
  // boundaries of nseg pair ranges of [lo, hi) with about units/nseg units each
  auto cuts = [&](int lo, int hi, int64_t units, int nseg) {
    std::vector<int> cut{lo};
    if (nseg > 1) {
      int64_t acc = 0;
      int made = 1;
      for (int p = lo; p < hi && made < nseg; ++p) {
        acc += (pairs[p].k + 15) / 16;
        if (acc * nseg >= units * made && p + 1 < hi) {
          cut.push_back(p + 1);
          ++made;
        }
      }
    }
    cut.push_back(hi);
    return cut;
  };
  // split the pair lists of long reductions when the big tiles alone cannot fill the device
  const bool split = pol.split_big > 0 && nbig > 0 && nbig < pol.split_big;
  const int64_t target = split ? std::max<int64_t>(8, (ubig + pol.split_big - 1) / pol.split_big) : 0;
  for (size_t ti : bigtasks) {
    const GemmTask& t = tasks[ti];
    const int lo = t.pair_begin, hi = t.pair_begin + t.pair_count;
    const int64_t units = units_of(lo, hi);
    int nseg = 1;
    if (split)
      nseg = (int)std::min<int64_t>({64, (int64_t)t.pair_count, std::max<int64_t>(1, (units + target - 1) / target)});
    const std::vector<int> cut = cuts(lo, hi, units, nseg);
    const int nparts = (int)cut.size() - 1;
    const int E = edge_of(t);
    std::vector<BigTile>& big = E == kGemmTile ? out.big : out.big2;
    std::vector<RedTile>& red = E == kGemmTile ? out.red : out.red2;
    int& nslots = E == kGemmTile ? out.nslots : out.nslots2;
    if (nparts == 1) {
      // the column tiles of a row band follow each other: they read the same A panels
      for (int r = 0; r < t.m; r += E)
        for (int c = 0; c < t.n; c += E) big.push_back({(int32_t)ti, r, c, lo, hi, -1});
      continue;
    }
    // split reduction: slots of tile (r, c) are contiguous (summed in order by the reduce kernel); the tiles are
    // listed row band by row band and pair range by pair range, so that consecutive tiles share their A panels
    const int ncol = (t.n + E - 1) / E;
    const int slot_base = nslots;
    for (int r = 0; r < t.m; r += E)
      for (int c = 0; c < t.n; c += E) {
        red.push_back({(int32_t)ti, r, c, nslots, nparts, 0});
        nslots += nparts;
      }
    for (int r = 0, rb = 0; r < t.m; r += E, ++rb)
      for (int q = 0; q < nparts; ++q)
        for (int c = 0, cb = 0; c < t.n; c += E, ++cb)
          big.push_back({(int32_t)ti, r, c, cut[q], cut[q + 1], slot_base + (rb * ncol + cb) * nparts + q});
  }
  // the split 32x32 tiles: a tile whose reduction is much longer than the list's share per work-group
  // (split_small work-groups in all) is cut into pair ranges of >= split_seg units, one work-group each
  const int64_t starget =
      pol.split_small > 0 ? std::max<int64_t>(pol.split_seg, (usplit + pol.split_small - 1) / pol.split_small) : 0;
  for (size_t ti : splittasks) {
    const GemmTask& t = tasks[ti];
    const int lo = t.pair_begin, hi = t.pair_begin + t.pair_count;
    const int64_t units = units_of(lo, hi);
    int nseg = 1;
    if (starget > 0) nseg = (int)std::min<int64_t>({64, (int64_t)t.pair_count, std::max<int64_t>(1, units / starget)});
    const std::vector<int> cut = cuts(lo, hi, units, nseg);
    const int nparts = (int)cut.size() - 1;
    for (int r = 0; r < t.m; r += kSmallTile)
      for (int c = 0; c < t.n; c += kSmallTile) {
        if (nparts == 1) {
          out.split.push_back({(int32_t)ti, r, c, lo, hi, -1});
          continue;
        }
        out.split_red.push_back({(int32_t)ti, r, c, out.split_slots, nparts, 0});
        for (int q = 0; q < nparts; ++q)
          out.split.push_back({(int32_t)ti, r, c, cut[q], cut[q + 1], out.split_slots + q});
        out.split_slots += nparts;
      }
  }
}

void set_gemm_policy(Step& st, const GemmPolicy& pol) {
  GemmTiling g;
  gemm_tiles(st.host.gtasks, st.host.gpairs, pol, g);
  st.policy = pol;
  st.max_groups = pol.max_groups;
  // the 64x64 and 128x128 tiles grouped by their task's operand layouts (one kernel per layout), order kept within
  auto group_by_layout = [&](std::vector<BigTile>& tl_list, int (&counts)[4]) {
    const int32_t order[4] = {kTaskContigA | kTaskContigB, kTaskContigA, kTaskContigB, 0};
    std::vector<BigTile> grouped;
    grouped.reserve(tl_list.size());
    for (int q = 0; q < 4; ++q) {
      const size_t before = grouped.size();
      for (const BigTile& tl : tl_list)
        if ((st.host.gtasks[tl.task].layout & (kTaskContigA | kTaskContigB)) == order[q]) grouped.push_back(tl);
      counts[q] = (int)(grouped.size() - before);
    }
    tl_list.swap(grouped);
  };
  group_by_layout(g.big, st.nwork_layout);
  group_by_layout(g.big2, st.nwork2_layout);
  st.nwork = (int)g.big.size();
  st.work = upload_vec(g.big);
  st.nwork_red = (int)g.red.size();
  st.work_red = upload_vec(g.red);
  st.partial.alloc((size_t)g.nslots * kGemmTile * kGemmTile * sizeof(double));
  st.nwork_small = (int)g.small.size();
  st.work_small = upload_vec(g.small);
  st.nwork_split = (int)g.split.size();
  st.work_split = upload_vec(g.split);
  st.nwork_split_red = (int)g.split_red.size();
  st.work_split_red = upload_vec(g.split_red);
  st.nwork2 = (int)g.big2.size();
  st.work2 = upload_vec(g.big2);
  st.nwork_red2 = (int)g.red2.size();
  st.work_red2 = upload_vec(g.red2);
  st.partial2.alloc((size_t)g.nslots2 * kGemmTile2 * kGemmTile2 * sizeof(double));
  st.split_partial.alloc((size_t)g.split_slots * kSmallTile * kSmallTile * sizeof(double));
}

bool make_fused_axpy_step(const ContractSym& first, const ContractSym& second, int y, int w1, int w2, int z,
                          int accumulate, Step& out) {
  auto scalar_b = [](const ContractSym& s) {
    for (const GemmTask& t : s.tasks) {
      if (t.n != 1) return false;
      for (int p = t.pair_begin; p < t.pair_begin + t.pair_count; ++p)
        if (s.pairs[p].k != 1) return false;
    }
    return true;
  };
  if (!scalar_b(first) || !scalar_b(second)) return false;
  std::unordered_map<int64_t, size_t> xtask;  // X block offset -> task of `first` that writes it
  for (size_t i = 0; i < first.tasks.size(); ++i) xtask[first.tasks[i].c_off] = i;
  std::vector<AxpyTask> tasks;
  std::vector<AxpyPair> pairs;
  std::vector<Chunk> chunks;
  for (size_t ti = 0; ti < second.tasks.size(); ++ti) {
    const GemmTask& t = second.tasks[ti];
    AxpyTask at;
    at.c_off = t.c_off;
    at.len = t.m;
    at.pair_begin = (int32_t)pairs.size();
    at.inst = 0;
    for (int q = t.pair_begin; q < t.pair_begin + t.pair_count; ++q) {
      const GemmPair& g2 = second.pairs[q];
      auto it = xtask.find(g2.a_off);
      if (it == xtask.end()) return false;  // X read from inside a block: not a plain chain
      const GemmTask& t1 = first.tasks[it->second];
      for (int p = t1.pair_begin; p < t1.pair_begin + t1.pair_count; ++p) {
        const GemmPair& g1 = first.pairs[p];
        AxpyPair ap;
        ap.v_off = g1.a_off;
        ap.v_stride = g1.sa_m * g2.sa_m;
        ap.s_off = g1.b_off;
        ap.s2_off = g2.b_off;
        ap.pad = 0;
        pairs.push_back(ap);
      }
    }
    at.pair_count = (int32_t)pairs.size() - at.pair_begin;
    for (int s = 0; s < at.len; s += kAxpyChunk) chunks.push_back({(int32_t)ti, s});
    tasks.push_back(at);
  }
  out = Step();
  out.kind = Step::AXPY;
  out.a = y;
  out.b = w1;
  out.b2 = w2;
  out.c = z;
  out.accumulate = accumulate;
  out.alpha = 1.0;
  out.flops = first.flops + second.flops;
  out.nwork = (int)chunks.size();
  out.work = upload_vec(chunks);
  out.tasks = upload_vec(tasks);
  out.pairs = upload_vec(pairs);
  out.host.atasks = std::move(tasks);
  out.host.apairs = std::move(pairs);
  out.host.chunks = std::move(chunks);
  return true;
}

Step make_copy_step(const std::vector<CopySeg>& segs, int x, int y, double alpha, double beta) {
  Step st;
  st.kind = Step::COPY;
  st.a = x;
  st.c = y;
  st.alpha = alpha;
  st.beta = beta;
  std::vector<Chunk> chunks;
  for (size_t i = 0; i < segs.size(); ++i)
    for (int s = 0; s < segs[i].len; s += kAxpyChunk) chunks.push_back({(int32_t)i, s});
  st.nwork = (int)chunks.size();
  st.work = upload_vec(chunks);
  st.tasks = upload_vec(segs);
  return st;
}

namespace {
template <typename T, typename TC>
void launch_step_t(const Step& st, const Operands& ops, unsigned parts, bool x3, hipStream_t s) {
  if (st.kind == Step::GEMM) {
    if ((parts & 1u) && x3)
      launch_gemm_x3<TC>(st.work.as<BigTile>(), st.nwork_layout, st.work_red.as<RedTile>(), st.nwork_red, st.partial.p,
                         st.tasks.as<GemmTask>(), st.pairs.as<GemmPair>(), ops, st.alpha, st.accumulate, s,
                         st.max_groups);
    else if (parts & 1u)
      launch_gemm<T, TC>(st.work.as<BigTile>(), st.nwork_layout, st.work_red.as<RedTile>(), st.nwork_red, st.partial.p,
                         st.tasks.as<GemmTask>(), st.pairs.as<GemmPair>(), ops, st.alpha, st.accumulate, s,
                         st.max_groups);
    if (parts & 2u)
      launch_gemm_small<T, TC>(st.work_small.as<GemmTile>(), st.nwork_small, st.tasks.as<GemmTask>(),
                               st.pairs.as<GemmPair>(), ops, st.alpha, st.accumulate, s, st.max_groups);
    if (parts & 4u)
      launch_gemm_split<T, TC>(st.work_split.as<BigTile>(), st.nwork_split, st.work_split_red.as<RedTile>(),
                               st.nwork_split_red, st.split_partial.p, st.tasks.as<GemmTask>(),
                               st.pairs.as<GemmPair>(), ops, st.alpha, st.accumulate, s, st.max_groups);
    if ((parts & 8u) && x3)
      launch_gemm_x3_128<TC>(st.work2.as<BigTile>(), st.nwork2_layout, st.work_red2.as<RedTile>(), st.nwork_red2,
                             st.partial2.p, st.tasks.as<GemmTask>(), st.pairs.as<GemmPair>(), ops, st.alpha,
                             st.accumulate, s, st.max_groups);
    else if (parts & 8u)
      launch_gemm128<T, TC>(st.work2.as<BigTile>(), st.nwork2_layout, st.work_red2.as<RedTile>(), st.nwork_red2,
                            st.partial2.p, st.tasks.as<GemmTask>(), st.pairs.as<GemmPair>(), ops, st.alpha,
                            st.accumulate, s, st.max_groups);
  } else if (parts & 1u) {
    launch_axpy<T, TC>(st.work.as<Chunk>(), st.nwork, st.tasks.as<AxpyTask>(), st.pairs.as<AxpyPair>(), ops,
                       st.alpha, st.accumulate, s, st.max_groups);
  }
}
}  // namespace

void launch_step(const Step& st, const Operands& ops, bool f32, bool cf32, hipStream_t s, unsigned parts, bool x3) {
  if (!f32) launch_step_t<double, double>(st, ops, parts, false, s);
  else if (cf32) launch_step_t<float, float>(st, ops, parts, x3, s);
  else launch_step_t<float, double>(st, ops, parts, x3, s);
}

std::vector<KernelPart> step_parts(const Step& st, int ea, int ec) {
  std::vector<KernelPart> out;
  if (st.kind == Step::AXPY) {
    int64_t fl = 0, by = 0;
    for (const AxpyTask& t : st.host.atasks) {
      fl += 2ll * t.len * t.pair_count;
      by += (int64_t)ea * t.len * t.pair_count + (int64_t)ec * t.len;
    }
    if (st.nwork > 0) out.push_back({KernelPart::AXPY, 1u, st.nwork, fl, by});
    return out;
  }
  if (st.kind != Step::GEMM) return out;
  // exact split of the step's flops over the three tile lists: a tile covers rows x cols of its block
  auto tile_cost = [&](int task, int row0, int col0, int lo, int hi, int edge, int64_t& fl, int64_t& by) {
    const GemmTask& t = st.host.gtasks[task];
    const int64_t r = std::min(edge, t.m - row0), c = std::min(edge, t.n - col0);
    int64_t ks = 0;
    for (int p = lo; p < hi; ++p) ks += st.host.gpairs[p].k;
    fl += 2 * r * c * ks;
    by += ea * ((r + c) * ks);
  };
  GemmTiling g;
  gemm_tiles(st.host.gtasks, st.host.gpairs, st.policy, g);
  int64_t fl = 0, by = 0;
  for (const BigTile& tl : g.big) tile_cost(tl.task, tl.row0, tl.col0, tl.pair_lo, tl.pair_hi, kGemmTile, fl, by);
  for (const BigTile& tl : g.big) {
    if (tl.slot >= 0) continue;
    const GemmTask& t = st.host.gtasks[tl.task];
    by += ec * (int64_t)std::min(kGemmTile, t.m - tl.row0) * std::min(kGemmTile, t.n - tl.col0);
  }
  for (const RedTile& rt : g.red) {
    const GemmTask& t = st.host.gtasks[rt.task];
    by += ec * (int64_t)std::min(kGemmTile, t.m - rt.row0) * std::min(kGemmTile, t.n - rt.col0);
  }
  if (!g.big.empty()) out.push_back({KernelPart::GEMM64, 1u, (int)g.big.size(), fl, by});
  fl = by = 0;
  for (const GemmTile& tl : g.small) {
    const GemmTask& t = st.host.gtasks[tl.task];
    tile_cost(tl.task, tl.row0, tl.col0, t.pair_begin, t.pair_begin + t.pair_count, kSmallTile, fl, by);
    by += ec * (int64_t)std::min(kSmallTile, t.m - tl.row0) * std::min(kSmallTile, t.n - tl.col0);
  }
  if (!g.small.empty()) out.push_back({KernelPart::GEMM_SMALL, 2u, (int)g.small.size(), fl, by});
  fl = by = 0;
  for (const BigTile& tl : g.split) {
    const GemmTask& t = st.host.gtasks[tl.task];
    tile_cost(tl.task, tl.row0, tl.col0, tl.pair_lo, tl.pair_hi, kSmallTile, fl, by);
    if (tl.slot < 0) by += ec * (int64_t)std::min(kSmallTile, t.m - tl.row0) * std::min(kSmallTile, t.n - tl.col0);
  }
  for (const RedTile& rt : g.split_red) {
    const GemmTask& t = st.host.gtasks[rt.task];
    by += ec * (int64_t)std::min(kSmallTile, t.m - rt.row0) * std::min(kSmallTile, t.n - rt.col0);
  }
  if (!g.split.empty()) out.push_back({KernelPart::GEMM_SPLIT, 4u, (int)g.split.size(), fl, by});
  fl = by = 0;
  for (const BigTile& tl : g.big2) {
    tile_cost(tl.task, tl.row0, tl.col0, tl.pair_lo, tl.pair_hi, kGemmTile2, fl, by);
    if (tl.slot < 0) {
      const GemmTask& t = st.host.gtasks[tl.task];
      by += ec * (int64_t)std::min(kGemmTile2, t.m - tl.row0) * std::min(kGemmTile2, t.n - tl.col0);
    }
  }
  for (const RedTile& rt : g.red2) {
    const GemmTask& t = st.host.gtasks[rt.task];
    by += ec * (int64_t)std::min(kGemmTile2, t.m - rt.row0) * std::min(kGemmTile2, t.n - rt.col0);
  }
  if (!g.big2.empty()) out.push_back({KernelPart::GEMM128, 8u, (int)g.big2.size(), fl, by});
  return out;
}

const char* part_kernel_name(int kind) {
  switch (kind) {
    case KernelPart::GEMM64: return "gemm_grouped_kernel";
    case KernelPart::GEMM_X3: return "gemm_x3_kernel";
    case KernelPart::GEMM_SMALL: return "gemm_small_kernel";
    case KernelPart::GEMM_SPLIT: return "gemm_split_kernel";
    case KernelPart::AXPY: return "axpy_grouped_kernel";
    case KernelPart::GEMM128: return "gemm128_kernel";
    case KernelPart::GEMM_X3_128: return "gemm_x3_128_kernel";
    default: return "copy_kernel";
  }
}

void run_plan(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s) {
  run_plan_timed(P, inputs, out, s, {});
}

void run_plan_timed(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                    const std::vector<hipEvent_t>& events) {
  run_plan_ex(P, inputs, out, s, events, ~0u);
}

void Plan::strip() {
  for (DevBuf& b : slot_buf) b.release();
  conv_src.clear();
  stripped = true;
}

bool Plan::restore() {
  if (!stripped) return false;
  bool any = false;
  for (size_t i = 0; i < slot_layout.size(); ++i) {
    const bool inter = slot_input[i] == -1, f32_copy = slot_input[i] >= 0 && slot_prec[i] == 1;
    if (!(inter || f32_copy) || slot_buf[i].p) continue;
    slot_buf[i].alloc(std::max<int64_t>(slot_layout[i].nelem, 1) * (slot_prec[i] ? sizeof(float) : sizeof(double)));
    any = true;
  }
  stripped = false;
  return any;
}

std::vector<void*> slot_pointers(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                                 unsigned convert_mask) {
  P.restore();  // a restored plan's FP32 operand copies hold nothing (conv_src was cleared by strip)
  const int ns = (int)P.slot_layout.size();
  if ((int)P.conv_src.size() != ns) P.conv_src.assign(ns, nullptr);
  std::vector<void*> ptr(ns, nullptr);
  for (int i = 0; i < ns; ++i) {
    if (P.slot_input[i] >= 0) {
      const double* src = inputs[P.slot_input[i]];
      if (P.slot_prec[i] == 1) {
        if ((convert_mask & (1u << P.slot_input[i])) || P.conv_src[i] != src) {
          launch_convert<double, float>(src, P.slot_buf[i].as<float>(), P.slot_layout[i].nelem, s);
          P.conv_src[i] = src;
        }
        ptr[i] = P.slot_buf[i].p;
      } else {
        ptr[i] = const_cast<double*>(src);
      }
    } else if (P.slot_input[i] == -2) {
      ptr[i] = out;
    } else {
      ptr[i] = P.slot_buf[i].p;
    }
  }
  return ptr;
}

namespace {
Operands step_operands(const Step& st, const std::vector<void*>& ptr) {
  return Operands{ptr[st.a], st.b >= 0 ? ptr[st.b] : nullptr, ptr[st.b2 >= 0 ? st.b2 : (st.b >= 0 ? st.b : st.a)],
                  ptr[st.c], 0};
}
}  // namespace

void run_plan_ex(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                 const std::vector<hipEvent_t>& events, unsigned convert_mask) {
  std::vector<void*> ptr = slot_pointers(P, inputs, out, s, convert_mask);
  for (size_t si = 0; si < P.steps.size(); ++si) {
    Step& st = P.steps[si];
    if (!events.empty()) hip_check(hipEventRecord(events[si], s), "event");
    const bool f32 = P.slot_prec[st.a] == 1;
    const bool cf32 = P.slot_prec[st.c] == 1;
    switch (st.kind) {
      case Step::GEMM:
      case Step::AXPY:
        launch_step(st, step_operands(st, ptr), f32, cf32, s, kAllParts, P.precision == MPSXX_BF16X3);
        break;
      case Step::COPY:
        if (!f32 && !cf32)
          launch_copy<double, double>(st.work.as<Chunk>(), st.nwork, st.tasks.as<CopySeg>(),
                                      (const double*)ptr[st.a], (double*)ptr[st.c], st.alpha, st.beta, s);
        else if (f32 && cf32)
          launch_copy<float, float>(st.work.as<Chunk>(), st.nwork, st.tasks.as<CopySeg>(), (const float*)ptr[st.a],
                                    (float*)ptr[st.c], st.alpha, st.beta, s);
        else if (!f32 && cf32)
          launch_copy<double, float>(st.work.as<Chunk>(), st.nwork, st.tasks.as<CopySeg>(),
                                     (const double*)ptr[st.a], (float*)ptr[st.c], st.alpha, st.beta, s);
        else
          launch_copy<float, double>(st.work.as<Chunk>(), st.nwork, st.tasks.as<CopySeg>(), (const float*)ptr[st.a],
                                     (double*)ptr[st.c], st.alpha, st.beta, s);
        break;
    }
  }
  if (!events.empty()) hip_check(hipEventRecord(events[P.steps.size()], s), "event");
  hip_check(hipGetLastError(), "kernel launch");
}

PlanGraph::~PlanGraph() {
  if (exec) {
    (void)hipDeviceSynchronize();  // a replay of this graph may still be running (the plan was evicted meanwhile)
    (void)hipGraphExecDestroy(exec);
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (uploaded) {
    (void)hipEventSynchronize(uploaded);
    (void)hipEventDestroy(uploaded);
  }
  if (pinned) (void)hipHostFree(pinned);
}

namespace {
// the stream plan graphs are captured on (one per host thread; captures are thread-local)
hipStream_t capture_stream() {
  static thread_local hipStream_t cs = [] {
    hipStream_t x = nullptr;
    hip_check(hipStreamCreateWithFlags(&x, hipStreamNonBlocking), "stream");
    return x;
  }();
  return cs;
}
}  // namespace

namespace {
std::atomic<int>& sigma_graph_flag() {
  static std::atomic<int> on([] {
    const char* e = std::getenv("MPSXX_SIGMA_GRAPH");
    return (e && std::string(e) == "0") ? 0 : 1;
  }());
  return on;
}
}  // namespace

bool sigma_graphs_enabled() { return sigma_graph_flag().load() != 0; }
int set_sigma_graphs(int on) { return sigma_graph_flag().exchange(on ? 1 : 0); }

void run_plan_graph(Plan& P, const std::vector<const double*>& inputs, double* out, hipStream_t s,
                    unsigned convert_mask) {
  for (const Step& st : P.steps)
    if (st.kind == Step::COPY) {
      run_plan_ex(P, inputs, out, s, {}, convert_mask);
      return;
    }
  std::vector<void*> ptr = slot_pointers(P, inputs, out, s, convert_mask);  // FP32 copies converted here, on s
  const size_t ns = ptr.size();
  if (!P.graph) {
    auto g = std::make_shared<PlanGraph>();
    g->table.alloc(std::max<size_t>(ns, 1) * sizeof(void*));
    hip_check(hipHostMalloc(&g->pinned, std::max<size_t>(ns, 1) * sizeof(void*), hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipEventCreateWithFlags(&g->uploaded, hipEventDisableTiming), "event");
    P.graph = g;
  }
  PlanGraph& G = *P.graph;
  if (ptr != G.shadow) {
    hip_check(hipEventSynchronize(G.uploaded), "plan table sync");  // the previous upload may still read `pinned`
    std::memcpy(G.pinned, ptr.data(), ns * sizeof(void*));
    hip_check(hipMemcpyAsync(G.table.p, G.pinned, ns * sizeof(void*), hipMemcpyHostToDevice, s), "plan table");
    hip_check(hipEventRecord(G.uploaded, s), "event");
    G.shadow = ptr;
  }
  if (!G.exec) {
    // kernels only: every operand base comes from the table (one instance, so the table is indexed by slot)
    void** T = G.table.as<void*>();
    const hipStream_t cap = capture_stream();
    hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      for (const Step& st : P.steps) {
        const int b2 = st.b2 >= 0 ? st.b2 : (st.b >= 0 ? st.b : st.a);
        const Operands ops{T + st.a, st.b >= 0 ? T + st.b : nullptr, T + b2, T + st.c, 1};
        launch_step(st, ops, P.slot_prec[st.a] == 1, P.slot_prec[st.c] == 1, cap, kAllParts,
                    P.precision == MPSXX_BF16X3);
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cap, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(cap, &G.graph), "end capture");
    hip_check(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0), "graph instantiate");
  }
  hip_check(hipGraphLaunch(G.exec, s), "graph launch");
  hip_check(hipGetLastError(), "graph launch");
}

double time_launches(const std::function<void()>& launch, int reps, hipStream_t s) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  launch();  // warm
  hip_check(hipEventRecord(e0, s), "event");
  for (int r = 0; r < reps; ++r) launch();
  hip_check(hipEventRecord(e1, s), "event");
  hip_check(hipEventSynchronize(e1), "event sync");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

std::vector<PartTiming> profile_plan_parts(Plan& P, const std::vector<const double*>& inputs, double* out, int reps,
                                           hipStream_t s) {
  std::vector<void*> ptr = slot_pointers(P, inputs, out, s, 0u);
  std::vector<PartTiming> res;
  for (size_t si = 0; si < P.steps.size(); ++si) {
    const Step& st = P.steps[si];
    if (st.kind == Step::COPY) continue;
    const bool f32 = P.slot_prec[st.a] == 1, cf32 = P.slot_prec[st.c] == 1;
    const Operands ops = step_operands(st, ptr);
    const bool x3 = P.precision == MPSXX_BF16X3;
    for (KernelPart kp : step_parts(st, f32 ? 4 : 8, cf32 ? 4 : 8)) {
      const double ms = time_launches([&] { launch_step(st, ops, f32, cf32, s, kp.mask, x3); }, reps, s);
      if (x3 && kp.kind == KernelPart::GEMM64) kp.kind = KernelPart::GEMM_X3;
      if (x3 && kp.kind == KernelPart::GEMM128) kp.kind = KernelPart::GEMM_X3_128;
      res.push_back({(int)si, kp, ms});
    }
  }
  return res;
}

const char* kind_name(int k) { return k == Step::GEMM ? "gemm" : k == Step::AXPY ? "axpy" : "copy"; }

}  // namespace mpx
