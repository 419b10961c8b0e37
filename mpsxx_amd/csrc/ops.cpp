// Composite operations: each one is the reference's sequence of btas::Contract / btas::Gemm / btas::Tie / btas::Copy
// calls (cited per function), planned once per input structure and replayed as grouped device launches.
#include "ops.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>

#include "mpsxx_amd.h"
#include "noalloc.hpp"

namespace mpx {

OpStats& last_stats() {
  static thread_local OpStats st;
  return st;
}

LastRun& last_run() {
  static thread_local LastRun r;
  return r;
}

namespace {

enum OpId {
  OP_CONTRACT = 1, OP_SIGMA2, OP_SIGMA1, OP_SIGMAM, OP_SIGMA_OVL, OP_SIGMA_CONJ, OP_DIAG2, OP_DIAG1, OP_DIAGM,
  OP_RENORM_F, OP_RENORM_B, OP_RENORM_OVL_F, OP_RENORM_OVL_B, OP_RENORM_M_F, OP_RENORM_M_B, OP_TIE, OP_UPCAST
};

enum OutMode { OUT_ACCUMULATE = 0, OUT_OVERWRITE = 1 };

struct CachedPlan {
  Plan plan;
  int out_mode = OUT_ACCUMULATE;
  size_t charged = 0;  // device bytes counted against the cache budget at insertion
  uint64_t id = 0;     // identity of this build within the cache
  // the operand layouts the plan was built for, in key order (absent operands empty, present[k] = 0): a hit is
  // taken only when the caller's layouts have this structure, not on equal structure hashes alone
  std::vector<Layout> shape;
  std::vector<char> present;
  bool matches(const std::vector<const Layout*>& ls) const {
    if (ls.size() != shape.size()) return false;
    for (size_t k = 0; k < ls.size(); ++k) {
      if ((ls[k] != nullptr) != (present[k] != 0)) return false;
      if (ls[k] && !ls[k]->same_structure(shape[k])) return false;
    }
    return true;
  }
};

// MPSXX_PLAN_TRACE: host time of each plan build, split into the symbolic contractions, the device steps (tiling
// and task-list uploads) and the intermediate allocations, printed on stderr
struct BuildTrace {
  bool on = std::getenv("MPSXX_PLAN_TRACE") != nullptr;
  double sym_ms = 0.0, step_ms = 0.0, alloc_ms = 0.0;
  int64_t pairs = 0;
};
BuildTrace& btrace() {
  static thread_local BuildTrace t;
  return t;
}
struct TraceScope {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit TraceScope(double& a) : acc(a) {}
  ~TraceScope() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

// Plan cache keyed by (op, precision, structure hashes, extra ints).  Plans of large QC problems hold GB-sized
// intermediates, one set per site, and a sweep cycles through all of them, so a plain LRU over whole plans would
// evict exactly the plans the sweep needs next.  Instead, while the device memory held by the plans exceeds the
// budget (MPSXX_PLAN_CACHE_GB, default 31 % of the device's memory), the least recently used plans release their slot
// buffers and keep their task lists (the costly part to rebuild); a released plan allocates its buffers again from
// the device pool on its next use.  Whole plans are evicted, least recently used first, only beyond kMax plans or
// when the task lists of the released plans alone exceed the budget.  The kKeep most recently used plans are never
// released or evicted,
// so the plans an operation holds at once (a Davidson solve over its operator groups) stay valid.
class PlanCache {
 public:
  // The plan cached under `key` if it was built for operands of the structures `ls` (key order, nullptr for an
  // absent operand).  Where the key holds a plan of other structures (a structure-hash collision), the lookup moves
  // on to the key extended by one more word, and `key` returns the key the plan is (to be) cached under; plans
  // other callers hold stay valid.
  CachedPlan* find(std::vector<uint64_t>& key, const std::vector<const Layout*>& ls) {
    auto it = map_.find(key);
    while (it != map_.end() && !it->second.first->matches(ls)) {
      ++stats_.collisions;
      key.push_back(0x9e3779b97f4a7c15ull);
      it = map_.find(key);
    }
    if (it == map_.end()) {
      miss_t0_ = std::chrono::steady_clock::now();  // a miss is followed by the build and insert()
      BuildTrace& bt = btrace();
      bt.sym_ms = bt.step_ms = bt.alloc_ms = 0.0;
      bt.pairs = 0;
      return nullptr;
    }
    ++stats_.hits;
    return touch(it);
  }
  // the plan cached under `key` for the structures `ls`, or nullptr; no build follows a miss (counted as neither)
  CachedPlan* lookup(const std::vector<uint64_t>& key, const std::vector<const Layout*>& ls) {
    auto it = map_.find(key);
    if (it == map_.end() || !it->second.first->matches(ls)) return nullptr;
    return touch(it);
  }
  // identity of the plan cached under `key` (0: none)
  uint64_t id_of(const std::vector<uint64_t>& key) const {
    auto it = map_.find(key);
    return it == map_.end() ? 0 : it->second.first->id;
  }
  // lookup that builds nothing: the plan with identity `id` if it is still cached under `key`, else nullptr
  CachedPlan* peek(const std::vector<uint64_t>& key, uint64_t id) {
    auto it = map_.find(key);
    if (it == map_.end() || it->second.first->id != id) return nullptr;
    return touch(it);
  }
  CachedPlan* insert(const std::vector<uint64_t>& key, std::unique_ptr<CachedPlan> p,
                     const std::vector<const Layout*>& ls) {
    ++stats_.builds;
    p->id = ++next_id_;
    p->shape.assign(ls.size(), Layout());
    p->present.assign(ls.size(), 0);
    for (size_t k = 0; k < ls.size(); ++k)
      if (ls[k]) {
        p->shape[k] = *ls[k];
        p->present[k] = 1;
      }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - miss_t0_).count();
    stats_.build_ms += ms;
    const BuildTrace& bt = btrace();
    if (bt.on)
      std::fprintf(stderr, "[plan] op %d build %.2f ms: symbolic %.2f steps %.2f alloc %.2f, %lld pairs, %zu steps\n",
                   (int)key[0], ms, bt.sym_ms, bt.step_ms, bt.alloc_ms, (long long)bt.pairs, p->plan.steps.size());
    const size_t add = p->plan.device_bytes();
    make_room(add);
    lru_.push_front(key);
    CachedPlan* raw = p.get();
    raw->charged = add;
    map_[key] = {std::move(p), lru_.begin()};
    bytes_ += add;
    return raw;
  }
  void clear() {
    map_.clear();
    lru_.clear();
    bytes_ = 0;
  }
  PlanCacheStats stats() const {
    PlanCacheStats st = stats_;
    st.plans = (int64_t)map_.size();
    st.device_bytes = (int64_t)bytes_;
    return st;
  }

 private:
  // make `add` more bytes fit: release buffers of the least recently used plans, then evict whole plans
  void make_room(size_t add) {
    auto evict_back = [&] {
      auto victim = lru_.back();
      lru_.pop_back();
      auto it = map_.find(victim);
      bytes_ -= it->second.first->charged;
      map_.erase(it);
      ++stats_.evictions;
    };
    while (map_.size() > kKeep && map_.size() >= kMax) evict_back();
    size_t from_back = 0;
    for (auto it = lru_.rbegin(); it != lru_.rend() && bytes_ + add > budget(); ++it, ++from_back) {
      if (map_.size() - from_back <= kKeep) break;
      CachedPlan* cp = map_.find(*it)->second.first.get();
      if (cp->plan.stripped) continue;
      cp->plan.strip();
      ++stats_.releases;
      bytes_ -= cp->charged;
      cp->charged = cp->plan.device_bytes();
      bytes_ += cp->charged;
    }
    // whole plans go only while the released plans' task lists alone exceed the budget: the kKeep most recent plans
    // (the ones in use) are not counted, since evicting other plans would not shrink them
    auto kept = [&] {
      size_t b = 0, n = 0;
      for (auto it = lru_.begin(); it != lru_.end() && n < kKeep; ++it, ++n) b += map_.find(*it)->second.first->charged;
      return b;
    };
    while (map_.size() > kKeep && bytes_ - kept() + add > budget()) evict_back();
  }
  using Map = std::map<std::vector<uint64_t>,
                       std::pair<std::unique_ptr<CachedPlan>, std::list<std::vector<uint64_t>>::iterator>>;
  // a hit: most recently used, slot buffers allocated again if they were released
  CachedPlan* touch(Map::iterator it) {
    lru_.splice(lru_.begin(), lru_, it->second.second);
    CachedPlan* cp = it->second.first.get();
    if (cp->plan.stripped) {
      cp->plan.restore();
      ++stats_.restores;
      bytes_ -= cp->charged;
      cp->charged = cp->plan.device_bytes();
      bytes_ += cp->charged;
      make_room(0);
    }
    return cp;
  }
  PlanCacheStats stats_;
  uint64_t next_id_ = 0;
  std::chrono::steady_clock::time_point miss_t0_ = std::chrono::steady_clock::now();
  static size_t budget() {
    static const size_t b = [] {
      if (const char* e = std::getenv("MPSXX_PLAN_CACHE_GB"))
        return (size_t)(std::max(1e-3, std::atof(e)) * 1073741824.0);  // fractions of a GB allowed (tests)
      return device_memory_fraction(0.31, 64.0);
    }();
    return b;
  }
  static constexpr size_t kMax = 1024;
  static constexpr size_t kKeep = kPlanCacheKeep;
  size_t bytes_ = 0;
  Map map_;
  std::list<std::vector<uint64_t>> lru_;
};

PlanCache& cache() {
  static PlanCache c;
  return c;
}


}  // namespace

PlanCacheStats plan_cache_stats() { return cache().stats(); }
void plan_cache_clear() { cache().clear(); }

namespace {
thread_local std::vector<uint64_t> t_plan_key;  // key of the plan the last plan_sigma_* call returned
}

void record_last_run(const std::vector<uint64_t>& key, const std::vector<Tensor*>& ins) {
  LastRun& r = last_run();
  r.key = key;
  r.plan_id = cache().id_of(key);
  r.tensors.clear();
  r.inputs.clear();
  r.hashes.clear();
  for (const Tensor* t : ins) {
    r.tensors.push_back(t);
    r.inputs.push_back(t->ptr());
    r.hashes.push_back(t->L.hash);
  }
}

void record_last_run(const std::vector<uint64_t>& key, std::initializer_list<const Tensor*> ins) {
  std::vector<Tensor*> v;
  for (const Tensor* t : ins) v.push_back(const_cast<Tensor*>(t));
  record_last_run(key, v);
}

Plan& last_plan(std::vector<const double*>& inputs) {
  LastRun& r = last_run();
  if (r.key.empty()) throw Error(MPSXX_ERR_ARG, "profile: no hot-path call to profile on this thread");
  for (size_t i = 0; i < r.tensors.size(); ++i)
    if (!Tensor::alive(r.tensors[i]) || r.tensors[i]->ptr() != r.inputs[i] || r.tensors[i]->L.hash != r.hashes[i])
      throw Error(MPSXX_ERR_ARG, "profile: an input of the last hot-path call was freed or changed since");
  CachedPlan* cp = cache().peek(r.key, r.plan_id);
  if (!cp) throw Error(MPSXX_ERR_ARG, "profile: the plan of the last hot-path call was evicted from the cache");
  for (size_t i = 0; i < r.tensors.size(); ++i)
    if (i >= cp->shape.size() || !cp->present[i] || !r.tensors[i]->L.same_structure(cp->shape[i]))
      throw Error(MPSXX_ERR_ARG, "profile: an input of the last hot-path call changed its block structure since");
  inputs = r.inputs;
  return cp->plan;
}

namespace {

class PlanBuilder {
 public:
  explicit PlanBuilder(int prec) {
    if (prec != MPSXX_FP64 && prec != MPSXX_FP32 && prec != MPSXX_BF16X3) throw Error(MPSXX_ERR_ARG, "unknown precision");
    P_.precision = prec;
  }
  // blocks with both sides <= this run as 32x32 tiles in this plan (negative: the default policy's)
  void set_small_limit(int v) { small_limit_ = v; }
  int small_limit_ = -1;

  int input(const Layout& L, int which) {
    const int id = (int)P_.slot_layout.size();
    P_.slot_layout.push_back(L);
    P_.slot_input.push_back(which);
    P_.slot_prec.push_back(P_.precision != MPSXX_FP64 ? 1 : 0);
    P_.slot_buf.emplace_back();
    if (P_.precision != MPSXX_FP64) P_.slot_buf.back().alloc(std::max<int64_t>(L.nelem, 1) * sizeof(float));
    P_.bytes += L.nelem * 8;
    return id;
  }

  // intermediate = Contract(A[ca], B[cb]); with `mask`, only the output blocks that mask's quantum numbers allow (a
  // layout with the output's sector counts onto which the intermediate is up-cast afterwards)
  int contract(int a, bool ca_conj, const std::vector<int>& ca, int b, bool cb_conj, const std::vector<int>& cb,
               const Layout* mask = nullptr) {
    View va{&P_.slot_layout[a], ca_conj}, vb{&P_.slot_layout[b], cb_conj};
    ContractSym sym;
    {
      TraceScope ts(btrace().sym_ms);
      sym = contract_symbolic(va, ca, vb, cb, nullptr, mask);
    }
    btrace().pairs += (int64_t)sym.pairs.size();
    const int id = add_inter(sym.C, P_.precision != MPSXX_FP64 ? 1 : 0);
    TraceScope ts(btrace().step_ms);
    P_.steps.push_back(make_contract_step(sym, a, b, id, 0, 1.0, policy()));
    P_.flops += sym.flops;
    return id;
  }

  // output (+)= Contract(A[ca], B[cb]); out_existing may be null (empty output)
  void contract_out(int a, bool ca_conj, const std::vector<int>& ca, int b, bool cb_conj, const std::vector<int>& cb,
                    const Layout* out_existing, double alpha = 1.0) {
    View va{&P_.slot_layout[a], ca_conj}, vb{&P_.slot_layout[b], cb_conj};
    ContractSym sym;
    {
      TraceScope ts(btrace().sym_ms);
      sym = contract_symbolic(va, ca, vb, cb, out_existing);
    }
    btrace().pairs += (int64_t)sym.pairs.size();
    const int id = add_output(sym.C);
    TraceScope ts(btrace().step_ms);
    P_.steps.push_back(make_contract_step(sym, a, b, id, 1, alpha, policy()));
    P_.flops += sym.flops;
    P_.bytes += sym.C.nelem * 8;
  }

  // intermediate = Contract(Contract(Y[ca1], W1[cb1])[ca2], W2[cb2]) as one fused coefficient pass when both are
  // contractions with a scalar (MPO) block; the inner intermediate's structure is still derived (it fixes the
  // reference's block pattern) but never stored.  Falls back to two steps otherwise.
  int contract2(int y, const std::vector<int>& ca1, int w1, const std::vector<int>& cb1, const std::vector<int>& ca2,
                int w2, const std::vector<int>& cb2) {
    static const bool nofuse = std::getenv("MPSXX_NO_FUSE") != nullptr;
    View vy{&P_.slot_layout[y], false}, vw1{&P_.slot_layout[w1], false};
    ContractSym s1, s2;
    {
      TraceScope ts(btrace().sym_ms);
      s1 = contract_symbolic(vy, ca1, vw1, cb1, nullptr);
      View vx{&s1.C, false}, vw2{&P_.slot_layout[w2], false};
      s2 = contract_symbolic(vx, ca2, vw2, cb2, nullptr);
    }
    btrace().pairs += (int64_t)(s1.pairs.size() + s2.pairs.size());
    TraceScope ts(btrace().step_ms);
    Step fused;
    const int prec = P_.precision != MPSXX_FP64 ? 1 : 0;
    if (!nofuse && make_fused_axpy_step(s1, s2, y, w1, w2, (int)P_.slot_layout.size(), 0, fused)) {
      const int id = add_inter(s2.C, prec);
      P_.steps.push_back(std::move(fused));
      P_.flops += s1.flops + s2.flops;
      return id;
    }
    const int x = add_inter(s1.C, prec);
    P_.steps.push_back(make_contract_step(s1, y, w1, x, 0, 1.0, policy()));
    const int z = add_inter(s2.C, prec);
    P_.steps.push_back(make_contract_step(s2, x, w2, z, 0, 1.0, policy()));
    P_.flops += s1.flops + s2.flops;
    return z;
  }

  int tie(int x, const std::vector<int>& t) {
    std::vector<CopySeg> segs;
    Layout Y = tie_layout(P_.slot_layout[x], t, segs);
    const int id = add_inter(Y, 0);
    P_.steps.push_back(make_copy_step(segs, x, id, 1.0, 0.0));
    return id;
  }

  // output := up-casting copy of slot x onto (q, qshape) of `target` (btas::Copy(x, y, 1): y's blocks are cleared
  // first, then every x block allowed in y is copied; STBLAS.h:374-388, 48-60)
  void upcast_out(int x, const Layout& target) {
    const Layout& X = P_.slot_layout[x];
    Layout Y;
    Y.set_shape(target.rank, target.qsym, target.q, target.qs, X.ds);
    if (X.rank != target.rank) throw Error(MPSXX_ERR_SHAPE, "Copy: rank mismatch");
    for (int i = 0; i < X.rank; ++i)
      if (X.ds[i].size() != target.ds[i].size())
        throw Error(MPSXX_ERR_SHAPE, "Copy(SPARSE): x and y must have the same shape as long as up-casting is specified.");
    for (int64_t b = 0; b < X.nblocks(); ++b)
      if (Y.allowed(X.index(b))) Y.add_block(X.index(b));
    Y.finalize();
    std::vector<CopySeg> segs;
    for (int64_t b = 0; b < X.nblocks(); ++b) {
      const int64_t j = Y.find(X.tag[b]);
      if (j < 0 || X.bsize[b] == 0) continue;
      segs.push_back({Y.off[j], X.off[b], (int32_t)X.bsize[b], 1});
    }
    const int id = add_output(Y);
    P_.steps.push_back(make_copy_step(segs, x, id, 1.0, 0.0));
    P_.bytes += Y.nelem * 8;
  }

  Plan finish() { return std::move(P_); }

 private:
  // tile policy of the plan's GEMM steps
  GemmPolicy policy() const {
    GemmPolicy p = default_policy(P_.precision);
    if (small_limit_ >= 0) p.small_limit = small_limit_;
    return p;
  }

  int add_inter(const Layout& L, int prec) {
    const int id = (int)P_.slot_layout.size();
    P_.slot_layout.push_back(L);
    P_.slot_input.push_back(-1);
    P_.slot_prec.push_back(prec);
    TraceScope ts(btrace().alloc_ms);
    P_.slot_buf.emplace_back(std::max<int64_t>(L.nelem, 1) * (prec ? sizeof(float) : sizeof(double)));
    return id;
  }
  int add_output(const Layout& L) {
    const int id = (int)P_.slot_layout.size();
    P_.slot_layout.push_back(L);
    P_.slot_input.push_back(-2);
    P_.slot_prec.push_back(0);
    P_.slot_buf.emplace_back();
    P_.out = L;
    return id;
  }
  Plan P_;
};

void require(const Tensor& t, const char* what) {
  if (t.empty) throw Error(MPSXX_ERR_ARG, std::string(what) + ": input array is empty");
}

std::vector<uint64_t> make_key(int op, int prec, std::initializer_list<const Tensor*> ins, const Tensor* out,
                               std::initializer_list<int64_t> extra = {}) {
  std::vector<uint64_t> k;
  k.push_back((uint64_t)op);
  k.push_back((uint64_t)prec);
  for (const Tensor* t : ins) k.push_back(t->L.hash);
  k.push_back(out && !out->empty ? out->L.hash : 0ull);
  for (int64_t e : extra) k.push_back((uint64_t)e);
  return k;
}

// the operand layouts a plan cached for (ins..., out) is checked against on a hit; an empty output is absent
std::vector<const Layout*> key_layouts(const std::vector<Tensor*>& ins, const Tensor* out) {
  std::vector<const Layout*> ls;
  for (const Tensor* t : ins) ls.push_back(&t->L);
  ls.push_back(out && !out->empty ? &out->L : nullptr);
  return ls;
}

// Look up or build a plan, shape the output, run.
template <typename BuildFn>
void execute(std::vector<uint64_t> key, BuildFn build, int out_mode, std::vector<Tensor*> ins, Tensor& out,
             hipStream_t s, unsigned convert_mask = ~0u) {
  const std::vector<const Layout*> ls = key_layouts(ins, &out);
  CachedPlan* cp = cache().find(key, ls);
  if (!cp) {
    auto np = std::make_unique<CachedPlan>();
    np->plan = build();
    np->out_mode = out_mode;
    cp = cache().insert(key, std::move(np), ls);
  }
  Plan& P = cp->plan;
  if (out_mode == OUT_OVERWRITE || out.empty) out.set_layout(P.out, s);
  else if (!out.L.same_structure(P.out)) out.relayout(P.out, s);
  std::vector<const double*> ptrs;
  for (Tensor* t : ins) ptrs.push_back(t->ptr());
  run_plan_ex(P, ptrs, out.ptr(), s, {}, convert_mask);
  last_stats().flops = P.flops;
  last_stats().bytes = P.bytes;
  record_last_run(key, ins);
}

const Layout* existing(const Tensor& t) { return t.empty ? nullptr : &t.L; }

}  // namespace

void clear_plan_cache() {
  cache().clear();
  clear_merged_operator_cache();
  last_run() = LastRun();
}

// ---------------------------------------------------------------------------------------------------------------------
// generic machinery

void op_contract(double alpha, Tensor& a, bool conj_a, const std::vector<int>& ca, Tensor& b, bool conj_b,
                 const std::vector<int>& cb, double beta, Tensor& c, hipStream_t s) {
  require(a, "Contract");
  require(b, "Contract");
  if (!c.empty && beta != 1.0) op_scal(beta, c, s);  // QSTBLAS.h:206 (Scal(beta, c))
  auto key = make_key(OP_CONTRACT, 0, {&a, &b}, &c, {conj_a, conj_b, (int64_t)ca.size()});
  for (int x : ca) key.push_back(x);
  for (int x : cb) key.push_back(x + 100);
  uint64_t ab;
  std::memcpy(&ab, &alpha, 8);
  key.push_back(ab);
  const Layout* ex = existing(c);
  execute(key, [&] {
    PlanBuilder B(MPSXX_FP64);
    const int sa = B.input(a.L, 0), sb = B.input(b.L, 1);
    B.contract_out(sa, conj_a, ca, sb, conj_b, cb, ex, alpha);
    return B.finish();
  }, OUT_ACCUMULATE, {&a, &b}, c, s);
}

void op_gemm(int ta, int tb, double alpha, Tensor& a, bool conj_a, Tensor& b, bool conj_b, double beta, Tensor& c,
             hipStream_t s) {
  require(a, "Gemm");
  require(b, "Gemm");
  int NC;
  if (!c.empty) NC = c.L.rank;
  else throw Error(MPSXX_ERR_ARG, "Gemm: the output rank must be known (create the output with its structure)");
  std::vector<int> ca, cb;
  gemm_indices(ta, tb, a.L.rank, b.L.rank, NC, ca, cb);
  op_contract(alpha, a, conj_a, ca, b, conj_b, cb, beta, c, s);
}

void op_tie(Tensor& x, const std::vector<int>& t, Tensor& y, hipStream_t s) {
  require(x, "Tie");
  auto key = make_key(OP_TIE, 0, {&x}, nullptr, {(int64_t)t.size()});
  for (int v : t) key.push_back(v);
  execute(key, [&] {
    PlanBuilder B(MPSXX_FP64);
    const int sx = B.input(x.L, 0);
    std::vector<CopySeg> segs;
    Layout Y = tie_layout(x.L, t, segs);
    (void)sx;
    // a one-step plan whose output is the tie result
    Plan P = B.finish();
    P.slot_layout.push_back(Y);
    P.slot_input.push_back(-2);
    P.slot_prec.push_back(0);
    P.slot_buf.emplace_back();
    P.out = Y;
    P.steps.push_back(make_copy_step(segs, 0, 1, 1.0, 0.0));
    return P;
  }, OUT_OVERWRITE, {&x}, y, s);
}

void op_copy_upcast(Tensor& x, Tensor& y, hipStream_t s) {
  require(x, "Copy");
  if (y.empty) throw Error(MPSXX_ERR_ARG, "Copy(UpCast): target must carry its quantum numbers");
  auto key = make_key(OP_UPCAST, 0, {&x}, &y);
  Layout target = y.L;
  execute(key, [&] {
    PlanBuilder B(MPSXX_FP64);
    const int sx = B.input(x.L, 0);
    B.upcast_out(sx, target);
    return B.finish();
  }, OUT_OVERWRITE, {&x}, y, s);
}

// ---------------------------------------------------------------------------------------------------------------------
// BLAS-1 (QSTBLAS.h:40-93, STBLAS.h:94-145)

namespace {

struct SegList {
  std::vector<CopySeg> segs;
  std::vector<Chunk> chunks;
};

// segments y[dst] <- x[src] of the blocks x and y share, and their kAxpyChunk-element chunks (r's lists are
// overwritten, keeping their capacity)
void match_segments_into(const Layout& x, const Layout& y, SegList& r) {
  r.segs.clear();
  r.chunks.clear();
  for (int64_t b = 0; b < x.nblocks(); ++b) {
    const int64_t j = y.find(x.tag[b]);
    if (j < 0 || x.bsize[b] == 0) continue;
    r.segs.push_back({y.off[j], x.off[b], (int32_t)x.bsize[b], 1});
  }
  for (size_t i = 0; i < r.segs.size(); ++i)
    for (int s = 0; s < r.segs[i].len; s += kAxpyChunk) r.chunks.push_back({(int32_t)i, s});
}

SegList match_segments(const Layout& x, const Layout& y) {
  SegList r;
  match_segments_into(x, y, r);
  return r;
}

struct Scratch {
  DevBuf partial, result;
  size_t npart = 0;
  double* get_partial(size_t n) {
    if (n > npart) {
      npart = std::max<size_t>(n, 4096);
      partial.alloc(npart * sizeof(double));
    }
    if (!result.p) result.alloc(64 * sizeof(double));
    return partial.as<double>();
  }
};

Scratch& scratch() {
  static Scratch s;
  return s;
}

void check_same_q(const Tensor& x, const Tensor& y, const char* what) {
  if (x.L.qsym && y.L.qsym) {
    if (x.L.q != y.L.q) throw Error(MPSXX_ERR_QN, std::string(what) + "(QSPARSE): quantum numbers of x and y must be the same.");
    if (!x.L.same_qshape(y.L)) throw Error(MPSXX_ERR_QN, std::string(what) + "(QSPARSE): quantum shapes of x and y must be the same.");
  }
}

}  // namespace

double op_dotc(Tensor& x, Tensor& y, hipStream_t s) {
  if (x.empty || y.empty) return 0.0;
  check_same_q(x, y, "Dotc");
  SegList m = match_segments(y.L, x.L);  // dst = x offset, src = y offset
  double* part = scratch().get_partial(m.chunks.size() + 1);
  double* d_out = scratch().result.as<double>();
  DevBuf dsegs = upload_vec(m.segs), dch = upload_vec(m.chunks);
  launch_dot(dch.as<Chunk>(), (int)m.chunks.size(), dsegs.as<CopySeg>(), x.ptr(), y.ptr(), part, d_out, s);
  double h = 0.0;
  hip_check(hipMemcpyAsync(&h, d_out, sizeof(double), hipMemcpyDeviceToHost, s), "dot result");
  hip_check(hipStreamSynchronize(s), "dot sync");
  return h;
}

double op_dotc_noalloc(Tensor& x, Tensor& y, hipStream_t s) {
  if (x.empty || y.empty) return 0.0;
  check_same_q(x, y, "Dotc");
  static thread_local SegList m;  // host lists, reused from call to call
  match_segments_into(y.L, x.L, m);  // dst = x offset, src = y offset, as op_dotc
  // scratch layout: segments | chunks | partial sums | result, each part 256-byte aligned
  auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
  const size_t nseg = m.segs.size(), nch = m.chunks.size();
  const size_t off_ch = up(nseg * sizeof(CopySeg)), off_part = off_ch + up(nch * sizeof(Chunk));
  const size_t off_res = off_part + up((nch + 1) * sizeof(double)), total = off_res + 256;
  if (total > kNoAllocScratchBytes)
    throw Error(MPSXX_ERR_ARG, "Dotc (no allocation): " + std::to_string(nseg) + " block segments in " +
                                   std::to_string(nch) + " chunks need " + std::to_string(total) +
                                   " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                                   " the library holds per device");
  char* base = static_cast<char*>(noalloc_scratch());
  if (nseg)
    hip_check(hipMemcpyAsync(base, m.segs.data(), nseg * sizeof(CopySeg), hipMemcpyHostToDevice, s), "dot segments");
  if (nch)
    hip_check(hipMemcpyAsync(base + off_ch, m.chunks.data(), nch * sizeof(Chunk), hipMemcpyHostToDevice, s),
              "dot chunks");
  launch_dot(reinterpret_cast<const Chunk*>(base + off_ch), (int)nch, reinterpret_cast<const CopySeg*>(base),
             x.ptr(), y.ptr(), reinterpret_cast<double*>(base + off_part), reinterpret_cast<double*>(base + off_res),
             s);
  hip_check(hipGetLastError(), "dot launch");
  double h = 0.0;
  hip_check(hipMemcpyAsync(&h, base + off_res, sizeof(double), hipMemcpyDeviceToHost, s), "dot result");
  hip_check(hipStreamSynchronize(s), "dot sync");
  return h;
}

double op_nrm2(Tensor& x, hipStream_t s) { return std::sqrt(op_dotc(x, x, s)); }

void op_axpy(double alpha, Tensor& x, Tensor& y, hipStream_t s) {
  if (x.empty) return;
  if (y.empty) {
    Layout l = x.L;
    l.tag.clear();
    l.bidx.clear();
    l.finalize();
    y.set_layout(l, s);
  } else {
    check_same_q(x, y, "Axpy");
  }
  // y.reserve(each x block) (STBLAS.h:135-145)
  bool missing = false;
  for (int64_t b = 0; b < x.L.nblocks() && !missing; ++b) missing = y.L.find(x.L.tag[b]) < 0;
  if (missing) {
    Layout u = y.L;
    for (int64_t b = 0; b < x.L.nblocks(); ++b) u.add_block(x.L.index(b));
    u.finalize();
    y.relayout(u, s);
  }
  SegList m = match_segments(x.L, y.L);
  DevBuf dsegs = upload_vec(m.segs), dch = upload_vec(m.chunks);
  launch_copy<double, double>(dch.as<Chunk>(), (int)m.chunks.size(), dsegs.as<CopySeg>(), x.ptr(), y.ptr(), alpha,
                              1.0, s);
  hip_check(hipStreamSynchronize(s), "axpy sync");
}

void op_scal(double alpha, Tensor& x, hipStream_t s) {
  if (x.empty) return;
  launch_scale(x.ptr(), x.L.nelem, alpha, s);
}

void op_copy(Tensor& x, Tensor& y, hipStream_t s) {
  if (x.empty) {
    y.empty = true;
    y.data.release();
    y.L = Layout();
    return;
  }
  y.set_layout(x.L, s);
  hip_check(hipMemcpyAsync(y.ptr(), x.ptr(), x.L.nelem * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
}

// ---------------------------------------------------------------------------------------------------------------------
// sigma vectors (src/compute_sigma_vector.hpp)

Plan build_plan_sigma_twosite(const Layout& lmpo, const Layout& rmpo, const Layout& lopr, const Layout& ropr,
                              const Layout& wfn, const Layout* out, int prec) {
  PlanBuilder B(prec);
  const int s_lmpo = B.input(lmpo, 0), s_rmpo = B.input(rmpo, 1), s_lopr = B.input(lopr, 2);
  const int s_ropr = B.input(ropr, 3), s_wfn = B.input(wfn, 4);
  // compute_sigma_vector.hpp:63-69
  const int s1 = B.contract(s_lopr, false, {2}, s_wfn, false, {0});
  // the two MPO-coefficient contractions (scalar MPO blocks) run as one fused pass
  const int s3 = B.contract2(s1, {1, 2}, s_lmpo, {0, 2}, {4, 1}, s_rmpo, {0, 2});
  B.contract_out(s3, false, {4, 1}, s_ropr, false, {1, 2}, out);
  return B.finish();
}

// A sigma plan of precision `prec` made from the cached FP64 plan of the same key and structures, when there is one
// and both precisions tile alike: the FP64 plan's steps (the costly symbolic contraction, tiling and task lists)
// copied, its inputs and intermediates given FP32 storage (the output stays FP64), as PlanBuilder lays out a
// low-precision sigma plan.  A mixed-precision Davidson solve builds its FP64 plans first, so its FP32 or split-BF16
// plans come from here.  Returns false (nothing made) otherwise.
bool plan_from_fp64(std::vector<uint64_t> key, const std::vector<const Layout*>& ls, int prec, Plan& out) {
  static const bool off = std::getenv("MPSXX_NO_PLAN_REUSE") != nullptr;
  if (off || prec == MPSXX_FP64) return false;
  const GemmPolicy p64 = default_policy(MPSXX_FP64), pp = default_policy(prec);
  if (p64.small_limit != pp.small_limit || p64.split_units != pp.split_units || p64.max_groups != pp.max_groups ||
      p64.split_big != pp.split_big || p64.split_small != pp.split_small || p64.split_seg != pp.split_seg ||
      p64.big_tile_min != pp.big_tile_min)
    return false;
  key[1] = MPSXX_FP64;
  key.back() = 0;  // FP64 plans are shared by equal-structure groups (copy 0)
  CachedPlan* src = cache().lookup(key, ls);
  if (!src) return false;
  const Plan& P = src->plan;
  for (const Step& st : P.steps)
    if (st.kind == Step::COPY) return false;
  Plan Q;
  Q.precision = prec;
  Q.out = P.out;
  Q.flops = P.flops;
  Q.bytes = P.bytes;
  Q.slot_layout = P.slot_layout;
  Q.slot_input = P.slot_input;
  for (size_t i = 0; i < P.slot_layout.size(); ++i) {
    const bool low = P.slot_input[i] != -2;  // inputs and intermediates in FP32, the output in FP64
    Q.slot_prec.push_back(low ? 1 : 0);
    Q.slot_buf.emplace_back();
    if (low) Q.slot_buf.back().alloc(std::max<int64_t>(P.slot_layout[i].nelem, 1) * sizeof(float));
  }
  for (const Step& st : P.steps) Q.steps.push_back(clone_step(st));
  out = std::move(Q);
  return true;
}

Plan& plan_sigma_twosite(const Layout& lmpo, const Layout& rmpo, const Layout& lopr, const Layout& ropr,
                         const Layout& wfn, const Layout* out, int prec, int copy) {
  std::vector<uint64_t> key{OP_SIGMA2, (uint64_t)prec, lmpo.hash, rmpo.hash, lopr.hash, ropr.hash, wfn.hash,
                            out ? out->hash : 0ull, (uint64_t)copy};
  const std::vector<const Layout*> ls{&lmpo, &rmpo, &lopr, &ropr, &wfn, out};
  CachedPlan* cp = cache().find(key, ls);
  t_plan_key = key;
  if (cp) return cp->plan;
  auto np = std::make_unique<CachedPlan>();
  if (!plan_from_fp64(key, ls, prec, np->plan))
    np->plan = build_plan_sigma_twosite(lmpo, rmpo, lopr, ropr, wfn, out, prec);
  return cache().insert(key, std::move(np), ls)->plan;
}

void op_sigma_twosite(Tensor& lmpo, Tensor& rmpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec,
                      hipStream_t s, unsigned convert_mask) {
  for (Tensor* t : {&lmpo, &rmpo, &lopr, &ropr, &wfn}) require(*t, "compute_sigma_vector");
  Plan& P = plan_sigma_twosite(lmpo.L, rmpo.L, lopr.L, ropr.L, wfn.L, existing(sgv), prec);
  if (sgv.empty) sgv.set_layout(P.out, s);
  else if (!sgv.L.same_structure(P.out)) sgv.relayout(P.out, s);
  run_plan_ex(P, {lmpo.ptr(), rmpo.ptr(), lopr.ptr(), ropr.ptr(), wfn.ptr()}, sgv.ptr(), s, {}, convert_mask);
  last_stats().flops = P.flops;
  last_stats().bytes = P.bytes;
  record_last_run(t_plan_key, {&lmpo, &rmpo, &lopr, &ropr, &wfn});
}

Plan& plan_sigma_onesite(const Layout& mpo, const Layout& lopr, const Layout& ropr, const Layout& wfn,
                         const Layout* out, int prec, int copy) {
  std::vector<uint64_t> key{OP_SIGMA1, (uint64_t)prec, mpo.hash, lopr.hash, ropr.hash, wfn.hash,
                            out ? out->hash : 0ull, (uint64_t)copy};
  const std::vector<const Layout*> ls{&mpo, &lopr, &ropr, &wfn, out};
  CachedPlan* cp = cache().find(key, ls);
  t_plan_key = key;
  if (cp) return cp->plan;
  auto np = std::make_unique<CachedPlan>();
  if (!plan_from_fp64(key, ls, prec, np->plan)) {
    PlanBuilder B(prec);
    const int sm = B.input(mpo, 0), sl = B.input(lopr, 1), sr = B.input(ropr, 2), sw = B.input(wfn, 3);
    // compute_sigma_vector.hpp:17-21: Gemm(N,N,lopr,wfn0) -> Contract(mpo) -> Contract(ropr)
    const int s1 = B.contract(sl, false, {2}, sw, false, {0});
    const int s2 = B.contract(s1, false, {1, 2}, sm, false, {0, 2});
    B.contract_out(s2, false, {3, 1}, sr, false, {1, 2}, out);
    np->plan = B.finish();
  }
  return cache().insert(key, std::move(np), ls)->plan;
}

void op_sigma_onesite(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec, hipStream_t s) {
  for (Tensor* t : {&mpo, &lopr, &ropr, &wfn}) require(*t, "compute_sigma_vector");
  Plan& P = plan_sigma_onesite(mpo.L, lopr.L, ropr.L, wfn.L, existing(sgv), prec);
  if (sgv.empty) sgv.set_layout(P.out, s);
  else if (!sgv.L.same_structure(P.out)) sgv.relayout(P.out, s);
  run_plan_ex(P, {mpo.ptr(), lopr.ptr(), ropr.ptr(), wfn.ptr()}, sgv.ptr(), s, {}, ~0u);
  last_stats().flops = P.flops;
  last_stats().bytes = P.bytes;
  record_last_run(t_plan_key, {&mpo, &lopr, &ropr, &wfn});
}

Plan& plan_sigma_merged(const Layout& lopr, const Layout& ropr, const Layout& wfn, const Layout* out, int prec,
                        int copy) {
  std::vector<uint64_t> key{OP_SIGMAM, (uint64_t)prec, lopr.hash, ropr.hash, wfn.hash, out ? out->hash : 0ull,
                            (uint64_t)copy};
  const std::vector<const Layout*> ls{&lopr, &ropr, &wfn, out};
  CachedPlan* cp = cache().find(key, ls);
  t_plan_key = key;
  if (cp) return cp->plan;
  auto np = std::make_unique<CachedPlan>();
  if (!plan_from_fp64(key, ls, prec, np->plan)) {
    PlanBuilder B(prec);
    // the merged sigma's 33..64-row blocks run faster on 64x64 than on 32x32 tiles (the converged M=1024 QC state:
    // FP32 4.0 -> 3.85 ms, split-BF16x3 3.1 -> 2.95 ms), while the two-site sigma's keep the default of 64 (its
    // L=32 M=512 Hubbard form would take 60 instead of 40 us); MPSXX_SMALL_GEMM_LIMIT, when set, applies to both
    if (!std::getenv("MPSXX_SMALL_GEMM_LIMIT")) B.set_small_limit(kMergedSmallLimit);
    const int sl = B.input(lopr, 0), sr = B.input(ropr, 1), sw = B.input(wfn, 2);
    // compute_sigma_vector.hpp:94-95: Gemm(N,N,lopr,wfn0) (K=1), Gemm(N,T,scr1,ropr) (K=2)
    const int s1 = B.contract(sl, false, {2}, sw, false, {0});
    B.contract_out(s1, false, {1, 2}, sr, false, {1, 2}, out);
    np->plan = B.finish();
  }
  return cache().insert(key, std::move(np), ls)->plan;
}

void op_sigma_merged(Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, int prec, hipStream_t s,
                     unsigned convert_mask) {
  for (Tensor* t : {&lopr, &ropr, &wfn}) require(*t, "compute_sigma_vector");
  if (lopr.L.rank != 3 || ropr.L.rank != 3 || wfn.L.rank != 2)
    throw Error(MPSXX_ERR_ARG, "merged sigma: expects 3-index operators and a 2-index wavefunction");
  Plan& P = plan_sigma_merged(lopr.L, ropr.L, wfn.L, existing(sgv), prec);
  if (sgv.empty) sgv.set_layout(P.out, s);
  else if (!sgv.L.same_structure(P.out)) sgv.relayout(P.out, s);
  run_plan_ex(P, {lopr.ptr(), ropr.ptr(), wfn.ptr()}, sgv.ptr(), s, {}, convert_mask);
  last_stats().flops = P.flops;
  last_stats().bytes = P.bytes;
  record_last_run(t_plan_key, {&lopr, &ropr, &wfn});
}

void op_sigma_overlap(Tensor& lsov, Tensor& rsov, Tensor& wfn, Tensor& sgv, hipStream_t s) {
  for (Tensor* t : {&lsov, &rsov, &wfn}) require(*t, "compute_sigma_vector");
  if (lsov.L.rank != 2 || rsov.L.rank != 2) throw Error(MPSXX_ERR_ARG, "overlap sigma: expects 2-index overlaps");
  const Layout* ex = existing(sgv);
  const int N = wfn.L.rank;
  execute(make_key(OP_SIGMA_OVL, 0, {&lsov, &rsov, &wfn}, &sgv), [&] {
    PlanBuilder B(MPSXX_FP64);
    const int sl = B.input(lsov.L, 0), sr = B.input(rsov.L, 1), sw = B.input(wfn.L, 2);
    // compute_sigma_vector.hpp:32-33, 80-81, 106-107: Gemm(N,N,lopr,wfn0), Gemm(N,T,scr1,ropr)
    const int s1 = B.contract(sl, false, {1}, sw, false, {0});
    B.contract_out(s1, false, {N - 1}, sr, false, {1}, ex);
    return B.finish();
  }, OUT_ACCUMULATE, {&lsov, &rsov, &wfn}, sgv, s);
}

void op_sigma_conj(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& wfn, Tensor& sgv, hipStream_t s) {
  for (Tensor* t : {&mpo, &lopr, &ropr, &wfn}) require(*t, "compute_sigma_vector_conj");
  // compute_sigma_vector.hpp:45-50.  The output is the conjugated view of sgv0: its quantum numbers are checked
  // negated, and an empty sgv0 receives the data with the conjugate's quantum numbers flipped back.
  Layout exneg;
  const Layout* ex = nullptr;
  if (!sgv.empty) {
    exneg = sgv.L;
    exneg.q = -exneg.q;
    for (auto& v : exneg.qs)
      for (QN& x : v) x = -x;
    exneg.finalize();
    ex = &exneg;
  }
  auto key = make_key(OP_SIGMA_CONJ, 0, {&mpo, &lopr, &ropr, &wfn}, &sgv);
  const std::vector<const Layout*> ls = key_layouts({&mpo, &lopr, &ropr, &wfn}, &sgv);
  CachedPlan* cp = cache().find(key, ls);
  if (!cp) {
    PlanBuilder B(MPSXX_FP64);
    const int sm = B.input(mpo.L, 0), sl = B.input(lopr.L, 1), sr = B.input(ropr.L, 2), sw = B.input(wfn.L, 3);
    // Gemm(Trans, ConjTrans, lopr, wfn0): A contracts its first index, B (conjugated) its last
    const int s1 = B.contract(sl, false, {0}, sw, true, {2});
    const int s2 = B.contract(s1, false, {0, 2}, sm, false, {0, 1});
    B.contract_out(s2, false, {1, 3}, sr, false, {0, 1}, ex);
    auto np = std::make_unique<CachedPlan>();
    np->plan = B.finish();
    cp = cache().insert(key, std::move(np), ls);
  }
  Plan& P = cp->plan;
  Layout outL = P.out;
  outL.q = -outL.q;
  for (auto& v : outL.qs)
    for (QN& x : v) x = -x;
  outL.finalize();
  if (sgv.empty) sgv.set_layout(outL, s);
  else if (!sgv.L.same_structure(outL)) sgv.relayout(outL, s);
  run_plan(P, {mpo.ptr(), lopr.ptr(), ropr.ptr(), wfn.ptr()}, sgv.ptr(), s);
}

// ---------------------------------------------------------------------------------------------------------------------
// diagonal (src/compute_diagonal_elements.hpp)

void op_diag_twosite(Tensor& lmpo, Tensor& rmpo, Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s) {
  for (Tensor* t : {&lmpo, &rmpo, &lopr, &ropr}) require(*t, "compute_diagonal_elements");
  if (diag.empty) throw Error(MPSXX_ERR_ARG, "compute_diagonal_elements: diag must carry the wavefunction's q/qshape");
  Layout target = diag.L;
  execute(make_key(OP_DIAG2, 0, {&lmpo, &rmpo, &lopr, &ropr}, &diag), [&] {
    PlanBuilder B(MPSXX_FP64);
    const int a = B.input(lmpo.L, 0), b = B.input(rmpo.L, 1), c = B.input(lopr.L, 2), d = B.input(ropr.L, 3);
    // compute_diagonal_elements.hpp:47-60
    const int lm = B.tie(a, {1, 2}), rm = B.tie(b, {1, 2}), lo = B.tie(c, {0, 2}), ro = B.tie(d, {0, 2});
    const int s1 = B.contract(lo, false, {1}, lm, false, {0});
    const int s2 = B.contract(s1, false, {2}, rm, false, {0});
    const int s3 = B.contract(s2, false, {3}, ro, false, {1});
    B.upcast_out(s3, target);
    return B.finish();
  }, OUT_OVERWRITE, {&lmpo, &rmpo, &lopr, &ropr}, diag, s);
}

void op_diag_onesite(Tensor& mpo, Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s) {
  for (Tensor* t : {&mpo, &lopr, &ropr}) require(*t, "compute_diagonal_elements");
  if (diag.empty) throw Error(MPSXX_ERR_ARG, "compute_diagonal_elements: diag must carry the wavefunction's q/qshape");
  Layout target = diag.L;
  execute(make_key(OP_DIAG1, 0, {&mpo, &lopr, &ropr}, &diag), [&] {
    PlanBuilder B(MPSXX_FP64);
    const int a = B.input(mpo.L, 0), c = B.input(lopr.L, 1), d = B.input(ropr.L, 2);
    // compute_diagonal_elements.hpp:21-31
    const int m = B.tie(a, {1, 2}), lo = B.tie(c, {0, 2}), ro = B.tie(d, {0, 2});
    const int s1 = B.contract(lo, false, {1}, m, false, {0});
    const int s2 = B.contract(s1, false, {2}, ro, false, {1});
    B.upcast_out(s2, target);
    return B.finish();
  }, OUT_OVERWRITE, {&mpo, &lopr, &ropr}, diag, s);
}

void op_diag_merged(Tensor& lopr, Tensor& ropr, Tensor& diag, hipStream_t s) {
  for (Tensor* t : {&lopr, &ropr}) require(*t, "compute_diagonal_elements");
  if (diag.empty) throw Error(MPSXX_ERR_ARG, "compute_diagonal_elements: diag must carry the wavefunction's q/qshape");
  Layout target = diag.L;
  execute(make_key(OP_DIAGM, 0, {&lopr, &ropr}, &diag), [&] {
    PlanBuilder B(MPSXX_FP64);
    const int c = B.input(lopr.L, 0), d = B.input(ropr.L, 1);
    // compute_diagonal_elements.hpp:72-79: Gemm(N,T,lopr_diag,ropr_diag); only the blocks the up-casting copy keeps
    // (those the wavefunction's quantum numbers allow) are multiplied
    const int lo = B.tie(c, {0, 2}), ro = B.tie(d, {0, 2});
    const int s1 = B.contract(lo, false, {1}, ro, false, {1}, &target);
    B.upcast_out(s1, target);
    return B.finish();
  }, OUT_OVERWRITE, {&lopr, &ropr}, diag, s);
}

// ---------------------------------------------------------------------------------------------------------------------
// renormalization (src/renormalize.hpp)

void op_renormalize(bool fwd, Tensor& mpo, Tensor& opr0, Tensor& bra, Tensor& ket, Tensor& opr1, int prec,
                    hipStream_t s) {
  for (Tensor* t : {&mpo, &opr0, &bra, &ket}) require(*t, "renormalize");
  const Layout* ex = existing(opr1);
  execute(make_key(fwd ? OP_RENORM_F : OP_RENORM_B, prec, {&mpo, &opr0, &bra, &ket}, &opr1), [&] {
    PlanBuilder B(prec);
    const int sm = B.input(mpo.L, 0), so = B.input(opr0.L, 1), sb = B.input(bra.L, 2), sk = B.input(ket.L, 3);
    if (fwd) {  // renormalize.hpp:20-23
      const int s1 = B.contract(so, false, {0}, sb, true, {0});
      const int s2 = B.contract(s1, false, {0, 2}, sm, false, {0, 1});
      B.contract_out(s2, false, {0, 2}, sk, false, {0, 1}, ex);
    } else {    // renormalize.hpp:27-30
      const int s1 = B.contract(sb, true, {2}, so, false, {0});
      const int s2 = B.contract(s1, false, {1, 2}, sm, false, {1, 3});
      B.contract_out(s2, false, {3, 1}, sk, false, {1, 2}, ex);
    }
    return B.finish();
  }, OUT_ACCUMULATE, {&mpo, &opr0, &bra, &ket}, opr1, s);
}

void op_renormalize_overlap(bool fwd, Tensor& sov0, Tensor& bra, Tensor& ket, Tensor& sov1, hipStream_t s) {
  for (Tensor* t : {&sov0, &bra, &ket}) require(*t, "renormalize");
  const Layout* ex = existing(sov1);
  const int N = bra.L.rank;
  if (N != 2 && N != 3) throw Error(MPSXX_ERR_ARG, "renormalize(overlap): bra must have 2 or 3 indices");
  execute(make_key(fwd ? OP_RENORM_OVL_F : OP_RENORM_OVL_B, 0, {&sov0, &bra, &ket}, &sov1), [&] {
    PlanBuilder B(MPSXX_FP64);
    const int so = B.input(sov0.L, 0), sb = B.input(bra.L, 1), sk = B.input(ket.L, 2);
    if (fwd) {
      // renormalize.hpp:45-46 / 87-88: Gemm(T,N,sov0,bra*), Gemm(T,N,scr1,ket0)
      const int s1 = B.contract(so, false, {0}, sb, true, {0});
      std::vector<int> c1, c2;
      for (int k = 0; k < N - 1; ++k) { c1.push_back(k); c2.push_back(k); }
      B.contract_out(s1, false, c1, sk, false, c2, ex);
    } else {
      // renormalize.hpp:50-51 / 92-93: Gemm(N,N,bra*,sov0), Gemm(N,T,scr1,ket0)
      const int s1 = B.contract(sb, true, {N - 1}, so, false, {0});
      std::vector<int> c1, c2;
      for (int k = 1; k < N; ++k) { c1.push_back(k); c2.push_back(k); }
      B.contract_out(s1, false, c1, sk, false, c2, ex);
    }
    return B.finish();
  }, OUT_ACCUMULATE, {&sov0, &bra, &ket}, sov1, s);
}

void op_renormalize_merged(bool fwd, Tensor& opr0, Tensor& bra, Tensor& ket, Tensor& opr1, int prec,
                           hipStream_t s) {
  for (Tensor* t : {&opr0, &bra, &ket}) require(*t, "renormalize");
  if (bra.L.rank != 2 || ket.L.rank != 2 || opr0.L.rank != 3)
    throw Error(MPSXX_ERR_ARG, "renormalize(merged): expects a 3-index operator and 2-index bra/ket");
  const Layout* ex = existing(opr1);
  execute(make_key(fwd ? OP_RENORM_M_F : OP_RENORM_M_B, prec, {&opr0, &bra, &ket}, &opr1), [&] {
    PlanBuilder B(prec);
    const int so = B.input(opr0.L, 0), sb = B.input(bra.L, 1), sk = B.input(ket.L, 2);
    if (fwd) {  // renormalize.hpp:66-67: Gemm(T,N,bra*,opr0), Gemm(N,N,scr1,ket0)
      const int s1 = B.contract(sb, true, {0}, so, false, {0});
      B.contract_out(s1, false, {2}, sk, false, {0}, ex);
    } else {    // renormalize.hpp:71-72: Gemm(N,N,bra*,opr0), Gemm(N,T,scr1,ket0)
      const int s1 = B.contract(sb, true, {1}, so, false, {0});
      B.contract_out(s1, false, {2}, sk, false, {1}, ex);
    }
    return B.finish();
  }, OUT_ACCUMULATE, {&opr0, &bra, &ket}, opr1, s);
}

}  // namespace mpx
