// Quantum-number merging of block-sparse indices and the merged block operators.
//
// MergeInfo follows btas::QSTmergeInfo (include/btas/QSPARSE/QSTmergeInfo.h:44-95): the merged indices are packed
// row-major (first index slowest), each distinct packed quantum number becomes one merged sector in order of first
// appearance, and the packed sectors of one merged sector are stacked in packed order.  QSTmerge
// (QSTmerge.h:31-140) creates a merged block when at least one of its sub-blocks is stored and zero-fills the rest.
#include <algorithm>
#include <array>
#include <list>
#include <memory>
#include <mutex>
#include <cstdlib>
#include <map>
#include <unordered_map>

#include "mpsxx_amd.h"
#include "ops.hpp"

namespace mpx {


MergeInfo merge_info(const std::vector<std::vector<QN>>& qs, const std::vector<std::vector<int>>& ds) {
  MergeInfo M;
  std::vector<QN> qp{QN()};
  std::vector<int> dp{1};
  for (size_t i = 0; i < qs.size(); ++i) {
    if (qs[i].size() != ds[i].size()) throw Error(MPSXX_ERR_ARG, "QSTmergeInfo: qshape and dshape have different block size");
    std::vector<QN> q2;
    std::vector<int> d2;
    for (size_t a = 0; a < qp.size(); ++a)
      for (size_t b = 0; b < qs[i].size(); ++b) {
        q2.push_back(qp[a] + qs[i][b]);
        d2.push_back(dp[a] * ds[i][b]);
      }
    qp.swap(q2);
    dp.swap(d2);
    M.radix.push_back((int)qs[i].size());
  }
  std::map<std::pair<int, int>, int> idx;
  for (size_t k = 0; k < qp.size(); ++k) {
    auto key = std::make_pair(qp[k].p, qp[k].s);
    auto it = idx.find(key);
    int j;
    if (it == idx.end()) {
      j = (int)M.qm.size();
      idx[key] = j;
      M.qm.push_back(qp[k]);
      M.dm.push_back(0);
    } else {
      j = it->second;
    }
    M.p2m.push_back(j);
    M.poff.push_back(M.dm[j]);
    M.pdim.push_back(dp[k]);
    M.dm[j] += dp[k];
  }
  return M;
}

namespace {

int packed_index(const MergeInfo& M, const int32_t* idx) {
  int p = 0;
  for (size_t i = 0; i < M.radix.size(); ++i) p = p * M.radix[i] + idx[i];
  return p;
}

// Sub-rectangle copy segments: a source block viewed as (R x C) row-major placed at (ro, co) of a target block with
// row length ldc.
void rect_segments(int64_t src_off, int64_t R, int64_t C, int64_t dst_off, int64_t ro, int64_t co, int64_t ldc,
                   std::vector<CopySeg>& segs) {
  if (R == 0 || C == 0) return;
  if (co == 0 && C == ldc) {
    segs.push_back({dst_off + ro * ldc, src_off, (int32_t)(R * C), 1});
    return;
  }
  for (int64_t r = 0; r < R; ++r) segs.push_back({dst_off + (ro + r) * ldc + co, src_off + r * C, (int32_t)C, 1});
}

}  // namespace

// Structure and copy segments of QSTmerge(rows_info, a, cols_info): merge the first nr and the last nc indices of X
// (nr or nc may be 1, which leaves that side unchanged).
Layout merge_layout(const Layout& X, int nr, int nc, std::vector<CopySeg>& segs) {
  const int N = X.rank;
  if (nr < 1 || nc < 1 || nr + nc > N) throw Error(MPSXX_ERR_ARG, "QSTmerge: bad index counts");
  std::vector<std::vector<QN>> rq(X.qs.begin(), X.qs.begin() + nr), cq(X.qs.end() - nc, X.qs.end());
  std::vector<std::vector<int>> rd(X.ds.begin(), X.ds.begin() + nr), cd(X.ds.end() - nc, X.ds.end());
  // a single index is left as is (no renumbering of sectors that share a quantum number)
  auto identity = [](const std::vector<QN>& q, const std::vector<int>& d) {
    MergeInfo M;
    M.qm = q;
    M.dm = d;
    M.radix = {(int)q.size()};
    for (size_t k = 0; k < q.size(); ++k) {
      M.p2m.push_back((int)k);
      M.poff.push_back(0);
      M.pdim.push_back(d[k]);
    }
    return M;
  };
  MergeInfo MR = nr > 1 ? merge_info(rq, rd) : identity(rq[0], rd[0]);
  MergeInfo MC = nc > 1 ? merge_info(cq, cd) : identity(cq[0], cd[0]);
  std::vector<std::vector<QN>> qs;
  std::vector<std::vector<int>> ds;
  qs.push_back(MR.qm);
  ds.push_back(MR.dm);
  for (int i = nr; i < N - nc; ++i) { qs.push_back(X.qs[i]); ds.push_back(X.ds[i]); }
  qs.push_back(MC.qm);
  ds.push_back(MC.dm);
  Layout Y;
  Y.set_shape((int)qs.size(), X.qsym, X.q, qs, ds);
  std::vector<int32_t> yi(Y.rank);
  auto target = [&](const int32_t* xi) {
    const int pr = packed_index(MR, xi), pc = packed_index(MC, xi + (N - nc));
    yi[0] = MR.p2m[pr];
    for (int i = nr; i < N - nc; ++i) yi[1 + i - nr] = xi[i];
    yi[Y.rank - 1] = MC.p2m[pc];
  };
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    target(X.index(b));
    if (Y.allowed(yi.data())) Y.add_block(yi.data());
  }
  Y.finalize();
  segs.clear();
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    const int32_t* xi = X.index(b);
    target(xi);
    const int64_t yb = Y.find(Y.compute_tag(yi.data()));
    if (yb < 0) continue;
    const int pr = packed_index(MR, xi), pc = packed_index(MC, xi + (N - nc));
    int64_t mid = 1;
    for (int i = nr; i < N - nc; ++i) mid *= X.ds[i][xi[i]];
    const int64_t dr = MR.pdim[pr], dc = MC.pdim[pc];
    const int64_t Dc = MC.dm[MC.p2m[pc]];
    // target element (r, m, c) at ((ro + r) * mid + m) * Dc + co + c; source (r, m, c) at (r * mid + m) * dc + c
    rect_segments(X.off[b], dr * mid, dc, Y.off[yb], (int64_t)MR.poff[pr] * mid, MC.poff[pc], Dc, segs);
  }
  return Y;
}

namespace {

bool scalar_blocks(const Layout& W) {
  for (int64_t b = 0; b < W.nblocks(); ++b)
    if (W.bsize[b] != 1) return false;
  return W.nblocks() > 0;
}

// Merged operator from a scalar-block MPO without materialising the contraction's pair list.  The environment's blocks
// are grouped into slices (bra sector, ket sector) over the contracted operator index w, the MPO entries into columns
// (the three other MPO indices) listing (w, entry) in ascending w.  A sub-block (slice, column) of the contraction
// exists iff the two w sets intersect (Contract's output rule, STBLAS.h:331-350); all w of a slice (of a column) carry
// one quantum number, so only slices and columns of the same class are intersected, as bit sets.  The kernel sums
// each sub-block over the column's entries in ascending w, the reference's accumulation order, and writes it straight
// to its place in the merged block (QSTmerge placement), so the contraction intermediate is never stored.
}  // namespace

bool merged_operator_plan(bool fwd, const Layout& W, const Layout& E, MergedOpPlan& P) {
  if (W.rank != 4 || E.rank != 3 || !scalar_blocks(W)) return false;
  const int wc = fwd ? 0 : 3;  // contracted MPO index (E's is 1)
  const int nw = E.nsec(1);
  if (W.nsec(wc) != nw) throw Error(MPSXX_ERR_QN, "Contract: contracted sector counts differ");
  for (int k = 0; k < nw; ++k) {
    if (E.qs[1][k] != -W.qs[wc][k]) throw Error(MPSXX_ERR_QN, "contraction of quantum numbers failed");
    if (E.ds[1][k] != W.ds[wc][k]) throw Error(MPSXX_ERR_SHAPE, "Contract: contracted dense shapes differ");
    if (E.ds[1][k] != 1) return false;
  }
  // the three free MPO indices of a column: forward (nb, nk, w'), backward (w, nb, nk)
  const int f0 = fwd ? 1 : 0, f1 = fwd ? 2 : 1, f2 = fwd ? 3 : 2;
  const int n1 = E.nsec(0), n2 = E.nsec(2);
  // slices
  std::vector<int32_t> slice_of((size_t)n1 * n2, -1);
  std::vector<int32_t> sl_e0, sl_e2;
  std::vector<QN> sl_q;
  for (int64_t b = 0; b < E.nblocks(); ++b) {
    const int32_t* e = E.index(b);
    int32_t& id = slice_of[(size_t)e[0] * n2 + e[2]];
    if (id < 0) {
      id = (int32_t)sl_e0.size();
      sl_e0.push_back(e[0]);
      sl_e2.push_back(e[2]);
      sl_q.push_back(E.qs[1][e[1]]);
    }
  }
  const int nsl = (int)sl_e0.size();
  const int words = (nw + 63) / 64;
  std::vector<int32_t> slice_blk((size_t)std::max(nsl, 1) * nw, -1);
  std::vector<uint64_t> sl_bits((size_t)std::max(nsl, 1) * words, 0);
  for (int64_t b = 0; b < E.nblocks(); ++b) {
    const int32_t* e = E.index(b);
    const int32_t id = slice_of[(size_t)e[0] * n2 + e[2]];
    slice_blk[(size_t)id * nw + e[1]] = (int32_t)b;
    sl_bits[(size_t)id * words + e[1] / 64] |= uint64_t(1) << (e[1] % 64);
  }
  // columns, entries in MPO tag order (ascending w within a column: w is the slowest index forward, the fastest
  // backward, and the column key leaves it out)
  const int64_t nc1 = W.nsec(f1), nc2 = W.nsec(f2);
  std::unordered_map<int64_t, int32_t> col_of;
  std::vector<int64_t> col_key;
  std::vector<int32_t> cnt;
  std::vector<QN> col_q;
  for (int64_t b = 0; b < W.nblocks(); ++b) {
    const int32_t* x = W.index(b);
    const int64_t key = ((int64_t)x[f0] * nc1 + x[f1]) * nc2 + x[f2];
    auto it = col_of.find(key);
    if (it == col_of.end()) {
      it = col_of.emplace(key, (int32_t)col_key.size()).first;
      col_key.push_back(key);
      cnt.push_back(0);
      col_q.push_back(-W.qs[wc][x[wc]]);
    }
    ++cnt[it->second];
  }
  const int ncol = (int)col_key.size();
  std::vector<int32_t> col_begin(ncol + 1, 0);
  for (int c = 0; c < ncol; ++c) col_begin[c + 1] = col_begin[c] + cnt[c];
  std::vector<int32_t> ent_w(W.nblocks());
  std::vector<int64_t> ent_off(W.nblocks());
  std::vector<uint64_t> col_bits((size_t)std::max(ncol, 1) * words, 0);
  {
    std::vector<int32_t> fill(col_begin.begin(), col_begin.end() - 1);
    for (int64_t b = 0; b < W.nblocks(); ++b) {
      const int32_t* x = W.index(b);
      const int32_t c = col_of[((int64_t)x[f0] * nc1 + x[f1]) * nc2 + x[f2]];
      ent_w[fill[c]] = x[wc];
      ent_off[fill[c]] = W.off[b];
      ++fill[c];
      col_bits[(size_t)c * words + x[wc] / 64] |= uint64_t(1) << (x[wc] % 64);
    }
  }
  // classes by the contracted quantum number
  std::map<std::pair<int, int>, std::pair<std::vector<int32_t>, std::vector<int32_t>>> cls;
  for (int i = 0; i < nsl; ++i) cls[{sl_q[i].p, sl_q[i].s}].first.push_back(i);
  for (int c = 0; c < ncol; ++c) cls[{col_q[c].p, col_q[c].s}].second.push_back(c);
  std::vector<std::pair<const std::vector<int32_t>*, int32_t>> work;  // (columns of the class, slice)
  for (auto& kv : cls)
    for (int32_t i : kv.second.first) work.push_back({&kv.second.second, i});
  // merged row / column sectors
  MergeInfo MR, MC;
  if (fwd) {
    MR = merge_info({E.qs[0], W.qs[1]}, {E.ds[0], W.ds[1]});
    MC = merge_info({E.qs[2], W.qs[2]}, {E.ds[2], W.ds[2]});
  } else {
    MR = merge_info({W.qs[1], E.qs[0]}, {W.ds[1], E.ds[0]});
    MC = merge_info({W.qs[2], E.qs[2]}, {W.ds[2], E.ds[2]});
  }
  // (slice, column) pairs with intersecting w sets, on several host threads in contiguous ranges of the work list
  const int nth = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)host_threads(), (int64_t)work.size() / 64));
  std::vector<std::vector<std::pair<int32_t, int32_t>>> part(nth);
  parallel_for(nth, [&](int t) {
    const size_t lo = work.size() * t / nth, hi = work.size() * (t + 1) / nth;
    for (size_t k = lo; k < hi; ++k) {
      const int32_t i = work[k].second;
      const uint64_t* a = &sl_bits[(size_t)i * words];
      for (int32_t c : *work[k].first) {
        const uint64_t* bb = &col_bits[(size_t)c * words];
        for (int x = 0; x < words; ++x)
          if (a[x] & bb[x]) {
            part[t].push_back({i, c});
            break;
          }
      }
    }
  });
  Layout Y;
  Y.set_shape(3, true, E.q + W.q, {MR.qm, W.qs[fwd ? 3 : 0], MC.qm}, {MR.dm, W.ds[fwd ? 3 : 0], MC.dm});
  struct Place { int32_t slice, col, pr, pc, w; };
  std::vector<Place> pl;
  for (auto& v : part)
    for (auto& sc : v) {
      const int64_t key = col_key[sc.second];
      const int32_t x0 = (int32_t)(key / (nc1 * nc2)), x1 = (int32_t)(key / nc2 % nc1), x2 = (int32_t)(key % nc2);
      Place p;
      p.slice = sc.first;
      p.col = sc.second;
      if (fwd) {  // column (nb, nk, w'): rows (lb, nb), cols (lk, nk)
        p.pr = sl_e0[sc.first] * MR.radix[1] + x0;
        p.pc = sl_e2[sc.first] * MC.radix[1] + x1;
        p.w = x2;
      } else {    // column (w, nb, nk): rows (nb, rb), cols (nk, rk)
        p.pr = x1 * MR.radix[1] + sl_e0[sc.first];
        p.pc = x2 * MC.radix[1] + sl_e2[sc.first];
        p.w = x0;
      }
      int32_t yi[3] = {MR.p2m[p.pr], p.w, MC.p2m[p.pc]};
      if (!Y.allowed(yi)) continue;
      Y.add_block(yi);
      pl.push_back(p);
    }
  Y.finalize();
  std::vector<MergedOpTask>& tasks = P.tasks;
  std::vector<Chunk>& chunks = P.chunks;
  tasks.clear();
  chunks.clear();
  tasks.reserve(pl.size());
  for (const Place& p : pl) {
    int32_t yi[3] = {MR.p2m[p.pr], p.w, MC.p2m[p.pc]};
    const int64_t yb = Y.find(Y.compute_tag(yi));
    MergedOpTask t;
    t.rows = E.ds[0][sl_e0[p.slice]];
    t.cols = E.ds[2][sl_e2[p.slice]];
    t.ld = MC.dm[yi[2]];
    t.y_off = Y.off[yb] + (int64_t)MR.poff[p.pr] * t.ld + MC.poff[p.pc];
    t.slice = p.slice;
    t.col = p.col;
    t.pad = 0;
    for (int st = 0; st < t.rows * t.cols; st += kAxpyChunk) chunks.push_back({(int32_t)tasks.size(), st});
    tasks.push_back(t);
  }
  P.Y = std::move(Y);
  P.nw = nw;
  P.slice_blk = std::move(slice_blk);
  P.col_begin = std::move(col_begin);
  P.ent_w = std::move(ent_w);
  P.ent_off = std::move(ent_off);
  P.env_off = E.off;
  return true;
}

namespace {

// Device copy of a scalar-MPO merged-operator plan.  Its structure depends only on the direction and the MPO and
// environment layouts, which repeat from one sweep to the next once the bond dimensions settle, so the plans are kept
// in an LRU cache of bounded device size (1% of device memory, at most kMergedKeep plans).
struct MergedOpDevice {
  Layout mpo, opr;  // the operand layouts the plan was built for (a hit needs these structures, not only the hashes)
  Layout Y;
  int nw = 0, nchunks = 0;
  DevBuf tasks, chunks, slice_blk, env_off, col_begin, ent_w, ent_off;
  size_t bytes() const {
    return tasks.bytes + chunks.bytes + slice_blk.bytes + env_off.bytes + col_begin.bytes + ent_w.bytes + ent_off.bytes;
  }
};

constexpr size_t kMergedKeep = 256;

struct MergedOpCache {
  typedef std::array<uint64_t, 3> Key;
  typedef std::list<std::pair<Key, std::shared_ptr<MergedOpDevice>>> List;
  std::mutex mu;
  List lru;  // most recently used first
  std::map<Key, List::iterator> idx;
  size_t bytes = 0;
  std::shared_ptr<MergedOpDevice> find(const Key& k, const Layout& mpo, const Layout& opr) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = idx.find(k);
    if (it == idx.end()) return nullptr;
    const MergedOpDevice& d = *it->second->second;
    if (!d.mpo.same_structure(mpo) || !d.opr.same_structure(opr)) return nullptr;  // hash collision: rebuild
    lru.splice(lru.begin(), lru, it->second);
    return it->second->second;
  }
  // insert, replacing a plan of other structures under the same key (callers hold their plans by shared_ptr)
  void insert(const Key& k, std::shared_ptr<MergedOpDevice> v) {
    static const size_t budget = device_memory_fraction(0.01, 2.0);
    std::lock_guard<std::mutex> lk(mu);
    auto old = idx.find(k);
    if (old != idx.end()) {
      bytes -= old->second->second->bytes();
      lru.erase(old->second);
      idx.erase(old);
    }
    bytes += v->bytes();
    lru.emplace_front(k, std::move(v));
    idx[k] = lru.begin();
    while (lru.size() > 1 && (bytes > budget || lru.size() > kMergedKeep)) {
      bytes -= lru.back().second->bytes();
      idx.erase(lru.back().first);
      lru.pop_back();
    }
  }
  void clear() {
    std::lock_guard<std::mutex> lk(mu);
    idx.clear();
    lru.clear();
    bytes = 0;
  }
};

MergedOpCache& merged_cache() {
  static MergedOpCache c;
  return c;
}

bool merged_operator_scalar(bool fwd, Tensor& xmpo, Tensor& xopr, Tensor& mopr, hipStream_t s) {
  const MergedOpCache::Key key{fwd ? 1u : 0u, xmpo.L.hash, xopr.L.hash};
  std::shared_ptr<MergedOpDevice> D = merged_cache().find(key, xmpo.L, xopr.L);
  if (!D) {
    MergedOpPlan P;
    if (!merged_operator_plan(fwd, xmpo.L, xopr.L, P)) return false;
    D = std::make_shared<MergedOpDevice>();
    D->mpo = xmpo.L;
    D->opr = xopr.L;
    D->Y = std::move(P.Y);
    D->nw = P.nw;
    D->nchunks = (int)P.chunks.size();
    if (D->nchunks > 0) {
      D->tasks = upload_vec(P.tasks);
      D->chunks = upload_vec(P.chunks);
      D->slice_blk = upload_vec(P.slice_blk);
      D->env_off = upload_vec(P.env_off);
      D->col_begin = upload_vec(P.col_begin);
      D->ent_w = upload_vec(P.ent_w);
      D->ent_off = upload_vec(P.ent_off);
    }
    merged_cache().insert(key, D);
  }
  mopr.set_layout(D->Y, s);  // zero: sub-blocks no pair reaches stay zero, as QSTmerge leaves them
  if (D->nchunks > 0) {
    launch_merged_operator(D->chunks.as<Chunk>(), D->nchunks, D->tasks.as<MergedOpTask>(), D->slice_blk.as<int32_t>(),
                           D->nw, D->env_off.as<int64_t>(), D->col_begin.as<int32_t>(), D->ent_w.as<int32_t>(),
                           D->ent_off.as<int64_t>(), xopr.ptr(), xmpo.ptr(), mopr.ptr(), s);
    hip_check(hipStreamSynchronize(s), "merged operator");
  }
  return true;
}

}  // namespace

void merged_plan_cache_clear() { merged_cache().clear(); }

void clear_merged_operator_cache() { merged_cache().clear(); }

// Merged block operator (reference src/optimize.hpp:273-368), one operator group:
//   forward : L~(Lb, w', Lk) = merge_{(lb,nb)} merge_{(lk,nk)} sum_w L(lb,w,lk) W(w,nb,nk,w')
//   backward: R~(Rb, w, Rk)  = merge_{(nb,rb)} merge_{(nk,rk)} sum_w' W(w,nb,nk,w') R(rb,w',rk)
// The contraction is the reference's Contract-by-symbols (one K=1 coefficient step), the permutation and both merges
// become one placement copy.
void op_merged_operator(bool fwd, Tensor& xmpo, Tensor& xopr, Tensor& mopr, hipStream_t s) {
  if (xmpo.empty || xopr.empty) throw Error(MPSXX_ERR_ARG, "compute_merged_operator: empty input");
  if (merged_operator_scalar(fwd, xmpo, xopr, mopr, s)) return;
  const Layout& W = xmpo.L;
  const Layout& E = xopr.L;
  // symbolic contraction: forward Contract(L, {1}, W, {0}) -> (lb, lk, nb, nk, w');
  //                       backward Contract(W, {3}, R, {1}) -> (w, nb, nk, rb, rk)
  View vw{&W, false}, ve{&E, false};
  ContractSym sym = fwd ? contract_symbolic(ve, {1}, vw, {0}, nullptr) : contract_symbolic(vw, {3}, ve, {1}, nullptr);
  const Layout& X = sym.C;
  MergeInfo MR, MC;
  if (fwd) {
    MR = merge_info({E.qs[0], W.qs[1]}, {E.ds[0], W.ds[1]});
    MC = merge_info({E.qs[2], W.qs[2]}, {E.ds[2], W.ds[2]});
  } else {
    MR = merge_info({W.qs[1], E.qs[0]}, {W.ds[1], E.ds[0]});
    MC = merge_info({W.qs[2], E.qs[2]}, {W.ds[2], E.ds[2]});
  }
  Layout Y;
  Y.set_shape(3, true, X.q, {MR.qm, fwd ? W.qs[3] : W.qs[0], MC.qm}, {MR.dm, fwd ? W.ds[3] : W.ds[0], MC.dm});
  int32_t yi[3];
  // (row sector index pair, col sector index pair, operator) of an X block
  auto decode = [&](const int32_t* xi, int& pr, int& pc, int& w, int64_t& R, int64_t& C) {
    if (fwd) {  // (lb, lk, nb, nk, w')
      pr = xi[0] * MR.radix[1] + xi[2];
      pc = xi[1] * MC.radix[1] + xi[3];
      w = xi[4];
      R = X.ds[0][xi[0]];
      C = X.ds[1][xi[1]];
    } else {    // (w, nb, nk, rb, rk)
      pr = xi[1] * MR.radix[1] + xi[3];
      pc = xi[2] * MC.radix[1] + xi[4];
      w = xi[0];
      R = X.ds[3][xi[3]];
      C = X.ds[4][xi[4]];
    }
  };
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    int pr, pc, w;
    int64_t R, C;
    decode(X.index(b), pr, pc, w, R, C);
    yi[0] = MR.p2m[pr];
    yi[1] = w;
    yi[2] = MC.p2m[pc];
    if (Y.allowed(yi)) Y.add_block(yi);
  }
  Y.finalize();
  std::vector<CopySeg> segs;
  for (int64_t b = 0; b < X.nblocks(); ++b) {
    int pr, pc, w;
    int64_t R, C;
    decode(X.index(b), pr, pc, w, R, C);
    yi[0] = MR.p2m[pr];
    yi[1] = w;
    yi[2] = MC.p2m[pc];
    const int64_t yb = Y.find(Y.compute_tag(yi));
    if (yb < 0) continue;
    rect_segments(X.off[b], R, C, Y.off[yb], MR.poff[pr], MC.poff[pc], MC.dm[MC.p2m[pc]], segs);
  }
  // plan: [0] mpo, [1] operator, [2] contraction intermediate, [3] output
  Plan P;
  P.slot_layout = {W, E, X, Y};
  P.slot_input = {0, 1, -1, -2};
  P.slot_prec = {0, 0, 0, 0};
  P.slot_buf.resize(4);
  P.slot_buf[2].alloc(std::max<int64_t>(X.nelem, 1) * sizeof(double));
  P.out = Y;
  P.steps.push_back(fwd ? make_contract_step(sym, 1, 0, 2, 0, 1.0) : make_contract_step(sym, 0, 1, 2, 0, 1.0));
  P.steps.push_back(make_copy_step(segs, 2, 3, 1.0, 0.0));
  P.flops = sym.flops;
  mopr.set_layout(Y, s);
  run_plan(P, {xmpo.ptr(), xopr.ptr()}, mopr.ptr(), s);
  hip_check(hipStreamSynchronize(s), "merged operator");
}

// Merged overlap (reference src/optimize.hpp:326-328, 363-365): S (x) I on the site, merged like the operators.
//   forward : S~(Lb, Lk) = merge_{(lb,n)} merge_{(lk,n')} S(lb,lk) I(n,n')
//   backward: S~(Rb, Rk) = merge_{(n,rb)} merge_{(n',rk)} I(n,n') S(rb,rk)
// I is the site identity (the QSTArray filled with 1.0 on its allowed blocks, all diagonal for distinct site QNs).
void op_merged_overlap(bool fwd, Tensor& xmpo, Tensor& xsov, Tensor& msov, hipStream_t s) {
  if (xmpo.empty || xsov.empty) throw Error(MPSXX_ERR_ARG, "compute_merged_operator: empty input");
  const Layout& W = xmpo.L;
  const Layout& S = xsov.L;
  const std::vector<QN>& qn = W.qs[1];
  const std::vector<QN>& qk = W.qs[2];
  MergeInfo MR, MC;
  if (fwd) {
    MR = merge_info({S.qs[0], qn}, {S.ds[0], W.ds[1]});
    MC = merge_info({S.qs[1], qk}, {S.ds[1], W.ds[2]});
  } else {
    MR = merge_info({qn, S.qs[0]}, {W.ds[1], S.ds[0]});
    MC = merge_info({qk, S.qs[1]}, {W.ds[2], S.ds[1]});
  }
  // identity blocks (a, b) allowed: qn[a] + qk[b] == 0
  std::vector<std::pair<int, int>> eye;
  for (size_t a = 0; a < qn.size(); ++a)
    for (size_t b = 0; b < qk.size(); ++b)
      if (qn[a] + qk[b] == QN() && W.ds[1][a] == 1 && W.ds[2][b] == 1) eye.push_back({(int)a, (int)b});
  Layout Y;
  Y.set_shape(2, true, S.q, {MR.qm, MC.qm}, {MR.dm, MC.dm});
  struct Place { int64_t src, R, C; int pr, pc; };
  std::vector<Place> pl;
  for (int64_t b = 0; b < S.nblocks(); ++b) {
    const int32_t* si = S.index(b);
    for (auto& e : eye) {
      Place p;
      p.src = S.off[b];
      p.R = S.ds[0][si[0]];
      p.C = S.ds[1][si[1]];
      p.pr = fwd ? si[0] * MR.radix[1] + e.first : e.first * MR.radix[1] + si[0];
      p.pc = fwd ? si[1] * MC.radix[1] + e.second : e.second * MC.radix[1] + si[1];
      int32_t yi[2] = {MR.p2m[p.pr], MC.p2m[p.pc]};
      if (!Y.allowed(yi)) continue;
      Y.add_block(yi);
      pl.push_back(p);
    }
  }
  Y.finalize();
  std::vector<CopySeg> segs;
  for (const Place& p : pl) {
    int32_t yi[2] = {MR.p2m[p.pr], MC.p2m[p.pc]};
    const int64_t yb = Y.find(Y.compute_tag(yi));
    rect_segments(p.src, p.R, p.C, Y.off[yb], MR.poff[p.pr], MC.poff[p.pc], MC.dm[MC.p2m[p.pc]], segs);
  }
  Plan P;
  P.slot_layout = {S, Y};
  P.slot_input = {0, -2};
  P.slot_prec = {0, 0};
  P.slot_buf.resize(2);
  P.out = Y;
  P.steps.push_back(make_copy_step(segs, 0, 1, 1.0, 0.0));
  msov.set_layout(Y, s);
  run_plan(P, {xsov.ptr()}, msov.ptr(), s);
  hip_check(hipStreamSynchronize(s), "merged overlap");
}

// QSTmerge of the first nr and last nc indices (nr, nc >= 1).
void op_qst_merge(Tensor& x, int nr, int nc, Tensor& y, hipStream_t s) {
  if (x.empty) throw Error(MPSXX_ERR_ARG, "QSTmerge: empty input");
  std::vector<CopySeg> segs;
  Layout Y = merge_layout(x.L, nr, nc, segs);
  Plan P;
  P.slot_layout = {x.L, Y};
  P.slot_input = {0, -2};
  P.slot_prec = {0, 0};
  P.slot_buf.resize(2);
  P.out = Y;
  P.steps.push_back(make_copy_step(segs, 0, 1, 1.0, 0.0));
  y.set_layout(Y, s);
  run_plan(P, {x.ptr()}, y.ptr(), s);
  hip_check(hipStreamSynchronize(s), "merge");
}

}  // namespace mpx
