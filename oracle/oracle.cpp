// ORACLE -- test infrastructure only (CPU, FP64).  Imported by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg as the checker; never part of the product path.
//
// A clean-room CPU restatement of the reference's block-sparse algorithm (naokin/mpsxx), following the reference
// step by step so that block patterns and FP64 values can be compared with the device path:
//   * QSTArray model: blocks keyed by row-major tag, allowed iff sum of sector QNs == total
//     (include/btas/QSPARSE/QSTArray.h:34-36; include/btas/SPARSE/STArray.h:338-348), conjugate() negates QNs
//     (QSTArray.h:293-300);
//   * Contract: get_contract_jobs (include/btas/common/btas_contract_shape.h:164-204) -> explicit Permute copies of
//     A and B (QSTREINDEX.h:22-50, reindex.h:20-42) -> block-pair GEMM over (row, col) with at least one matching
//     contracted block, C block zero-initialised then accumulated pair by pair (STBLAS.h:298-359,
//     T_arguments.h:517-577), C's qshape/q from gemm_contract_qshape (btas_contract_qshape.h:60-108);
//   * Tie (STREINDEX.h:95-147), Copy with up-cast (STBLAS.h:48-60, 374-388), Dotc/Axpy/Scal/Nrm2 (QSTBLAS.h:40-93,
//     STBLAS.h:94-145);
//   * the hot-path compositions of src/compute_sigma_vector.hpp, src/compute_diagonal_elements.hpp and
//     src/renormalize.hpp, with the same index lists;
//   * Davidson (src/diagonalize.hpp:283-449) with a Jacobi eigen-solver standing in for LAPACKE_dsyev.
// Dense kernels are a plain cache-friendly row-major dgemm (the reference calls cblas_dgemm, blas/gemm_impl.h:63).
// OpenMP over output blocks mirrors parallel_call (T_arguments.h:584-599).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpsxx_amd.h"  // descriptor struct only

namespace orc {

struct QN {
  int p = 0, s = 0;
  QN() = default;
  QN(int a, int b) : p(a), s(b) {}
  QN operator+(const QN& o) const { return QN(p + o.p, s + o.s); }
  QN operator-() const { return QN(-p, -s); }
  bool operator==(const QN& o) const { return p == o.p && s == o.s; }
  bool operator!=(const QN& o) const { return !(*this == o); }
};

typedef std::vector<double> Block;

struct QST {
  int rank = 0;
  bool qsym = true;
  QN q;
  std::vector<std::vector<QN>> qs;
  std::vector<std::vector<int>> ds;
  std::map<long, std::shared_ptr<Block>> blocks;
  bool empty() const { return rank == 0; }

  long nsec(int i) const { return (long)ds[i].size(); }
  long tag(const std::vector<int>& idx) const {
    long t = 0;
    for (int i = 0; i < rank; ++i) t = t * nsec(i) + idx[i];
    return t;
  }
  std::vector<int> index(long t) const {
    std::vector<int> idx(rank);
    for (int i = rank - 1; i >= 0; --i) {
      idx[i] = (int)(t % nsec(i));
      t /= nsec(i);
    }
    return idx;
  }
  bool allowed(const std::vector<int>& idx) const {
    if (!qsym) return true;
    QN s;
    for (int i = 0; i < rank; ++i) s = s + qs[i][idx[i]];
    return s == q;
  }
  long bsize(const std::vector<int>& idx) const {
    long n = 1;
    for (int i = 0; i < rank; ++i) n *= ds[i][idx[i]];
    return n;
  }
  // STArray::reserve: existing block, or a new zero block when allowed, else nullptr (STArray.h:448-471)
  Block* reserve(long t) {
    auto it = blocks.find(t);
    std::vector<int> idx = index(t);
    if (!allowed(idx)) return nullptr;
    if (it != blocks.end()) return it->second.get();
    auto b = std::make_shared<Block>(bsize(idx), 0.0);
    blocks[t] = b;
    return b.get();
  }
  void set_shape(int r, bool qs_, QN q_, std::vector<std::vector<QN>> qv, std::vector<std::vector<int>> dv) {
    rank = r;
    qsym = qs_;
    q = q_;
    qs = qv;
    ds = dv;
    if (!qsym) {
      qs.assign(rank, {});
      for (int i = 0; i < rank; ++i) qs[i].assign(ds[i].size(), QN());
    }
    blocks.clear();
  }
};

QST conjugate(const QST& a) {
  QST c = a;  // shares block pointers
  c.q = -a.q;
  for (auto& v : c.qs)
    for (QN& x : v) x = -x;
  return c;
}

// dense row-major reindex: out index i takes input index perm[i]
QST permute(const QST& a, const std::vector<int>& perm) {
  QST c;
  std::vector<std::vector<QN>> qs(a.rank);
  std::vector<std::vector<int>> ds(a.rank);
  for (int i = 0; i < a.rank; ++i) {
    qs[i] = a.qs[perm[i]];
    ds[i] = a.ds[perm[i]];
  }
  c.set_shape(a.rank, a.qsym, a.q, qs, ds);
  for (const auto& kv : a.blocks) {
    std::vector<int> ia = a.index(kv.first);
    std::vector<int> ic(a.rank);
    for (int i = 0; i < a.rank; ++i) ic[i] = ia[perm[i]];
    std::vector<long> da(a.rank), sa(a.rank), dc(a.rank);
    long s = 1;
    for (int i = a.rank - 1; i >= 0; --i) {
      da[i] = a.ds[i][ia[i]];
      sa[i] = s;
      s *= da[i];
    }
    for (int i = 0; i < a.rank; ++i) dc[i] = da[perm[i]];
    auto out = std::make_shared<Block>(s, 0.0);
    std::vector<long> cnt(a.rank, 0);
    for (long e = 0; e < s; ++e) {
      long src = 0;
      for (int i = 0; i < a.rank; ++i) src += cnt[i] * sa[perm[i]];
      (*out)[e] = (*kv.second)[src];
      for (int i = a.rank - 1; i >= 0; --i) {
        if (++cnt[i] < dc[i]) break;
        cnt[i] = 0;
      }
    }
    c.blocks[c.tag(ic)] = out;
  }
  return c;
}

// Optional CBLAS dgemm with 64-bit integers (an ILP64 BLAS such as numpy's bundled OpenBLAS), set by
// orc_set_dgemm for CPU timings: the reference's contractions call CBLAS dgemm per block pair (BTAS
// DENSE/BLAS_STL_vector.h through STBLAS.h:298-359).  Unset (the default) the loops below run.
typedef void (*cblas_dgemm64_fn)(int order, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                                 const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                                 int64_t ldc);
cblas_dgemm64_fn g_cblas_dgemm = nullptr;

// C(m x n) += alpha * A(m x k) * B(k x n), row-major
void dgemm(long m, long n, long k, double alpha, const double* A, const double* B, double* C) {
  if (g_cblas_dgemm) {
    g_cblas_dgemm(101 /* CblasRowMajor */, 111 /* CblasNoTrans */, 111, m, n, k, alpha, A, k, B, n, 1.0, C, n);
    return;
  }
  for (long i = 0; i < m; ++i) {
    double* c = C + i * n;
    for (long p = 0; p < k; ++p) {
      const double a = alpha * A[i * k + p];
      const double* b = B + p * n;
      for (long j = 0; j < n; ++j) c[j] += a * b[j];
    }
  }
}

struct Task {
  Block* c;
  long m, n;
  std::vector<std::pair<const Block*, const Block*>> ab;
  std::vector<long> ks;
  double flops;
};

int g_threads = 1;

// Contract (QSTCONTRACT.h:16-57); only 0 < K < NA, NB (every call on the hot path)
void contract(double alpha, const QST& a, const std::vector<int>& ca, const QST& b, const std::vector<int>& cb,
              double beta, QST& c) {
  const int NA = a.rank, NB = b.rank, K = (int)ca.size();
  if (K <= 0 || K >= NA || K >= NB || (int)cb.size() != K) throw std::runtime_error("oracle contract: unsupported K");
  // get_contract_jobs: A -> (uncontracted..., contracted...), B -> (contracted..., uncontracted...)
  std::vector<int> pa, pb;
  for (int i = 0; i < NA; ++i)
    if (std::find(ca.begin(), ca.end(), i) == ca.end()) pa.push_back(i);
  for (int x : ca) pa.push_back(x);
  for (int x : cb) pb.push_back(x);
  for (int i = 0; i < NB; ++i)
    if (std::find(cb.begin(), cb.end(), i) == cb.end()) pb.push_back(i);
  bool ida = true, idb = true;
  for (int i = 0; i < NA; ++i) ida &= (pa[i] == i);
  for (int i = 0; i < NB; ++i) idb &= (pb[i] == i);
  const QST ap = ida ? a : permute(a, pa);
  const QST bp = idb ? b : permute(b, pb);
  // gemm_contract_qshape (NoTrans, NoTrans)
  const bool qsym = a.qsym && b.qsym;
  for (int k = 0; k < K; ++k) {
    const auto& qa = ap.qs[NA - K + k];
    const auto& qb = bp.qs[k];
    if (qa.size() != qb.size()) throw std::runtime_error("oracle: contracted sector counts differ");
    for (size_t s = 0; s < qa.size(); ++s)
      if (qsym && qa[s] != -qb[s]) throw std::runtime_error("btas::gemm_contract_qshape contraction of quantum numbers failed");
  }
  std::vector<std::vector<QN>> qs;
  std::vector<std::vector<int>> ds;
  for (int i = 0; i < NA - K; ++i) { qs.push_back(ap.qs[i]); ds.push_back(ap.ds[i]); }
  for (int i = K; i < NB; ++i) { qs.push_back(bp.qs[i]); ds.push_back(bp.ds[i]); }
  const QN qc = ap.q + bp.q;
  if (!c.empty()) {
    if (qsym && c.q != qc) throw std::runtime_error("Gemm(QSPARSE): quantum number of c must equal to a.q() + b.q().");
    if (qsym && c.qs != qs) throw std::runtime_error("Gemm(QSPARSE): c must have the same quantum shape of [ a * b ].");
    for (auto& kv : c.blocks) {  // Scal(beta, c)
      auto nb = std::make_shared<Block>(*kv.second);
      for (double& x : *nb) x *= beta;
      kv.second = nb;
    }
  } else {
    c.set_shape(NA + NB - 2 * K, qsym, qc, qs, ds);
  }
  // rows/cols/K-tags of the permuted operands
  long colsA = 1, rowsA = 1, colsB = 1;
  for (int i = 0; i < NA - K; ++i) rowsA *= ap.nsec(i);
  for (int i = NA - K; i < NA; ++i) colsA *= ap.nsec(i);
  for (int i = K; i < NB; ++i) colsB *= bp.nsec(i);
  std::map<long, std::vector<std::pair<long, const Block*>>> arow, bcol;
  for (const auto& kv : ap.blocks) arow[kv.first / colsA].push_back({kv.first % colsA, kv.second.get()});
  for (const auto& kv : bp.blocks) bcol[kv.first % colsB].push_back({kv.first / colsB, kv.second.get()});
  std::vector<Task> tasks;
  for (const auto& ra : arow) {
    const long i = ra.first;
    for (const auto& cbk : bcol) {
      const long j = cbk.first;
      const long ij = i * colsB + j;
      std::vector<int> cidx = c.index(ij);
      if (!c.allowed(cidx)) continue;  // STBLAS.h:331
      Task t;
      for (const auto& x : ra.second)
        for (const auto& y : cbk.second)
          if (x.first == y.first) {
            t.ab.push_back({x.second, y.second});
            // K extent from A's contracted sectors
            std::vector<int> ia = ap.index(i * colsA + x.first);
            long kk = 1;
            for (int q = NA - K; q < NA; ++q) kk *= ap.ds[q][ia[q]];
            t.ks.push_back(kk);
          }
      if (t.ab.empty()) continue;
      long m = 1, n = 1;
      for (int q = 0; q < NA - K; ++q) m *= c.ds[q][cidx[q]];
      for (int q = NA - K; q < c.rank; ++q) n *= c.ds[q][cidx[q]];
      t.m = m;
      t.n = n;
      t.c = c.reserve(ij);
      if (!t.c) continue;
      // copy-on-write: C blocks may share storage with a conjugated reference
      t.flops = 0;
      for (long kk : t.ks) t.flops += (double)m * n * kk;
      tasks.push_back(std::move(t));
    }
  }
  std::sort(tasks.begin(), tasks.end(), [](const Task& x, const Task& y) { return x.flops > y.flops; });
#pragma omp parallel for schedule(guided) num_threads(g_threads)
  for (long t = 0; t < (long)tasks.size(); ++t) {
    Task& T = tasks[t];
    for (size_t p = 0; p < T.ab.size(); ++p)
      dgemm(T.m, T.n, T.ks[p], alpha, T.ab[p].first->data(), T.ab[p].second->data(), T.c->data());
  }
}

// btas::Gemm index lists
void gemm(int ta, int tb, double alpha, const QST& a, const QST& b, double beta, QST& c, int NC) {
  const int K = (a.rank + b.rank - NC) / 2;
  std::vector<int> ca, cb;
  for (int k = 0; k < K; ++k) ca.push_back(ta ? k : a.rank - K + k);
  for (int k = 0; k < K; ++k) cb.push_back(tb ? b.rank - K + k : k);
  contract(alpha, a, ca, b, cb, beta, c);
}

QST tie(const QST& x, const std::vector<int>& t) {
  const int N = x.rank, K = (int)t.size();
  std::vector<int> uniq;
  for (int i = 0; i < N; ++i) {
    const bool tied = std::find(t.begin(), t.end(), i) != t.end();
    if (!tied || i == t[0]) uniq.push_back(i);
  }
  QST y;
  std::vector<std::vector<int>> ds;
  for (int i : uniq) ds.push_back(x.ds[i]);
  y.set_shape((int)uniq.size(), false, QN(), {}, ds);
  for (const auto& kv : x.blocks) {
    std::vector<int> ix = x.index(kv.first);
    bool ok = true;
    for (int k = 1; k < K; ++k) ok &= ix[t[k]] == ix[t[0]];
    if (!ok) continue;
    std::vector<int> iy;
    for (int i : uniq) iy.push_back(ix[i]);
    std::vector<long> dx(N), sx(N);
    long s = 1;
    for (int i = N - 1; i >= 0; --i) { dx[i] = x.ds[i][ix[i]]; sx[i] = s; s *= dx[i]; }
    std::vector<long> dy, sy;
    for (int i : uniq) {
      dy.push_back(dx[i]);
      long st = sx[i];
      if (i == t[0]) { st = 0; for (int k = 0; k < K; ++k) st += sx[t[k]]; }
      sy.push_back(st);
    }
    long ny = 1;
    for (long d : dy) ny *= d;
    auto out = std::make_shared<Block>(ny, 0.0);
    std::vector<long> cnt(dy.size(), 0);
    for (long e = 0; e < ny; ++e) {
      long src = 0;
      for (size_t i = 0; i < dy.size(); ++i) src += cnt[i] * sy[i];
      (*out)[e] = (*kv.second)[src];
      for (int i = (int)dy.size() - 1; i >= 0; --i) {
        if (++cnt[i] < dy[i]) break;
        cnt[i] = 0;
      }
    }
    y.blocks[y.tag(iy)] = out;
  }
  return y;
}

// Copy(x, y, UpCast=1): y keeps q/qshape, its blocks are cleared, allowed x blocks are copied
void copy_upcast(const QST& x, QST& y) {
  y.ds = x.ds;
  y.blocks.clear();
  for (const auto& kv : x.blocks) {
    Block* b = y.reserve(kv.first);
    if (!b) continue;
    *b = *kv.second;
  }
}

double dotc(const QST& x, const QST& y) {
  double v = 0.0;
  for (const auto& kv : x.blocks) {
    auto it = y.blocks.find(kv.first);
    if (it == y.blocks.end()) continue;
    const Block& a = *kv.second;
    const Block& b = *it->second;
    for (size_t i = 0; i < a.size(); ++i) v += a[i] * b[i];
  }
  return v;
}

void scal(double alpha, QST& x) {
  for (auto& kv : x.blocks) {
    auto nb = std::make_shared<Block>(*kv.second);
    for (double& v : *nb) v *= alpha;
    kv.second = nb;
  }
}

void axpy(double alpha, const QST& x, QST& y) {
  if (y.empty()) { y.set_shape(x.rank, x.qsym, x.q, x.qs, x.ds); }
  for (const auto& kv : x.blocks) {
    Block* b = y.reserve(kv.first);
    if (!b) throw std::runtime_error("ST_Axpy_serial: reservation failed");
    // copy-on-write when shared
    auto it = y.blocks.find(kv.first);
    if (it->second.use_count() > 1) { it->second = std::make_shared<Block>(*it->second); b = it->second.get(); }
    for (size_t i = 0; i < b->size(); ++i) (*b)[i] += alpha * (*kv.second)[i];
  }
}

QST deep_copy(const QST& x) {
  QST y = x;
  for (auto& kv : y.blocks) kv.second = std::make_shared<Block>(*kv.second);
  return y;
}

// ---- quantum-number merge (QSTmergeInfo.h:44-95, QSTmerge.h:31-140) ---------------------------------------------

struct MergeInfo {
  std::vector<QN> qm;
  std::vector<int> dm;
  std::multimap<int, int> m2p;  // merged -> packed (ascending packed order)
  std::vector<int> pdim;
};

MergeInfo merge_info(const std::vector<std::vector<QN>>& qs, const std::vector<std::vector<int>>& ds) {
  MergeInfo M;
  std::vector<QN> qp{QN()};
  std::vector<int> dp{1};
  for (size_t i = 0; i < qs.size(); ++i) {
    std::vector<QN> q2;
    std::vector<int> d2;
    for (size_t a = 0; a < qp.size(); ++a)
      for (size_t b = 0; b < qs[i].size(); ++b) {
        q2.push_back(qp[a] + qs[i][b]);
        d2.push_back(dp[a] * ds[i][b]);
      }
    qp = q2;
    dp = d2;
  }
  // first-appearance numbering of distinct packed quantum numbers
  std::vector<QN> seen;
  for (size_t k = 0; k < qp.size(); ++k) {
    int j = -1;
    for (size_t t = 0; t < seen.size(); ++t)
      if (seen[t] == qp[k]) j = (int)t;
    if (j < 0) {
      j = (int)seen.size();
      seen.push_back(qp[k]);
      M.dm.push_back(0);
    }
    M.m2p.insert({j, (int)k});
    M.dm[j] += dp[k];
  }
  M.qm = seen;
  M.pdim = dp;
  return M;
}

// merge the first MR indices (row merge)
QST merge_rows(const MergeInfo& R, int MR, const QST& a) {
  const int N = a.rank, MC = N - MR;
  QST b;
  std::vector<std::vector<QN>> qs{R.qm};
  std::vector<std::vector<int>> ds{R.dm};
  for (int i = 0; i < MC; ++i) { qs.push_back(a.qs[MR + i]); ds.push_back(a.ds[MR + i]); }
  b.set_shape(1 + MC, a.qsym, a.q, qs, ds);
  long a_stride = 1, b_stride = 1;
  for (int i = MR; i < N; ++i) a_stride *= a.nsec(i);
  for (int i = 1; i < b.rank; ++i) b_stride *= b.nsec(i);
  for (int i = 0; i < b.nsec(0); ++i) {
    auto range = R.m2p.equal_range(i);
    if (range.first == range.second) continue;
    for (long j = 0; j < b_stride; ++j) {
      std::vector<int> bi = b.index(i * b_stride + j);
      if (!b.allowed(bi)) continue;
      const long ncol = b.bsize(bi) / b.ds[0][bi[0]];
      Block blk(b.bsize(bi), 0.0);
      bool nz = false;
      long row = 0;
      for (auto it = range.first; it != range.second; ++it) {
        const int irow = it->second, drow = R.pdim[irow];
        auto f = a.blocks.find(irow * a_stride + j);
        if (f != a.blocks.end()) {
          nz = true;
          std::copy(f->second->begin(), f->second->end(), blk.begin() + row * ncol);
        }
        row += drow;
      }
      if (nz) b.blocks[i * b_stride + j] = std::make_shared<Block>(blk);
    }
  }
  return b;
}

// merge the last MC indices (column merge)
QST merge_cols(const QST& a, const MergeInfo& C, int MC) {
  const int N = a.rank, MR = N - MC;
  QST b;
  std::vector<std::vector<QN>> qs;
  std::vector<std::vector<int>> ds;
  for (int i = 0; i < MR; ++i) { qs.push_back(a.qs[i]); ds.push_back(a.ds[i]); }
  qs.push_back(C.qm);
  ds.push_back(C.dm);
  b.set_shape(MR + 1, a.qsym, a.q, qs, ds);
  long a_stride = 1, b_stride = b.nsec(MR);
  for (int i = MR; i < N; ++i) a_stride *= a.nsec(i);
  long nrows = 1;
  for (int i = 0; i < MR; ++i) nrows *= b.nsec(i);
  for (long i = 0; i < nrows; ++i)
    for (long j = 0; j < b_stride; ++j) {
      auto range = C.m2p.equal_range((int)j);
      if (range.first == range.second) continue;
      std::vector<int> bi = b.index(i * b_stride + j);
      if (!b.allowed(bi)) continue;
      const long Dc = b.ds[MR][bi[MR]];
      const long R = b.bsize(bi) / Dc;
      Block blk(b.bsize(bi), 0.0);
      bool nz = false;
      long col = 0;
      for (auto it = range.first; it != range.second; ++it) {
        const int jcol = it->second, dcol = C.pdim[jcol];
        auto f = a.blocks.find(i * a_stride + jcol);
        if (f != a.blocks.end()) {
          nz = true;
          for (long r = 0; r < R; ++r)
            for (long c = 0; c < dcol; ++c) blk[r * Dc + col + c] = (*f->second)[r * dcol + c];
        }
        col += dcol;
      }
      if (nz) b.blocks[i * b_stride + j] = std::make_shared<Block>(blk);
    }
  return b;
}

// Contract by symbols (QSTCONTRACT.h:61-89): contract, then permute into the requested symbol order
QST contract_symbols(const QST& a, const std::vector<int>& sa, const QST& b, const std::vector<int>& sb,
                     const std::vector<int>& sc) {
  std::vector<int> ca, cb, axb;
  for (size_t i = 0; i < sa.size(); ++i) {
    auto it = std::find(sb.begin(), sb.end(), sa[i]);
    if (it != sb.end()) {
      ca.push_back((int)i);
      cb.push_back((int)(it - sb.begin()));
    } else {
      axb.push_back(sa[i]);
    }
  }
  for (int s : sb)
    if (std::find(sa.begin(), sa.end(), s) == sa.end()) axb.push_back(s);
  QST c;
  if (ca.empty()) {
    // no common symbol: the direct product c(a..., b...) = a(a...) b(b...) (QSTCONTRACT.h with K = 0)
    std::vector<std::vector<QN>> qs(a.qs);
    std::vector<std::vector<int>> ds(a.ds);
    qs.insert(qs.end(), b.qs.begin(), b.qs.end());
    ds.insert(ds.end(), b.ds.begin(), b.ds.end());
    c.set_shape(a.rank + b.rank, a.qsym && b.qsym, a.q + b.q, qs, ds);
    long nbt = 1;
    for (int i = 0; i < b.rank; ++i) nbt *= b.nsec(i);
    for (const auto& x : a.blocks)
      for (const auto& y : b.blocks) {
        const long t = x.first * nbt + y.first;
        if (!c.allowed(c.index(t))) continue;
        auto blk = std::make_shared<Block>(x.second->size() * y.second->size());
        for (size_t i = 0; i < x.second->size(); ++i)
          for (size_t j = 0; j < y.second->size(); ++j) (*blk)[i * y.second->size() + j] = (*x.second)[i] * (*y.second)[j];
        c.blocks[t] = blk;
      }
  } else {
    contract(1.0, a, ca, b, cb, 1.0, c);
  }
  if (axb == sc) return c;
  std::vector<int> perm;
  for (int s : sc) perm.push_back((int)(std::find(axb.begin(), axb.end(), s) - axb.begin()));
  return permute(c, perm);
}

// compute_merged_operator for one group (src/optimize.hpp:296-368)
QST merged_operator(bool fwd, const QST& mpo, const QST& opr) {
  if (fwd) {
    MergeInfo Mb = merge_info({opr.qs[0], mpo.qs[1]}, {opr.ds[0], mpo.ds[1]});
    MergeInfo Mk = merge_info({opr.qs[2], mpo.qs[2]}, {opr.ds[2], mpo.ds[2]});
    QST l = contract_symbols(opr, {0, 1, 2}, mpo, {1, 3, 4, 5}, {0, 3, 5, 2, 4});
    QST s = merge_rows(Mb, 2, l);
    return merge_cols(s, Mk, 2);
  }
  MergeInfo Mb = merge_info({mpo.qs[1], opr.qs[0]}, {mpo.ds[1], opr.ds[0]});
  MergeInfo Mk = merge_info({mpo.qs[2], opr.qs[2]}, {mpo.ds[1], opr.ds[0]});
  QST r = contract_symbols(mpo, {0, 1, 2, 3}, opr, {4, 3, 5}, {1, 4, 0, 2, 5});
  QST s = merge_rows(Mb, 2, r);
  return merge_cols(s, Mk, 2);
}

// merged overlap of compute_merged_operator (src/optimize.hpp:296-299,326-328 forward; :334-336,363-365 backward):
// I = QSTArray(zero, (mpo.qshape(1), mpo.qshape(2)), (mpo.dshape(1), mpo.dshape(2)), 1.0) on its allowed blocks;
// forward Contract(S, (0,1), I, (2,3), -> (0,2,1,3)), backward Contract(I, (0,1), S, (2,3), -> (0,2,1,3)), then
// QSTmerge of the first two and the last two indices with the operators' merge infos.
QST merged_overlap(bool fwd, const QST& mpo, const QST& sov) {
  QST I;
  I.set_shape(2, true, QN(), {mpo.qs[1], mpo.qs[2]}, {mpo.ds[1], mpo.ds[2]});
  for (long t = 0; t < I.nsec(0) * I.nsec(1); ++t) {
    std::vector<int> idx = I.index(t);
    if (!I.allowed(idx) || I.bsize(idx) == 0) continue;
    I.blocks[t] = std::make_shared<Block>(I.bsize(idx), 1.0);
  }
  if (fwd) {
    MergeInfo Mb = merge_info({sov.qs[0], mpo.qs[1]}, {sov.ds[0], mpo.ds[1]});
    MergeInfo Mk = merge_info({sov.qs[1], mpo.qs[2]}, {sov.ds[1], mpo.ds[2]});
    QST x = contract_symbols(sov, {0, 1}, I, {2, 3}, {0, 2, 1, 3});
    return merge_cols(merge_rows(Mb, 2, x), Mk, 2);
  }
  MergeInfo Mb = merge_info({mpo.qs[1], sov.qs[0]}, {mpo.ds[1], sov.ds[0]});
  MergeInfo Mk = merge_info({mpo.qs[2], sov.qs[1]}, {mpo.ds[2], sov.ds[1]});
  QST x = contract_symbols(I, {0, 1}, sov, {2, 3}, {0, 2, 1, 3});
  return merge_cols(merge_rows(Mb, 2, x), Mk, 2);
}

// ---- QN-blocked SVD with truncation (QSTLAPACK.h:122-288) and two-site canonicalize (canonicalize.hpp:27-46) ------

// One-sided Jacobi SVD of a row-major m x n matrix: U (m x k), s (k, descending), Vt (k x n), k = min(m, n).
// (The reference calls LAPACKE_dgesvd('S','S'); singular values agree to rounding, vectors up to sign/rotation in
// degenerate subspaces.)
void jacobi_svd(int m, int n, const double* A, std::vector<double>& U, std::vector<double>& s, std::vector<double>& Vt) {
  const bool tr = m < n;
  const int R = tr ? n : m, C = tr ? m : n;  // work on W (R x C), R >= C, columns orthogonalised
  std::vector<double> W((size_t)R * C), V((size_t)C * C, 0.0);
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) W[(size_t)i * C + j] = tr ? A[(size_t)j * n + i] : A[(size_t)i * n + j];
  for (int j = 0; j < C; ++j) V[(size_t)j * C + j] = 1.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < C; ++p)
      for (int q = p + 1; q < C; ++q) {
        double a = 0, b = 0, g = 0;
        for (int i = 0; i < R; ++i) {
          const double x = W[(size_t)i * C + p], y = W[(size_t)i * C + q];
          a += x * x;
          b += y * y;
          g += x * y;
        }
        if (std::fabs(g) <= 1e-15 * std::sqrt(a * b) || g == 0.0) continue;
        off = std::max(off, std::fabs(g) / std::sqrt(a * b));
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < R; ++i) {
          const double x = W[(size_t)i * C + p], y = W[(size_t)i * C + q];
          W[(size_t)i * C + p] = c * x - sn * y;
          W[(size_t)i * C + q] = sn * x + c * y;
        }
        for (int i = 0; i < C; ++i) {
          const double x = V[(size_t)i * C + p], y = V[(size_t)i * C + q];
          V[(size_t)i * C + p] = c * x - sn * y;
          V[(size_t)i * C + q] = sn * x + c * y;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> sv(C);
  for (int j = 0; j < C; ++j) {
    double a = 0;
    for (int i = 0; i < R; ++i) a += W[(size_t)i * C + j] * W[(size_t)i * C + j];
    sv[j] = std::sqrt(a);
  }
  std::vector<int> ord(C);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sv[x] > sv[y]; });
  const int k = C;
  s.resize(k);
  // left vectors of W: columns / s ; right vectors: columns of V
  std::vector<double> Lw((size_t)R * k, 0.0), Rw((size_t)C * k);
  for (int j = 0; j < k; ++j) {
    const int o = ord[j];
    s[j] = sv[o];
    for (int i = 0; i < R; ++i) Lw[(size_t)i * k + j] = sv[o] > 0 ? W[(size_t)i * C + o] / sv[o] : 0.0;
    for (int i = 0; i < C; ++i) Rw[(size_t)i * k + j] = V[(size_t)i * C + o];
  }
  if (!tr) {  // A = Lw s Rw^T
    U = Lw;
    Vt.assign((size_t)k * n, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < n; ++i) Vt[(size_t)j * n + i] = Rw[(size_t)i * k + j];
  } else {    // A^T = Lw s Rw^T -> A = Rw s Lw^T
    U = Rw;
    Vt.assign((size_t)k * n, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < n; ++i) Vt[(size_t)j * n + i] = Lw[(size_t)i * k + j];
  }
}

// QSTexpand of the rows of a (R, s) matrix into (l, n, s) and of the columns of (s, C) into (s, n, r)
QST expand_rows(const MergeInfo& Mr, const std::vector<std::vector<QN>>& qs, const std::vector<std::vector<int>>& ds,
                const QST& a) {
  QST b;
  b.set_shape(3, true, a.q, {qs[0], qs[1], a.qs[1]}, {ds[0], ds[1], a.ds[1]});
  for (const auto& kv : a.blocks) {
    std::vector<int> ai = a.index(kv.first);
    const long ncol = a.ds[1][ai[1]];
    long row = 0;
    auto range = Mr.m2p.equal_range(ai[0]);
    for (auto it = range.first; it != range.second; ++it) {
      const int p = it->second, dr = Mr.pdim[p];
      if (dr == 0) continue;
      std::vector<int> bi{p / (int)qs[1].size(), p % (int)qs[1].size(), ai[1]};
      Block* blk = b.reserve(b.tag(bi));
      if (blk) std::copy(kv.second->begin() + row * ncol, kv.second->begin() + (row + dr) * ncol, blk->begin());
      row += dr;
    }
  }
  return b;
}

QST expand_cols(const QST& a, const MergeInfo& Mc, const std::vector<std::vector<QN>>& qs,
                const std::vector<std::vector<int>>& ds) {
  QST b;
  b.set_shape(3, true, a.q, {a.qs[0], qs[0], qs[1]}, {a.ds[0], ds[0], ds[1]});
  for (const auto& kv : a.blocks) {
    std::vector<int> ai = a.index(kv.first);
    const long nrow = a.ds[0][ai[0]], Dc = a.ds[1][ai[1]];
    long col = 0;
    auto range = Mc.m2p.equal_range(ai[1]);
    for (auto it = range.first; it != range.second; ++it) {
      const int p = it->second, dc = Mc.pdim[p];
      if (dc == 0) continue;
      std::vector<int> bi{ai[0], p / (int)qs[1].size(), p % (int)qs[1].size()};
      Block* blk = b.reserve(b.tag(bi));
      if (blk)
        for (long r = 0; r < nrow; ++r)
          for (long c = 0; c < dc; ++c) (*blk)[r * dc + c] = (*kv.second)[r * Dc + col + c];
      col += dc;
    }
  }
  return b;
}

// Truncation cutoff over all singular values.
//   rule 0, the reference's (QSTLAPACK.h:209-213): cutoff = 1e-16, or the M-th largest value when 0 < M < count;
//     every value >= cutoff is kept.
//   rule 1, the stable variant: values within a relative 1e-12 of the M-th largest count as equal to it (exact
//     degeneracies split by rounding stay together on every backend), and the cutoff never drops below 1e-16 (values
//     at the rounding level are not kept just because the M-th largest is one of them).
double truncation_cutoff(std::vector<double> all, int M, int rule) {
  std::sort(all.rbegin(), all.rend());
  double cutoff = 1.0e-16;
  if (M > 0 && M < (int)all.size()) cutoff = rule == 0 ? all[M - 1] : std::max(cutoff, all[M - 1] * (1.0 - 1e-12));
  return cutoff;
}

// Rule 1 also reports singular values below max(m, n) eps |A|_F of their block (the backward-error level of an
// orthogonal decomposition) as exact zeros; rule 0 keeps them as computed.
void numerical_rank(int m, int n, std::vector<double>& s) {
  double fro = 0.0;
  for (double v : s) fro += v * v;
  const double thr = std::max(m, n) * 2.220446049250313e-16 * std::sqrt(fro);
  for (double& v : s)
    if (v < thr) v = 0.0;
}

double canonicalize_twosite(bool forward, const QST& wfnc, QST& lmps, QST& rmps, int M, int rule = 1) {
  MergeInfo Mr = merge_info({wfnc.qs[0], wfnc.qs[1]}, {wfnc.ds[0], wfnc.ds[1]});
  MergeInfo Mc = merge_info({wfnc.qs[2], wfnc.qs[3]}, {wfnc.ds[2], wfnc.ds[3]});
  QST a = merge_cols(merge_rows(Mr, 2, wfnc), Mc, 2);
  struct Dec { int i, j, m, n; std::vector<double> U, s, Vt; };
  std::vector<Dec> dec;
  for (const auto& kv : a.blocks) {
    std::vector<int> ai = a.index(kv.first);
    Dec d;
    d.i = ai[0];
    d.j = ai[1];
    d.m = a.ds[0][ai[0]];
    d.n = a.ds[1][ai[1]];
    jacobi_svd(d.m, d.n, kv.second->data(), d.U, d.s, d.Vt);
    if (rule == 1) numerical_rank(d.m, d.n, d.s);
    dec.push_back(std::move(d));
  }
  std::vector<double> all;
  for (auto& d : dec) all.insert(all.end(), d.s.begin(), d.s.end());
  const double cutoff = truncation_cutoff(all, M, rule);
  double dnorm = 0.0;
  // kept sectors in sector order (rows for LeftArrow, columns for RightArrow)
  const int nsec = forward ? a.nsec(0) : a.nsec(1);
  std::vector<QN> qnz;
  std::vector<int> dnz;
  std::vector<int> secmap(dec.size(), -1), kept(dec.size(), 0);
  for (size_t k = 0; k < dec.size(); ++k) {
    int r = 0;
    while (r < (int)dec[k].s.size() && dec[k].s[r] >= cutoff) ++r;
    kept[k] = r;
    for (int t = r; t < (int)dec[k].s.size(); ++t) dnorm += dec[k].s[t] * dec[k].s[t];
  }
  for (int q = 0; q < nsec; ++q)
    for (size_t k = 0; k < dec.size(); ++k)
      if ((forward ? dec[k].i : dec[k].j) == q && kept[k] > 0) {
        secmap[k] = (int)qnz.size();
        qnz.push_back(forward ? a.qs[0][q] : -a.qs[1][q]);
        dnz.push_back(kept[k]);
      }
  std::vector<QN> nq;
  for (const QN& q : qnz) nq.push_back(-q);
  QST um, vm;
  um.set_shape(2, true, forward ? QN() : a.q, {a.qs[0], nq}, {a.ds[0], dnz});
  vm.set_shape(2, true, forward ? a.q : QN(), {qnz, a.qs[1]}, {dnz, a.ds[1]});
  for (size_t k = 0; k < dec.size(); ++k) {
    if (secmap[k] < 0) continue;
    const Dec& d = dec[k];
    const int kk = (int)d.s.size(), r = kept[k];
    Block* ub = um.reserve(um.tag({d.i, secmap[k]}));
    Block* vb = vm.reserve(vm.tag({secmap[k], d.j}));
    for (int i = 0; i < d.m; ++i)
      for (int c = 0; c < r; ++c) (*ub)[(size_t)i * r + c] = d.U[(size_t)i * kk + c] * (forward ? 1.0 : d.s[c]);
    for (int c = 0; c < r; ++c)
      for (int j = 0; j < d.n; ++j) (*vb)[(size_t)c * d.n + j] = d.Vt[(size_t)c * d.n + j] * (forward ? d.s[c] : 1.0);
  }
  lmps = expand_rows(Mr, {wfnc.qs[0], wfnc.qs[1]}, {wfnc.ds[0], wfnc.ds[1]}, um);
  rmps = expand_cols(vm, Mc, {wfnc.qs[2], wfnc.qs[3]}, {wfnc.ds[2], wfnc.ds[3]});
  return dnorm;
}

// ---- hot path compositions ---------------------------------------------------------------------------------------

void sigma_twosite(const QST& lmpo, const QST& rmpo, const QST& lopr, const QST& ropr, const QST& wfn, QST& sgv) {
  QST s1, s2, s3;  // compute_sigma_vector.hpp:63-69
  contract(1.0, lopr, {2}, wfn, {0}, 1.0, s1);
  contract(1.0, s1, {1, 2}, lmpo, {0, 2}, 1.0, s2);
  contract(1.0, s2, {4, 1}, rmpo, {0, 2}, 1.0, s3);
  contract(1.0, s3, {4, 1}, ropr, {1, 2}, 1.0, sgv);
}

void sigma_onesite(const QST& mpo, const QST& lopr, const QST& ropr, const QST& wfn, QST& sgv) {
  QST s1, s2;  // compute_sigma_vector.hpp:17-21
  gemm(0, 0, 1.0, lopr, wfn, 1.0, s1, 4);
  contract(1.0, s1, {1, 2}, mpo, {0, 2}, 1.0, s2);
  contract(1.0, s2, {3, 1}, ropr, {1, 2}, 1.0, sgv);
}

void sigma_merged(const QST& lopr, const QST& ropr, const QST& wfn, QST& sgv) {
  QST s1;  // compute_sigma_vector.hpp:94-95
  gemm(0, 0, 1.0, lopr, wfn, 1.0, s1, 3);
  gemm(0, 1, 1.0, s1, ropr, 1.0, sgv, 2);
}

void sigma_overlap(const QST& lsov, const QST& rsov, const QST& wfn, QST& sgv) {
  QST s1;  // compute_sigma_vector.hpp:32-33, 80-81, 106-107
  gemm(0, 0, 1.0, lsov, wfn, 1.0, s1, wfn.rank);
  gemm(0, 1, 1.0, s1, rsov, 1.0, sgv, wfn.rank);
}

// compute_sigma_vector_conj (compute_sigma_vector.hpp:38-51): < phi | = < psi | H^*.  Gemm(Trans, ConjTrans, lopr,
// wfn0) contracts lopr's first index with the last index of the conjugated wfn0 (QSTBLAS.h:213-225,
// btas_contract_qshape.h:60-108), then Contract(scr1, {0,2}, mpo0, {0,1}) and Contract(scr2, {1,3}, ropr, {0,1})
// accumulate into sgv0.conjugate().  The reference's target is a temporary; here it is sgv0 itself: the result is
// computed with the conjugated quantum numbers and they are flipped back.
void sigma_conj(const QST& mpo, const QST& lopr, const QST& ropr, const QST& wfn, QST& sgv) {
  QST s1, s2;
  contract(1.0, lopr, {0}, conjugate(wfn), {2}, 1.0, s1);
  contract(1.0, s1, {0, 2}, mpo, {0, 1}, 1.0, s2);
  QST out = sgv.empty() ? QST() : conjugate(sgv);
  contract(1.0, s2, {1, 3}, ropr, {0, 1}, 1.0, out);
  sgv = conjugate(out);
}

// make_noise (canonicalize.hpp:11-24): nothing below 1e-8; otherwise wfnc += noise * wfpt, where wfpt holds every
// allowed block of wfnc's net dense shape (STArray.h:386-401: a sector's dimension is taken from the stored blocks,
// sectors without a stored block have dimension 0 and get no block) and is filled by QSTArray's generator
// constructor with std::bind(uniform_real_distribution(-1, 1), mt19937()).  That constructor hands the generator to
// every block by value (STArray.h:243-262, TArray.h:779-782), so each block is filled, in row-major order, by a fresh
// copy of the default-seeded engine: block element e holds the e-th draw of the sequence.
void make_noise(QST& w, double noise) {
  if (noise < 1.0e-8) return;
  std::vector<std::vector<int>> net(w.rank);
  for (int i = 0; i < w.rank; ++i) net[i].assign(w.ds[i].size(), 0);
  for (const auto& kv : w.blocks) {
    std::vector<int> idx = w.index(kv.first);
    for (int i = 0; i < w.rank; ++i) net[i][idx[i]] = w.ds[i][idx[i]];
  }
  QST p;
  p.set_shape(w.rank, w.qsym, w.q, w.qs, net);
  long nblk = 1;
  for (int i = 0; i < w.rank; ++i) nblk *= p.nsec(i);
  for (long t = 0; t < nblk; ++t) {
    std::vector<int> idx = p.index(t);
    if (!p.allowed(idx) || p.bsize(idx) == 0) continue;
    auto b = std::make_shared<Block>(p.bsize(idx));
    std::mt19937 rgen;  // default seed, a fresh copy per block
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (double& v : *b) v = dist(rgen);
    p.blocks[t] = b;
  }
  p.ds = w.ds;  // same sector dims as w wherever a block exists
  axpy(noise, p, w);
}

void diag_twosite(const QST& lmpo, const QST& rmpo, const QST& lopr, const QST& ropr, QST& diag) {
  QST lm = tie(lmpo, {1, 2}), rm = tie(rmpo, {1, 2}), lo = tie(lopr, {0, 2}), ro = tie(ropr, {0, 2});
  QST s1, s2, s3;  // compute_diagonal_elements.hpp:47-60
  contract(1.0, lo, {1}, lm, {0}, 1.0, s1);
  contract(1.0, s1, {2}, rm, {0}, 1.0, s2);
  contract(1.0, s2, {3}, ro, {1}, 1.0, s3);
  copy_upcast(s3, diag);
}

void diag_onesite(const QST& mpo, const QST& lopr, const QST& ropr, QST& diag) {
  QST m = tie(mpo, {1, 2}), lo = tie(lopr, {0, 2}), ro = tie(ropr, {0, 2});
  QST s1, s2;  // compute_diagonal_elements.hpp:21-31
  contract(1.0, lo, {1}, m, {0}, 1.0, s1);
  contract(1.0, s1, {2}, ro, {1}, 1.0, s2);
  copy_upcast(s2, diag);
}

void diag_merged(const QST& lopr, const QST& ropr, QST& diag) {
  QST lo = tie(lopr, {0, 2}), ro = tie(ropr, {0, 2});
  QST s1;  // compute_diagonal_elements.hpp:72-79
  gemm(0, 1, 1.0, lo, ro, 1.0, s1, 2);
  copy_upcast(s1, diag);
}

void renormalize(bool fwd, const QST& mpo, const QST& opr0, const QST& bra, const QST& ket, QST& opr1) {
  QST s1, s2;  // renormalize.hpp:18-31
  if (fwd) {
    contract(1.0, opr0, {0}, conjugate(bra), {0}, 1.0, s1);
    contract(1.0, s1, {0, 2}, mpo, {0, 1}, 1.0, s2);
    contract(1.0, s2, {0, 2}, ket, {0, 1}, 1.0, opr1);
  } else {
    contract(1.0, conjugate(bra), {2}, opr0, {0}, 1.0, s1);
    contract(1.0, s1, {1, 2}, mpo, {1, 3}, 1.0, s2);
    contract(1.0, s2, {3, 1}, ket, {1, 2}, 1.0, opr1);
  }
}

void renormalize_overlap(bool fwd, const QST& sov0, const QST& bra, const QST& ket, QST& sov1) {
  QST s1;  // renormalize.hpp:43-52, 85-94
  if (fwd) {
    gemm(1, 0, 1.0, sov0, conjugate(bra), 1.0, s1, bra.rank);
    gemm(1, 0, 1.0, s1, ket, 1.0, sov1, 2);
  } else {
    gemm(0, 0, 1.0, conjugate(bra), sov0, 1.0, s1, bra.rank);
    gemm(0, 1, 1.0, s1, ket, 1.0, sov1, 2);
  }
}

void renormalize_merged(bool fwd, const QST& opr0, const QST& bra, const QST& ket, QST& opr1) {
  QST s1;  // renormalize.hpp:64-73
  if (fwd) {
    gemm(1, 0, 1.0, conjugate(bra), opr0, 1.0, s1, 3);
    gemm(0, 0, 1.0, s1, ket, 1.0, opr1, 3);
  } else {
    gemm(0, 0, 1.0, conjugate(bra), opr0, 1.0, s1, 3);
    gemm(0, 1, 1.0, s1, ket, 1.0, opr1, 3);
  }
}

// ---- Davidson (src/diagonalize.hpp:283-449), one root, group loop inside multH ----------------------------------

// Jacobi eigen-decomposition of a small symmetric matrix (row-major n x n); eigenvalues ascending, vectors in columns
// (the convention of LAPACKE_dsyev with row-major storage, TLAPACK.h:23-47)
void jacobi_eig(int n, std::vector<double> a, std::vector<double>& w, std::vector<double>& v) {
  v.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
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
  std::vector<double> v2(n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = a[ord[j] * n + ord[j]];
    for (int i = 0; i < n; ++i) v2[i * n + j] = v[i * n + ord[j]];
  }
  v.swap(v2);
}

void normalize(QST& x) { scal(1.0 / std::sqrt(dotc(x, x)), x); }

void precondition(double eval, const QST& diag, QST& errv) {
  for (auto& kv : errv.blocks) {
    auto nb = std::make_shared<Block>(*kv.second);
    auto id = diag.blocks.find(kv.first);
    if (id != diag.blocks.end()) {
      for (size_t i = 0; i < nb->size(); ++i) {
        double denm = eval - (*id->second)[i];
        if (std::fabs(denm) < 1.0e-8) denm = 1.0e-8;
        (*nb)[i] /= denm;
      }
    } else {
      for (double& v : *nb) v *= 1.0 / eval;
    }
    kv.second = nb;
  }
}

struct Functor2 {
  std::vector<const QST*> lmpo, rmpo, lopr, ropr;
  int nsigma = 0;
  void operator()(const QST& c, QST& v) {
    for (size_t g = 0; g < lmpo.size(); ++g) sigma_twosite(*lmpo[g], *rmpo[g], *lopr[g], *ropr[g], c, v);
    ++nsigma;
  }
};

double davidson(Functor2& multH, const QST& diag, QST& wfnc, double tolerance) {
  const int max_ritz = 10;
  double eval = 0.0;
  std::vector<QST> trial(max_ritz), sigma(max_ritz);
  trial[0] = deep_copy(wfnc);
  int niter = 0, iconv = 0;
  while (iconv < 1 && niter < 4) {
    normalize(trial[0]);
    sigma[0] = QST();
    multH(trial[0], sigma[0]);
    for (int m = 1; m <= max_ritz; ++m) {
      std::vector<double> heff(m * m), ovlp(m * m);
      for (int i = 0; i < m; ++i) {
        heff[i * m + i] = dotc(trial[i], sigma[i]);
        ovlp[i * m + i] = dotc(trial[i], trial[i]);
        for (int j = 0; j < i; ++j) {
          const double hij = dotc(trial[i], sigma[j]);
          heff[i * m + j] = heff[j * m + i] = hij;
          const double sij = dotc(trial[i], trial[j]);
          ovlp[i * m + j] = ovlp[j * m + i] = sij;
        }
      }
      std::vector<double> rval, rvec;
      jacobi_eig(m, heff, rval, rvec);
      eval = rval[0];
      std::vector<QST> ts(m), ss(m);
      for (int i = 0; i < m; ++i) {
        ts[i] = deep_copy(trial[i]);
        ss[i] = deep_copy(sigma[i]);
        scal(rvec[i * m + i], trial[i]);
        scal(rvec[i * m + i], sigma[i]);
      }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
          if (i != j) {
            axpy(rvec[i * m + j], ts[i], trial[j]);
            axpy(rvec[i * m + j], ss[i], sigma[j]);
          }
      QST errv = deep_copy(sigma[0]);
      axpy(-eval, trial[0], errv);
      const double rnorm = dotc(errv, errv);
      if (std::getenv("ORACLE_DAVIDSON_TRACE")) fprintf(stderr, "oracle davidson m %d eval %.16f rnorm %.6e\n", m, eval, rnorm);
      if (rnorm < tolerance) { ++iconv; break; }
      if (m < max_ritz) {
        precondition(eval, diag, errv);
        for (int i = 0; i < m; ++i) {
          normalize(errv);
          axpy(-dotc(trial[i], errv), trial[i], errv);
        }
        normalize(errv);
        trial[m] = deep_copy(errv);
        sigma[m] = QST();
        multH(trial[m], sigma[m]);
      }
    }
    ++niter;
  }
  wfnc = deep_copy(trial[0]);
  normalize(wfnc);  // Functor::orthonormalize with no lower roots
  return eval;
}

// ---- two-site sweeps (dmrg/make_sweep.cpp:12-254; initial right-canonicalisation as make_random_mpss.cpp:204-251) ---

struct Sweeper {
  int L;
  std::vector<QST> mpo, mps, lenv, renv;
  int nsigma = 0;
};

QST boundary_env() {
  QST e;
  e.set_shape(3, true, QN(), {{QN()}, {QN()}, {QN()}}, {{1}, {1}, {1}});
  e.blocks[0] = std::make_shared<Block>(1, 1.0);
  return e;
}

void sweeper_init(Sweeper& S, int M) {
  const int L = S.L;
  for (int i = L - 1; i >= 1; --i) {
    QST psi, U, B;
    contract(1.0, S.mps[i - 1], {2}, S.mps[i], {0}, 1.0, psi);
    canonicalize_twosite(false, psi, U, B, M);
    normalize(U);
    S.mps[i - 1] = U;
    S.mps[i] = B;
  }
  S.lenv.assign(L + 1, QST());
  S.renv.assign(L + 1, QST());
  S.lenv[0] = boundary_env();
  S.renv[L] = boundary_env();
  for (int i = L - 1; i >= 1; --i) renormalize(false, S.mpo[i], S.renv[i + 1], S.mps[i], S.mps[i], S.renv[i]);
}

double sweeper_sweep(Sweeper& S, int M, double tol) {
  const int L = S.L;
  double emin = 1.0e8;
  auto opt = [&](int i, bool fwd) {
    QST psi;
    contract(1.0, S.mps[i], {2}, S.mps[i + 1], {0}, 1.0, psi);
    QST diag;
    diag.set_shape(4, true, psi.q, psi.qs, psi.ds);
    diag_twosite(S.mpo[i], S.mpo[i + 1], S.lenv[i], S.renv[i + 2], diag);
    Functor2 F;
    F.lmpo = {&S.mpo[i]};
    F.rmpo = {&S.mpo[i + 1]};
    F.lopr = {&S.lenv[i]};
    F.ropr = {&S.renv[i + 2]};
    const double e = davidson(F, diag, psi, 0.1 * tol);
    S.nsigma += F.nsigma;
    emin = std::min(emin, e);
    QST U, V;
    canonicalize_twosite(fwd, psi, U, V, M);
    S.mps[i] = U;
    S.mps[i + 1] = V;
    if (fwd) {
      QST e1;
      renormalize(true, S.mpo[i], S.lenv[i], S.mps[i], S.mps[i], e1);
      S.lenv[i + 1] = e1;
    } else {
      QST e1;
      renormalize(false, S.mpo[i + 1], S.renv[i + 2], S.mps[i + 1], S.mps[i + 1], e1);
      S.renv[i + 1] = e1;
    }
  };
  for (int i = 0; i < L - 1; ++i) opt(i, true);
  for (int i = L - 2; i >= 0; --i) opt(i, false);
  return emin;
}

}  // namespace orc

// ---- C interface (ctypes) -------------------------------------------------------------------------------------------

struct orc_tensor {
  orc::QST a;
};

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_set_threads(int n) { orc::g_threads = n > 0 ? n : 1; }
void orc_set_dgemm(void* fn) { orc::g_cblas_dgemm = reinterpret_cast<orc::cblas_dgemm64_fn>(fn); }

orc_tensor* orc_create(const mpsxx_qst_desc* d, const double* data) {
  auto* t = new orc_tensor;
  std::vector<std::vector<orc::QN>> qs(d->rank);
  std::vector<std::vector<int>> ds(d->rank);
  long k = 0;
  for (int i = 0; i < d->rank; ++i)
    for (int j = 0; j < d->nsec[i]; ++j, ++k) {
      ds[i].push_back(d->sec_dim[k]);
      if (d->qsym) qs[i].push_back(orc::QN(d->sec_q[2 * k], d->sec_q[2 * k + 1]));
    }
  t->a.set_shape(d->rank, d->qsym != 0, orc::QN(d->q_total[0], d->q_total[1]), qs, ds);
  std::vector<long> tags;
  for (long b = 0; b < d->nblocks; ++b) {
    std::vector<int> idx(d->block_idx + b * d->rank, d->block_idx + (b + 1) * d->rank);
    tags.push_back(t->a.tag(idx));
  }
  std::sort(tags.begin(), tags.end());
  tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
  long off = 0;
  for (long tg : tags) {
    std::vector<int> idx = t->a.index(tg);
    const long n = t->a.bsize(idx);
    auto blk = std::make_shared<orc::Block>(data ? std::vector<double>(data + off, data + off + n)
                                                 : std::vector<double>(n, 0.0));
    t->a.blocks[tg] = blk;
    off += n;
  }
  return t;
}

orc_tensor* orc_empty(void) { return new orc_tensor; }

void orc_free(orc_tensor* t) { delete t; }

void orc_info(orc_tensor* t, int32_t* rank, int32_t* qsym, int32_t* q, int64_t* nsec_total, int64_t* nblocks,
              int64_t* nelem) {
  const orc::QST& a = t->a;
  *rank = a.rank;
  *qsym = a.qsym;
  q[0] = a.q.p;
  q[1] = a.q.s;
  long ns = 0;
  for (int i = 0; i < a.rank; ++i) ns += a.nsec(i);
  *nsec_total = ns;
  *nblocks = (int64_t)a.blocks.size();
  long ne = 0;
  for (const auto& kv : a.blocks) ne += (long)kv.second->size();
  *nelem = ne;
}

void orc_layout(orc_tensor* t, int32_t* nsec, int32_t* sec_q, int32_t* sec_dim, int32_t* block_idx, double* data) {
  const orc::QST& a = t->a;
  long k = 0;
  for (int i = 0; i < a.rank; ++i) {
    nsec[i] = (int32_t)a.nsec(i);
    for (int j = 0; j < a.nsec(i); ++j, ++k) {
      sec_dim[k] = a.ds[i][j];
      sec_q[2 * k] = a.qs[i][j].p;
      sec_q[2 * k + 1] = a.qs[i][j].s;
    }
  }
  long b = 0, off = 0;
  for (const auto& kv : a.blocks) {
    std::vector<int> idx = a.index(kv.first);
    for (int i = 0; i < a.rank; ++i) block_idx[b * a.rank + i] = idx[i];
    std::memcpy(data + off, kv.second->data(), kv.second->size() * sizeof(double));
    off += (long)kv.second->size();
    ++b;
  }
}

#define ORC_GUARD(...)                \
  try {                               \
    __VA_ARGS__;                      \
    return 0;                         \
  } catch (const std::exception& e) { \
    g_err = e.what();                 \
    return 1;                         \
  }

int orc_sigma_twosite(orc_tensor* lmpo, orc_tensor* rmpo, orc_tensor* lopr, orc_tensor* ropr, orc_tensor* wfn,
                      orc_tensor* sgv) {
  ORC_GUARD(orc::sigma_twosite(lmpo->a, rmpo->a, lopr->a, ropr->a, wfn->a, sgv->a))
}
int orc_sigma_onesite(orc_tensor* mpo, orc_tensor* lopr, orc_tensor* ropr, orc_tensor* wfn, orc_tensor* sgv) {
  ORC_GUARD(orc::sigma_onesite(mpo->a, lopr->a, ropr->a, wfn->a, sgv->a))
}
int orc_sigma_merged(orc_tensor* lopr, orc_tensor* ropr, orc_tensor* wfn, orc_tensor* sgv) {
  ORC_GUARD(orc::sigma_merged(lopr->a, ropr->a, wfn->a, sgv->a))
}
int orc_sigma_overlap(orc_tensor* lsov, orc_tensor* rsov, orc_tensor* wfn, orc_tensor* sgv) {
  ORC_GUARD(orc::sigma_overlap(lsov->a, rsov->a, wfn->a, sgv->a))
}
int orc_diag_twosite(orc_tensor* lmpo, orc_tensor* rmpo, orc_tensor* lopr, orc_tensor* ropr, orc_tensor* diag) {
  ORC_GUARD(orc::diag_twosite(lmpo->a, rmpo->a, lopr->a, ropr->a, diag->a))
}
int orc_diag_onesite(orc_tensor* mpo, orc_tensor* lopr, orc_tensor* ropr, orc_tensor* diag) {
  ORC_GUARD(orc::diag_onesite(mpo->a, lopr->a, ropr->a, diag->a))
}
int orc_diag_merged(orc_tensor* lopr, orc_tensor* ropr, orc_tensor* diag) {
  ORC_GUARD(orc::diag_merged(lopr->a, ropr->a, diag->a))
}
int orc_renormalize(int fwd, orc_tensor* mpo, orc_tensor* opr0, orc_tensor* bra, orc_tensor* ket, orc_tensor* opr1) {
  ORC_GUARD(orc::renormalize(fwd != 0, mpo->a, opr0->a, bra->a, ket->a, opr1->a))
}
int orc_renormalize_overlap(int fwd, orc_tensor* sov0, orc_tensor* bra, orc_tensor* ket, orc_tensor* sov1) {
  ORC_GUARD(orc::renormalize_overlap(fwd != 0, sov0->a, bra->a, ket->a, sov1->a))
}
int orc_renormalize_merged(int fwd, orc_tensor* opr0, orc_tensor* bra, orc_tensor* ket, orc_tensor* opr1) {
  ORC_GUARD(orc::renormalize_merged(fwd != 0, opr0->a, bra->a, ket->a, opr1->a))
}
int orc_contract(double alpha, orc_tensor* a, int conj_a, const int32_t* ca, orc_tensor* b, int conj_b,
                 const int32_t* cb, int32_t K, double beta, orc_tensor* c) {
  ORC_GUARD({
    std::vector<int> va(ca, ca + K), vb(cb, cb + K);
    orc::contract(alpha, conj_a ? orc::conjugate(a->a) : a->a, va, conj_b ? orc::conjugate(b->a) : b->a, vb, beta,
                  c->a);
  })
}
int orc_tie(orc_tensor* x, const int32_t* t, int32_t K, orc_tensor* y) {
  ORC_GUARD(y->a = orc::tie(x->a, std::vector<int>(t, t + K)))
}
int orc_copy_upcast(orc_tensor* x, orc_tensor* y) { ORC_GUARD(orc::copy_upcast(x->a, y->a)) }
int orc_merged_overlap(int fwd, orc_tensor* mpo, orc_tensor* sov, orc_tensor* out) {
  ORC_GUARD(out->a = orc::merged_overlap(fwd != 0, mpo->a, sov->a))
}
int orc_merged_operator(int fwd, orc_tensor* mpo, orc_tensor* opr, orc_tensor* out) {
  ORC_GUARD(out->a = orc::merged_operator(fwd != 0, mpo->a, opr->a))
}
int orc_qst_merge(orc_tensor* x, int nr, int nc, orc_tensor* y) {
  ORC_GUARD({
    const orc::QST& a = x->a;
    std::vector<std::vector<orc::QN>> rq(a.qs.begin(), a.qs.begin() + nr), cq(a.qs.end() - nc, a.qs.end());
    std::vector<std::vector<int>> rd(a.ds.begin(), a.ds.begin() + nr), cd(a.ds.end() - nc, a.ds.end());
    orc::QST r = nr > 1 ? orc::merge_rows(orc::merge_info(rq, rd), nr, a) : a;
    y->a = nc > 1 ? orc::merge_cols(r, orc::merge_info(cq, cd), nc) : r;
  })
}
double orc_dotc(orc_tensor* x, orc_tensor* y) { return orc::dotc(x->a, y->a); }
int orc_axpy(double alpha, orc_tensor* x, orc_tensor* y) { ORC_GUARD(orc::axpy(alpha, x->a, y->a)) }

// Davidson over ngroups operator groups (two-site form)
int orc_davidson_twosite(int ngroups, orc_tensor** lmpo, orc_tensor** rmpo, orc_tensor** lopr, orc_tensor** ropr,
                         orc_tensor* diag, orc_tensor* wfnc, double tol, double* energy, int32_t* nsigma) {
  ORC_GUARD({
    orc::Functor2 F;
    for (int g = 0; g < ngroups; ++g) {
      F.lmpo.push_back(&lmpo[g]->a);
      F.rmpo.push_back(&rmpo[g]->a);
      F.lopr.push_back(&lopr[g]->a);
      F.ropr.push_back(&ropr[g]->a);
    }
    *energy = orc::davidson(F, diag->a, wfnc->a, tol);
    *nsigma = F.nsigma;
  })
}

int orc_canonicalize_twosite(int fwd, orc_tensor* wfnc, orc_tensor* lmps, orc_tensor* rmps, int M, double* dnorm) {
  ORC_GUARD(*dnorm = orc::canonicalize_twosite(fwd != 0, wfnc->a, lmps->a, rmps->a, M))
}
// canonicalize.hpp:27-46 with noise (make_noise on a copy of wfnc) and a truncation rule (0 reference, 1 stable)
int orc_canonicalize_twosite_ex(int fwd, orc_tensor* wfnc, orc_tensor* lmps, orc_tensor* rmps, int M, double noise,
                                int rule, double* dnorm) {
  ORC_GUARD({
    orc::QST w = orc::deep_copy(wfnc->a);
    orc::make_noise(w, noise);
    *dnorm = orc::canonicalize_twosite(fwd != 0, w, lmps->a, rmps->a, M, rule);
  })
}
int orc_truncation_cutoff(int n, const double* values, int M, int rule, double* cutoff) {
  ORC_GUARD(*cutoff = orc::truncation_cutoff(std::vector<double>(values, values + n), M, rule))
}
int orc_make_noise(orc_tensor* w, double noise) { ORC_GUARD(orc::make_noise(w->a, noise)) }
// the first n values of uniform_real_distribution<double>(-1, 1) on one default-seeded mt19937, in draw order (the
// per-site values of make_random_mpss.cpp:108-109,153,171)
int orc_mt19937_uniform(int n, double* out) {
  ORC_GUARD({
    std::mt19937 rgen;
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int i = 0; i < n; ++i) out[i] = dist(rgen);
  })
}
int orc_sigma_conj(orc_tensor* mpo, orc_tensor* lopr, orc_tensor* ropr, orc_tensor* wfn, orc_tensor* sgv) {
  ORC_GUARD(orc::sigma_conj(mpo->a, lopr->a, ropr->a, wfn->a, sgv->a))
}

// Full two-site DMRG on the CPU: nsweeps sweeps at bond dimension Ms[k]; energies[k] = sweep energy.
int orc_dmrg(int L, orc_tensor** mpo, orc_tensor** init, int nsweeps, const int32_t* Ms, double tol, double* energies,
             int32_t* nsigma) {
  ORC_GUARD({
    orc::Sweeper S;
    S.L = L;
    for (int i = 0; i < L; ++i) {
      S.mpo.push_back(mpo[i]->a);
      S.mps.push_back(orc::deep_copy(init[i]->a));
    }
    orc::sweeper_init(S, Ms[0]);
    for (int k = 0; k < nsweeps; ++k) energies[k] = orc::sweeper_sweep(S, Ms[k], tol);
    *nsigma = S.nsigma;
  })
}

int orc_jacobi_eig(int n, const double* a, double* w, double* v) {
  ORC_GUARD({
    std::vector<double> A(a, a + n * n), W, V;
    orc::jacobi_eig(n, A, W, V);
    std::memcpy(w, W.data(), n * sizeof(double));
    std::memcpy(v, V.data(), n * n * sizeof(double));
  })
}

}  // extern "C"
