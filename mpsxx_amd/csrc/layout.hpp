// Block-sparse array structure (no data): quantum-number sectors per index, stored blocks, canonical offsets.
//
// Mirrors the structural part of btas::QSTArray / btas::STArray (reference: include/btas/QSPARSE/QSTArray.h:16-328,
// include/btas/SPARSE/STArray.h:25-623): a block is addressed by its row-major "tag" over the per-index sector counts
// (STArray.h:338-348), a QSTArray block is allowed iff sum_i qshape[i][idx_i] == q_total (QSTArray.h:34-36), and
// conjugate() negates every quantum number (QSTArray.h:293-300).  Unlike the reference's std::map of heap blocks, the
// data of one array is a single contiguous buffer: blocks in ascending tag order, each dense row-major.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpx {

struct QN {
  int p = 0;  // particle number
  int s = 0;  // 2Sz
  QN() = default;
  QN(int p_, int s_) : p(p_), s(s_) {}
  QN operator+(const QN& o) const { return QN(p + o.p, s + o.s); }
  QN operator-(const QN& o) const { return QN(p - o.p, s - o.s); }
  QN operator-() const { return QN(-p, -s); }
  bool operator==(const QN& o) const { return p == o.p && s == o.s; }
  bool operator!=(const QN& o) const { return !(*this == o); }
  bool operator<(const QN& o) const { return p == o.p ? s < o.s : p < o.p; }  // reference fermion.h:100-101
  bool parity() const { return (p & 1) != 0; }
};

constexpr int kMaxRank = 8;

// Thrown by the host layer; the C ABI turns it into a status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Layout {
  int rank = 0;
  bool qsym = true;
  QN q;
  std::vector<std::vector<QN>> qs;   // [rank][nsec]
  std::vector<std::vector<int>> ds;  // [rank][nsec]
  // stored blocks, ascending tag
  std::vector<int64_t> tag;
  std::vector<int32_t> bidx;         // [nblocks * rank]
  std::vector<int64_t> off;          // [nblocks]
  std::vector<int64_t> bsize;        // [nblocks]
  int64_t nelem = 0;
  uint64_t hash = 0;

  int64_t nblocks() const { return (int64_t)tag.size(); }
  const int32_t* index(int64_t b) const { return bidx.data() + b * rank; }
  int nsec(int i) const { return (int)ds[i].size(); }

  int64_t compute_tag(const int32_t* idx) const {
    int64_t t = 0;
    for (int i = 0; i < rank; ++i) t = t * nsec(i) + idx[i];
    return t;
  }
  bool allowed(const int32_t* idx, bool conj = false) const {
    if (!qsym) return true;
    QN s;
    for (int i = 0; i < rank; ++i) s = s + qs[i][idx[i]];
    return conj ? (-s == -q) : (s == q);
  }
  int64_t find(int64_t t) const {
    auto it = std::lower_bound(tag.begin(), tag.end(), t);
    if (it == tag.end() || *it != t) return -1;
    return it - tag.begin();
  }
  int64_t block_elems(const int32_t* idx) const {
    int64_t n = 1;
    for (int i = 0; i < rank; ++i) n *= ds[i][idx[i]];
    return n;
  }
  bool same_qshape(const Layout& o, bool negate_other = false) const {
    if (rank != o.rank) return false;
    for (int i = 0; i < rank; ++i) {
      if (qs[i].size() != o.qs[i].size()) return false;
      for (size_t k = 0; k < qs[i].size(); ++k)
        if (qs[i][k] != (negate_other ? -o.qs[i][k] : o.qs[i][k])) return false;
    }
    return true;
  }
  bool same_dshape(const Layout& o) const { return ds == o.ds; }
  // Equal block structure (rank, symmetry, total and sector quantum numbers, sector dimensions, stored blocks), so
  // equal offsets: the structure hash is compared first, the fields decide.  Cost O(sectors + blocks).
  bool same_structure(const Layout& o) const {
    if (hash != o.hash || rank != o.rank || qsym != o.qsym || !(q == o.q) || nelem != o.nelem) return false;
    return ds == o.ds && qs == o.qs && tag == o.tag;
  }

  // Sort blocks by tag, drop duplicates, assign canonical offsets and hash the structure.
  void finalize();
  // Set the index structure without blocks.
  void set_shape(int rank_, bool qsym_, QN q_, std::vector<std::vector<QN>> qs_, std::vector<std::vector<int>> ds_);
  // Insert a block (before finalize); duplicates are dropped by finalize.
  void add_block(const int32_t* idx) { bidx.insert(bidx.end(), idx, idx + rank); tag.push_back(compute_tag(idx)); }
};

// Bits of the structure hash kept (all 64 by default).  Fewer bits make distinct structures share hashes, which tests
// use to exercise the structural checks of the plan caches; hashes of layouts finalized earlier are unchanged.
uint64_t structure_hash_mask();
void set_structure_hash_bits(int bits);

// Structure hash combine (FNV-1a 64).
inline uint64_t fnv(uint64_t h, uint64_t v) {
  for (int i = 0; i < 8; ++i) { h ^= (v >> (8 * i)) & 0xff; h *= 1099511628211ull; }
  return h;
}

}  // namespace mpx
