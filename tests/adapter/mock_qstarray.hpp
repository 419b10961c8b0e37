// Minimal stand-in for the btas::QSTArray API subset the drivers use (q, qshape, dshape, begin/end over
// (tag, shared_ptr<block>) in ascending tag order, index, size, resize, reserve, clear; tags run over the sector
// counts of qshape, as STArray.h:338-348), used to compile and exercise
// include/mpsxx_amd/mpsxx.hpp without the reference tree.  Quantum numbers: (particle, 2Sz) with p(), s().
#pragma once

#include <array>
#include <cstddef>
#include <map>
#include <memory>
#include <vector>

namespace mock {

struct Fermion {
  int p_ = 0, s_ = 0;
  Fermion() = default;
  Fermion(int p, int s) : p_(p), s_(s) {}
  int p() const { return p_; }
  int s() const { return s_; }
};

struct Block {
  std::vector<double> v;
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  size_t size() const { return v.size(); }
};

template <typename T, size_t N, class Q>
class QSTArray {
 public:
  typedef std::map<int, std::shared_ptr<Block>> map_t;
  typedef typename map_t::iterator iterator;
  typedef typename map_t::const_iterator const_iterator;

  const Q& q() const { return q_; }
  const std::array<std::vector<Q>, N>& qshape() const { return qs_; }
  const std::vector<Q>& qshape(int i) const { return qs_[i]; }
  const std::array<std::vector<int>, N>& dshape() const { return ds_; }
  const std::vector<int>& dshape(int i) const { return ds_[i]; }
  const_iterator begin() const { return blocks_.begin(); }
  const_iterator end() const { return blocks_.end(); }
  iterator end() { return blocks_.end(); }
  size_t size() const {
    size_t n = 0;
    for (auto& b : blocks_) n += b.second->size();
    return n;
  }
  std::array<int, N> index(int tag) const {
    std::array<int, N> idx;
    for (int i = (int)N - 1; i >= 0; --i) {
      idx[i] = tag % (int)qs_[i].size();
      tag /= (int)qs_[i].size();
    }
    return idx;
  }
  int tag(const std::array<int, N>& idx) const {
    int t = 0;
    for (size_t i = 0; i < N; ++i) t = t * (int)qs_[i].size() + idx[i];
    return t;
  }
  bool allowed(const std::array<int, N>& idx) const {
    int p = 0, s = 0;
    for (size_t i = 0; i < N; ++i) {
      p += qs_[i][idx[i]].p();
      s += qs_[i][idx[i]].s();
    }
    return p == q_.p() && s == q_.s();
  }
  // quantum numbers only, dense shapes unset (QSTArray(q, qshape))
  void resize(const Q& q, const std::array<std::vector<Q>, N>& qs) {
    q_ = q;
    qs_ = qs;
    for (auto& x : ds_) x.clear();
    blocks_.clear();
  }
  void resize(const Q& q, const std::array<std::vector<Q>, N>& qs, const std::array<std::vector<int>, N>& ds, bool) {
    q_ = q;
    qs_ = qs;
    ds_ = ds;
    blocks_.clear();
  }
  iterator reserve(int tag) {
    const std::array<int, N> idx = index(tag);
    if (!allowed(idx)) return blocks_.end();
    auto b = std::make_shared<Block>();
    size_t n = 1;
    for (size_t i = 0; i < N; ++i) n *= (size_t)ds_[i][idx[i]];
    b->v.assign(n, 0.0);
    return blocks_.emplace(tag, b).first;
  }
  void clear() {
    blocks_.clear();
    for (auto& x : qs_) x.clear();
    for (auto& x : ds_) x.clear();
    q_ = Q();
  }

 private:
  Q q_;
  std::array<std::vector<Q>, N> qs_;
  std::array<std::vector<int>, N> ds_;
  map_t blocks_;
};

}  // namespace mock
