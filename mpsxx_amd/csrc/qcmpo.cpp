// Quantum-chemistry MPO generator (host, FP64): H = sum_{pq,s} h_pq a+_ps a_qs
//                                                 + 1/2 sum_{pqrs,st} (pq|rs) a+_ps a+_rt a_st a_qs + ecore
// for one spatial orbital per site (site basis {0, a, b, ab}).
//
// Reference: src/gen_qc_operators.cpp:18-202 with the boundary operator sets of src/boundary_opinfo.cpp:47-140.
// The bond operators are the reference's: identity, the complete left Hamiltonian, the four single operators of
// every left site (normal), the four complementary single operators of every right site, and the pair operators of
// the smaller side (normal pairs on the left while it is the smaller block, complementary pairs otherwise: the
// reference's swap-sweep mode; without it pairs stay on the left), so D_w = 2 + 4K + 8n^2 - 2n with n = min(L, R).
//
// Construction (own design): every Hamiltonian term is written in the Jordan-Wigner tensor-product form
// coeff * O_0 (x) O_1 (x) ... (x) O_{K-1}, O_s = product, in operator order, of the local factor each operator puts on
// site s (its local matrix on its own site, the parity P on the sites left of it, I right of it).  At each bond a
// term sits in one channel, chosen by how many of its operators are left of the bond: none -> I, all -> H, one (or
// two, while pairs are kept on the left) -> the channel of its left factors ("normal", coefficient not yet applied),
// otherwise -> the channel of its right factors ("complementary", coefficient applied).  The MPO tensor of site k
// maps channels of bond k-1 to channels of bond k: transitions that keep a term normal or complementary copy the
// site factor, transitions that apply the coefficient accumulate coeff * factor.  Channel keys use sign-normalised
// local factors so that terms differing by an operator reordering share a channel.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

#include "mpsxx_amd.h"
#include "qcmpo.hpp"

namespace mpx {

namespace {

typedef std::array<double, 16> Mat;  // (bra, ket) row-major over {0, a, b, ab}

Mat mat_zero() { Mat m; m.fill(0.0); return m; }
Mat mat_id() { Mat m = mat_zero(); for (int i = 0; i < 4; ++i) m[i * 5] = 1.0; return m; }
Mat mat_parity() { Mat m = mat_zero(); m[0] = 1.0; m[5] = -1.0; m[10] = -1.0; m[15] = 1.0; return m; }
Mat mat_mul(const Mat& a, const Mat& b) {
  Mat c = mat_zero();
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) {
      const double x = a[i * 4 + k];
      if (x == 0.0) continue;
      for (int j = 0; j < 4; ++j) c[i * 4 + j] += x * b[k * 4 + j];
    }
  return c;
}
// local operators: spin 0 = alpha, 1 = beta; the beta operators pass the alpha orbital of their site
Mat local_op(int spin, bool dag) {
  Mat m = mat_zero();
  if (spin == 0) { m[1 * 4 + 0] = 1.0; m[3 * 4 + 2] = 1.0; }   // a+_alpha
  else { m[2 * 4 + 0] = 1.0; m[3 * 4 + 1] = -1.0; }            // a+_beta
  if (!dag) { Mat t = mat_zero(); for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) t[i * 4 + j] = m[j * 4 + i]; m = t; }
  return m;
}

// Interning of sign-normalised local matrices (first non-zero entry positive).
class MatTable {
 public:
  // returns (id, sign) with m = sign * table[id]
  std::pair<int, int> intern(const Mat& m) {
    int sign = 1;
    for (double x : m)
      if (x != 0.0) { sign = x > 0 ? 1 : -1; break; }
    Mat c = m;
    if (sign < 0) for (double& x : c) x = -x;
    for (size_t i = 0; i < mats_.size(); ++i)
      if (mats_[i] == c) return {(int)i, sign};
    mats_.push_back(c);
    return {(int)mats_.size() - 1, sign};
  }
  const Mat& operator[](int id) const { return mats_[id]; }
  int size() const { return (int)mats_.size(); }

 private:
  std::vector<Mat> mats_;
};

struct Term {
  double coeff;
  int p1, q1;            // one-body term: orbitals (p, q); two-body: -1
  int n;                 // 2 or 4 operators
  int nsite;             // distinct operator sites, ascending
  int site[4];
  int id[4], sgn[4];     // sign-normalised local factor of each operator site
  int cnt[4];            // operators on each operator site
  QN q[4];               // quantum number of the operators on each operator site
};

enum Cat : uint8_t { CAT_I = 0, CAT_H = 1, CAT_N = 2, CAT_C = 3 };

struct Key {
  uint8_t cat, np;
  uint16_t sid[4];  // (site << 10) | id
  bool operator==(const Key& o) const {
    if (cat != o.cat || np != o.np) return false;
    for (int i = 0; i < np; ++i)
      if (sid[i] != o.sid[i]) return false;
    return true;
  }
  bool operator<(const Key& o) const {
    if (cat != o.cat) return cat < o.cat;
    for (int i = 0; i < std::min(np, o.np); ++i)
      if (sid[i] != o.sid[i]) return sid[i] < o.sid[i];
    return np < o.np;
  }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = 1469598103934665603ull ^ k.cat ^ ((uint64_t)k.np << 8);
    for (int i = 0; i < k.np; ++i) h = (h ^ k.sid[i]) * 1099511628211ull;
    return (size_t)h;
  }
};

struct Chan {
  Cat cat;
  int sign;  // sign product of the keyed factors (left for normal, right for complementary)
  QN q;      // quantum number of the left operator
  Key key;
};

// channel of term t at bond b (b = -1: left boundary; b = K-1: right boundary)
Chan channel(const Term& t, int b, int K, bool swap_sweep) {
  Chan c;
  int nl = 0, nsl = 0;
  for (int j = 0; j < t.nsite; ++j)
    if (t.site[j] <= b) { nl += t.cnt[j]; ++nsl; }
  c.q = QN();
  for (int j = 0; j < nsl; ++j) c.q = c.q + t.q[j];
  c.key.np = 0;
  c.sign = 1;
  if (nl == 0) { c.cat = CAT_I; c.key.cat = CAT_I; return c; }
  if (nl == t.n) { c.cat = CAT_H; c.key.cat = CAT_H; return c; }
  const bool pairs_left = !swap_sweep || (b + 1 < K - b - 1);
  const bool normal = nl == 1 || (nl == 2 && pairs_left);
  c.cat = normal ? CAT_N : CAT_C;
  c.key.cat = c.cat;
  const int j0 = normal ? 0 : nsl, j1 = normal ? nsl : t.nsite;
  for (int j = j0; j < j1; ++j) {
    c.key.sid[c.key.np++] = (uint16_t)((t.site[j] << 10) | t.id[j]);
    c.sign *= t.sgn[j];
  }
  return c;
}

}  // namespace

QcMpo build_qc_mpo(int K, const double* h, const double* v, double ecore, bool swap_sweep) {
  if (K < 1 || K > 63) throw Error(MPSXX_ERR_ARG, "qc mpo: 1 <= norb <= 63 required");
  MatTable tab;
  const Mat P = mat_parity(), I = mat_id();
  // canonical spin-orbital terms: a+_a a+_b a_c a_d with a < b, c < d (two-body), a+_a a_c (one-body)
  const int S = 2 * K;
  std::unordered_map<uint64_t, double> two;
  two.reserve((size_t)K * K * K * K);
  auto so = [](int p, int s) { return 2 * p + s; };
  auto pack = [](int a, int b, int c, int d) {
    return ((uint64_t)a << 48) | ((uint64_t)b << 32) | ((uint64_t)c << 16) | (uint64_t)d;
  };
  for (int p = 0; p < K; ++p)
    for (int q = 0; q < K; ++q)
      for (int r = 0; r < K; ++r)
        for (int s = 0; s < K; ++s) {
          const double val = v[(((size_t)p * K + q) * K + r) * K + s];
          if (val == 0.0) continue;
          for (int sg = 0; sg < 2; ++sg)
            for (int tg = 0; tg < 2; ++tg) {
              int a = so(p, sg), b = so(r, tg), c = so(s, tg), d = so(q, sg);
              if (a == b || c == d) continue;
              double x = 0.5 * val;
              if (a > b) { std::swap(a, b); x = -x; }
              if (c > d) { std::swap(c, d); x = -x; }
              two[pack(a, b, c, d)] += x;
            }
        }
  std::vector<Term> terms;
  auto add_term = [&](double coeff, int n, const int* sos, const bool* dag) {
    Term t;
    t.coeff = coeff;
    t.p1 = n == 2 ? sos[0] / 2 : -1;
    t.q1 = n == 2 ? sos[1] / 2 : -1;
    t.n = n;
    t.nsite = 0;
    int sites[4];
    for (int i = 0; i < n; ++i) sites[i] = sos[i] / 2;
    std::vector<int> us(sites, sites + n);
    std::sort(us.begin(), us.end());
    us.erase(std::unique(us.begin(), us.end()), us.end());
    for (int u : us) {
      // factor of the term on site u: operators in order, P for every operator right of u, I left of u
      Mat m = I;
      int cnt = 0;
      QN q;
      for (int i = 0; i < n; ++i) {
        if (sites[i] == u) {
          m = mat_mul(m, local_op(sos[i] % 2, dag[i]));
          ++cnt;
          const int dp = dag[i] ? 1 : -1, ds = (sos[i] % 2 == 0) ? 1 : -1;
          q = q + QN(dp, dp * ds);
        } else if (sites[i] > u) {
          m = mat_mul(m, P);
        }
      }
      bool nz = false;
      for (double x : m) nz |= (x != 0.0);
      if (!nz) return;  // e.g. a+_a a+_a on one site
      auto is = tab.intern(m);
      t.site[t.nsite] = u;
      t.id[t.nsite] = is.first;
      t.sgn[t.nsite] = is.second;
      t.cnt[t.nsite] = cnt;
      t.q[t.nsite] = q;
      ++t.nsite;
    }
    terms.push_back(t);
  };
  for (int p = 0; p < K; ++p)
    for (int q = 0; q < K; ++q) {
      const double val = h[(size_t)p * K + q];
      if (val == 0.0) continue;
      for (int sg = 0; sg < 2; ++sg) {
        const int sos[2] = {so(p, sg), so(q, sg)};
        const bool dag[2] = {true, false};
        add_term(val, 2, sos, dag);
      }
    }
  {
    // deterministic term order (ascending packed index)
    std::vector<std::pair<uint64_t, double>> tv(two.begin(), two.end());
    std::sort(tv.begin(), tv.end());
    for (const auto& e : tv) {
      if (e.second == 0.0) continue;
      const int sos[4] = {(int)(e.first >> 48), (int)((e.first >> 32) & 0xffff), (int)((e.first >> 16) & 0xffff),
                          (int)(e.first & 0xffff)};
      const bool dag[4] = {true, true, false, false};
      add_term(e.second, 4, sos, dag);
    }
  }
  (void)S;

  // channels of every bond (bond b = after site b; bond -1 and K-1 are the boundaries)
  std::vector<std::unordered_map<Key, int, KeyHash>> chan_id(K + 1);
  std::vector<std::vector<Chan>> chans(K + 1);
  for (int b = -1; b < K; ++b) {
    std::map<Key, Chan> ordered;
    if (b == -1) {
      Chan c{CAT_I, 1, QN(), Key{CAT_I, 0, {0, 0, 0, 0}}};
      ordered.emplace(c.key, c);
    } else if (b == K - 1) {
      Chan c{CAT_H, 1, QN(), Key{CAT_H, 0, {0, 0, 0, 0}}};
      ordered.emplace(c.key, c);
    } else {
      std::unordered_map<Key, Chan, KeyHash> seen;
      for (const Term& t : terms) {
        Chan c = channel(t, b, K, swap_sweep);
        seen.emplace(c.key, c);
      }
      if (ecore != 0.0) {
        Chan hc{CAT_H, 1, QN(), Key{CAT_H, 0, {0, 0, 0, 0}}};
        seen.emplace(hc.key, hc);
      }
      for (auto& kv : seen) ordered.emplace(kv.first, kv.second);
    }
    auto& ids = chan_id[b + 1];
    for (auto& kv : ordered) {
      ids.emplace(kv.first, (int)chans[b + 1].size());
      chans[b + 1].push_back(kv.second);
    }
  }

  QcMpo out;
  out.K = K;
  out.ecore = ecore;
  out.sites.resize(K);
  const int idP = tab.intern(P).first, idI = tab.intern(I).first;
  for (int k = 0; k < K; ++k) {
    QcMpoSite& site = out.sites[k];
    const auto& cl = chans[k];
    const auto& cr = chans[k + 1];
    for (const Chan& c : cl) site.lq.push_back(c.q);
    for (const Chan& c : cr) site.rq.push_back(c.q);
    std::unordered_map<uint64_t, Mat> W;
    auto lr = [](int l, int r) { return ((uint64_t)(uint32_t)l << 32) | (uint32_t)r; };
    auto set = [&](int l, int r, const Mat& m) {
      auto it = W.find(lr(l, r));
      if (it == W.end()) W.emplace(lr(l, r), m);
      else if (it->second != m) throw Error(MPSXX_ERR_ARG, "qc mpo: inconsistent channel transition");
    };
    auto add = [&](int l, int r, double x, const Mat& m) {
      auto it = W.find(lr(l, r));
      if (it == W.end()) it = W.emplace(lr(l, r), mat_zero()).first;
      for (int e = 0; e < 16; ++e) it->second[e] += x * m[e];
    };
    auto find = [&](const std::unordered_map<Key, int, KeyHash>& ids, const Key& key) {
      auto it = ids.find(key);
      return it == ids.end() ? -1 : it->second;
    };
    const Key kI{CAT_I, 0, {0, 0, 0, 0}}, kH{CAT_H, 0, {0, 0, 0, 0}};
    const int Il = find(chan_id[k], kI), Ir = find(chan_id[k + 1], kI);
    const int Hl = find(chan_id[k], kH), Hr = find(chan_id[k + 1], kH);
    if (Il >= 0 && Ir >= 0) set(Il, Ir, I);
    if (Hl >= 0 && Hr >= 0) set(Hl, Hr, I);
    if (k == 0 && ecore != 0.0) add(0, Hr, ecore, I);
    for (const Term& t : terms) {
      const Chan cp = channel(t, k - 1, K, swap_sweep);
      const Chan cn = channel(t, k, K, swap_sweep);
      const int l = find(chan_id[k], cp.key), r = find(chan_id[k + 1], cn.key);
      if (l < 0 || r < 0) throw Error(MPSXX_ERR_ARG, "qc mpo: missing channel");
      if (cp.cat == CAT_I && cn.cat == CAT_I) continue;  // identity transition set above
      if (cp.cat == CAT_H) continue;
      // the term's factor on site k: Ok = ok_sgn * canon, canon = table[ok_id]
      int ok_id = idI, ok_sgn = 1;
      int j = 0;
      while (j < t.nsite && t.site[j] < k) ++j;
      if (j < t.nsite && t.site[j] == k) {
        ok_id = t.id[j];
        ok_sgn = t.sgn[j];
      } else {
        int nl = 0;
        for (int i = 0; i < t.nsite; ++i)
          if (t.site[i] < k) nl += t.cnt[i];
        if (nl % 2) ok_id = idP;
      }
      const Mat& canon = tab[ok_id];
      Mat Ok = canon;
      if (ok_sgn < 0) for (double& x : Ok) x = -x;
      const int sprev = (cp.cat == CAT_N) ? cp.sign : 1;
      auto record = [&](double mult) {
        out.apps.push_back({k, l, r, ok_id, mult * ok_sgn, t.p1, t.q1, t.coeff});
      };
      if (cn.cat == CAT_N) {
        set(l, r, canon);
      } else if (cn.cat == CAT_C) {
        if (cp.cat == CAT_C) {
          set(l, r, canon);
        } else {
          add(l, r, t.coeff * cn.sign * sprev, Ok);
          record(cn.sign * sprev);
        }
      } else {  // H
        if (cp.cat == CAT_C) {
          set(l, r, canon);
        } else {
          add(l, r, t.coeff * sprev, Ok);
          record(sprev);
        }
      }
    }
    std::vector<std::pair<uint64_t, const Mat*>> ent;
    for (auto& kv : W) ent.push_back({kv.first, &kv.second});
    std::sort(ent.begin(), ent.end(),
              [](const std::pair<uint64_t, const Mat*>& a, const std::pair<uint64_t, const Mat*>& b) {
                return a.first < b.first;
              });
    for (auto& e : ent) {
      bool nz = false;
      for (double x : *e.second) nz |= (x != 0.0);
      if (!nz) continue;
      site.wl.push_back((int)(e.first >> 32));
      site.wr.push_back((int)(e.first & 0xffffffffu));
      site.mat.push_back(*e.second);
    }
  }
  for (int i = 0; i < tab.size(); ++i) out.mats.push_back(tab[i]);
  return out;
}

void qc_mpo_gradient(const QcMpo& m, const std::vector<const Tensor*>& G, double* dh, double* dlambda, double* energy,
                     hipStream_t s) {
  const int K = m.K;
  if ((int)G.size() != K) throw Error(MPSXX_ERR_ARG, "qc gradient: one gradient tensor per site expected");
  // host copies of the gradient tensors: (wl, wr) -> 4x4 (bra, ket)
  std::vector<std::unordered_map<uint64_t, Mat>> g(K);
  for (int k = 0; k < K; ++k) {
    const Tensor& t = *G[k];
    if (t.empty) continue;
    if (t.L.rank != 4) throw Error(MPSXX_ERR_ARG, "qc gradient: gradient tensors are 4-index");
    std::vector<double> x(std::max<int64_t>(t.L.nelem, 1));
    hip_check(hipMemcpyAsync(x.data(), t.ptr(), t.L.nelem * sizeof(double), hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "download sync");
    for (int64_t b = 0; b < t.L.nblocks(); ++b) {
      const int32_t* idx = t.L.index(b);
      if (t.L.bsize[b] != 1) throw Error(MPSXX_ERR_SHAPE, "qc gradient: blocks of an MPO gradient are 1x1x1x1");
      const uint64_t key = ((uint64_t)(uint32_t)idx[0] << 32) | (uint32_t)idx[3];
      auto it = g[k].find(key);
      if (it == g[k].end()) it = g[k].emplace(key, mat_zero()).first;
      it->second[idx[1] * 4 + idx[2]] = x[t.L.off[b]];
    }
  }
  std::fill(dh, dh + (size_t)K * K, 0.0);
  double dl = 0.0, e = m.ecore;
  for (const QcTermApp& a : m.apps) {
    auto it = g[a.site].find(((uint64_t)(uint32_t)a.wl << 32) | (uint32_t)a.wr);
    if (it == g[a.site].end()) continue;
    double d = 0.0;
    const Mat& f = m.mats[a.mat];
    for (int e16 = 0; e16 < 16; ++e16) d += it->second[e16] * f[e16];
    d *= a.mult;
    if (a.p >= 0) dh[(size_t)a.p * K + a.q] += d;
    else dl += a.coeff * d;
    e += a.coeff * d;
  }
  if (dlambda) *dlambda = dl;
  if (energy) *energy = e;
}

void qc_mpo_tensor(const QcMpoSite& site, Tensor& out, hipStream_t s) {
  static const QN sq[4] = {QN(0, 0), QN(1, 1), QN(1, -1), QN(2, 0)};
  std::vector<std::vector<QN>> qs(4);
  std::vector<std::vector<int>> ds(4);
  qs[0] = site.lq;
  for (int i = 0; i < 4; ++i) { qs[1].push_back(sq[i]); qs[2].push_back(-sq[i]); }
  for (const QN& q : site.rq) qs[3].push_back(-q);
  ds[0].assign(site.lq.size(), 1);
  ds[1].assign(4, 1);
  ds[2].assign(4, 1);
  ds[3].assign(site.rq.size(), 1);
  Layout L;
  L.set_shape(4, true, QN(), qs, ds);
  struct E { int32_t idx[4]; double v; };
  std::vector<E> el;
  for (size_t e = 0; e < site.mat.size(); ++e)
    for (int b = 0; b < 4; ++b)
      for (int k = 0; k < 4; ++k) {
        const double x = site.mat[e][b * 4 + k];
        if (x == 0.0) continue;
        E en{{site.wl[e], b, k, site.wr[e]}, x};
        if (!L.allowed(en.idx)) throw Error(MPSXX_ERR_QN, "qc mpo: entry violates quantum-number conservation");
        L.add_block(en.idx);
        el.push_back(en);
      }
  L.finalize();
  std::vector<double> data(std::max<int64_t>(L.nelem, 1), 0.0);
  for (const E& en : el) data[L.off[L.find(L.compute_tag(en.idx))]] = en.v;
  out.set_layout(L, s);
  if (L.nelem > 0)
    hip_check(hipMemcpyAsync(out.ptr(), data.data(), L.nelem * sizeof(double), hipMemcpyHostToDevice, s), "upload");
  hip_check(hipStreamSynchronize(s), "upload sync");
}

}  // namespace mpx
