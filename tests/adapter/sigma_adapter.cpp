// Calls the drop-in templates of include/mpsxx_amd/mpsxx.hpp with QSTArray-like host containers, as the reference's
// site optimizer does (src/optimize.hpp:60-75: diag built from the wavefunction's quantum numbers, sigma vectors into
// an empty container, then renormalize).  Inputs come from a simple text dump written by the test; the outputs
// (sigma accumulated twice, diag, renormalized left operator) are dumped the same way.
//   format per array: rank, q.p q.s, per index: nsec then (p s d) per sector, nblocks, per block: indices then values
#include <cstdio>
#include <fstream>
#include <iostream>

#include "mock_qstarray.hpp"
#include "mpsxx_amd/mpsxx.hpp"

using Q = mock::Fermion;

template <size_t N>
bool read_array(std::istream& in, mock::QSTArray<double, N, Q>& a) {
  size_t rank;
  int qp, qs;
  if (!(in >> rank >> qp >> qs) || rank != N) return false;
  std::array<std::vector<Q>, N> qsh;
  std::array<std::vector<int>, N> dsh;
  for (size_t i = 0; i < N; ++i) {
    int ns;
    in >> ns;
    for (int k = 0; k < ns; ++k) {
      int p, s, d;
      in >> p >> s >> d;
      qsh[i].push_back(Q(p, s));
      dsh[i].push_back(d);
    }
  }
  a.resize(Q(qp, qs), qsh, dsh, false);
  long nb;
  in >> nb;
  for (long b = 0; b < nb; ++b) {
    std::array<int, N> idx;
    for (size_t i = 0; i < N; ++i) in >> idx[i];
    auto it = a.reserve(a.tag(idx));
    if (it == a.end()) return false;
    for (size_t e = 0; e < it->second->size(); ++e) in >> it->second->data()[e];
  }
  return true;
}

template <size_t N>
void write_array(std::ostream& out, const mock::QSTArray<double, N, Q>& a) {
  out << N << " " << a.q().p() << " " << a.q().s() << "\n";
  for (size_t i = 0; i < N; ++i) {
    out << a.qshape(i).size();
    for (size_t k = 0; k < a.qshape(i).size(); ++k)
      out << " " << a.qshape(i)[k].p() << " " << a.qshape(i)[k].s() << " " << a.dshape(i)[k];
    out << "\n";
  }
  long nb = 0;
  for (auto it = a.begin(); it != a.end(); ++it) ++nb;
  out << nb << "\n";
  out.precision(17);
  for (auto it = a.begin(); it != a.end(); ++it) {
    auto idx = a.index(it->first);
    for (size_t i = 0; i < N; ++i) out << idx[i] << " ";
    for (size_t e = 0; e < it->second->size(); ++e) out << it->second->data()[e] << " ";
    out << "\n";
  }
}

// Every other drop-in overload, instantiated so that the compile test checks it against the container API (their
// device paths are covered through the C ABI by the other tests).
[[maybe_unused]] void instantiate_other_forms() {
  mock::QSTArray<double, 4, Q> w4, s4;
  mock::QSTArray<double, 3, Q> o3, a3, b3, c3;
  mock::QSTArray<double, 2, Q> s2, t2, u2, v2;
  mpsxx::compute_sigma_vector(w4, o3, a3, b3, c3);                  // one-site
  mpsxx::compute_sigma_vector(o3, a3, s2, t2);                      // merged
  mpsxx::compute_sigma_vector(s2, t2, b3, c3);                      // overlap, 3-index
  mpsxx::compute_sigma_vector(s2, t2, w4, s4);                      // overlap, 4-index
  mpsxx::compute_sigma_vector(s2, t2, u2, v2);                      // overlap, 2-index
  mpsxx::compute_sigma_vector_conj(w4, o3, a3, b3, c3);
  mpsxx::compute_diagonal_elements(w4, o3, a3, b3);                 // one-site
  mpsxx::compute_diagonal_elements(o3, a3, s2);                     // merged
  std::vector<mock::QSTArray<double, 4, Q>> wg;
  std::vector<mock::QSTArray<double, 3, Q>> og, ag;
  mpsxx::compute_diagonal_elements(wg, og, ag, b3);
  mpsxx::compute_diagonal_elements(og, ag, s2);
  mpsxx::renormalize(true, s2, a3, b3, t2);                         // overlap
  mpsxx::renormalize(false, o3, s2, t2, a3);                        // merged operator
  mpsxx::renormalize(true, s2, t2, u2, v2);                         // merged overlap
  mpsxx::canonicalize(true, c3, a3, b3, 10, 0.0);                   // one-site
  mpsxx::canonicalize(false, u2, s2, t2, 10, 1e-3);                 // merged
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: sigma_adapter in.txt out.txt\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  mock::QSTArray<double, 4, Q> lmpo, rmpo, wfn, sgv;
  mock::QSTArray<double, 4, Q> diag, diag2;
  mock::QSTArray<double, 3, Q> lopr, ropr, ket, lnext, lcan, rcan;
  mock::QSTArray<double, 2, Q> lsov, rsov;
  std::vector<std::vector<mock::QSTArray<double, 3, Q>>> mlH, mrH;
  std::vector<mock::QSTArray<double, 2, Q>> mlS, mrS;
  if (!read_array(in, lmpo) || !read_array(in, rmpo) || !read_array(in, lopr) || !read_array(in, ropr) ||
      !read_array(in, wfn) || !read_array(in, ket) || !read_array(in, lsov) || !read_array(in, rsov)) {
    std::cerr << "bad input\n";
    return 3;
  }
  try {
    mpsxx::compute_sigma_vector(lmpo, rmpo, lopr, ropr, wfn, sgv);  // sgv empty: shaped from the inputs
    mpsxx::compute_sigma_vector(lmpo, rmpo, lopr, ropr, wfn, sgv);  // accumulates: 2 H psi
    diag.resize(wfn.q(), wfn.qshape(), wfn.dshape(), false);         // structure only, no blocks
    mpsxx::compute_diagonal_elements(lmpo, rmpo, lopr, ropr, diag);
    mpsxx::renormalize(true, lmpo, lopr, ket, ket, lnext);
    // group-sum form with the same operators in two groups: 2 x diag
    std::vector<mock::QSTArray<double, 4, Q>> lg{lmpo, lmpo}, rg{rmpo, rmpo};
    std::vector<mock::QSTArray<double, 3, Q>> log{lopr, lopr}, rog{ropr, ropr};
    diag2.resize(wfn.q(), wfn.qshape());  // as QSTArray(q, qshape): no dense shapes

    mpsxx::compute_diagonal_elements(lg, rg, log, rog, diag2);
    // merged operators of one root and two groups (optimize.hpp:273-368), both directions
    std::vector<std::vector<mock::QSTArray<double, 3, Q>>> xl{{lopr, lopr}}, xr{{ropr, ropr}};
    std::vector<mock::QSTArray<double, 2, Q>> sl{lsov}, sr{rsov};
    mpsxx::compute_merged_operator(true, lg, xl, mlH, sl, mlS);
    mpsxx::compute_merged_operator(false, rg, xr, mrH, sr, mrS);
    // two-site canonicalize with noise, as optimize.hpp:210 calls it (wfnc carries the noise afterwards)
    mpsxx::canonicalize(true, wfn, lcan, rcan, 12, 1e-3);
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 4;
  }
  std::ofstream out(argv[2]);
  write_array(out, sgv);
  write_array(out, diag);
  write_array(out, lnext);
  write_array(out, diag2);
  write_array(out, mlH[0][0]);
  write_array(out, mlH[0][1]);
  write_array(out, mlS[0]);
  write_array(out, mrH[0][1]);
  write_array(out, mrS[0]);
  write_array(out, lcan);
  write_array(out, rcan);
  write_array(out, wfn);
  return 0;
}
