// Regenerates the FCIDUMP of the reference's test integrals (dmrg/tests/intgen.cpp run with N = 6, as in
// dmrg/tests/check.sh:3-4): std::mt19937 with its default seed, uniform_real_distribution<double>(0, 1), one draw per
// listed integral in the generator's order, printed as it prints them.  The reference DMRG log for these integrals
// (dmrg/tests/dmrg.out:1735) gives the ground energy -5.4268020092153462.
//
//   g++ -O2 -o /tmp/intgen tests/golden/make_intgen_fcidump.cpp && /tmp/intgen 6 > tests/golden/intgen6.FCIDUMP
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>

namespace {

struct Emitter {
  std::mt19937 gen;
  std::uniform_real_distribution<double> u{0.0, 1.0};
  void put(double scale, int a, int b, int c, int d) {
    const double x = scale * u(gen);
    std::printf("%24.16e%4d%4d%4d%4d\n", x, a, b, c, d);
  }
};

}  // namespace

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 6;
  std::printf(" &FCI NORB=%3d,NELEC=%3d,MS2= 0,\n  ORBSYM=", n, n);
  for (int i = 0; i < n; ++i) std::printf("1,");
  std::printf("\n  ISYM=1,\n /\n");
  Emitter e;
  // two-electron integrals, 1-based orbital labels; nesting i > j > k > l
  for (int i = 1; i <= n; ++i) {
    e.put(1.0, i, i, i, i);
    for (int j = 1; j < i; ++j) {
      e.put(0.5, i, i, i, j);
      e.put(0.2, i, i, j, j);
      e.put(0.2, i, j, i, j);
      e.put(0.5, i, j, j, j);
      for (int k = 1; k < j; ++k) {
        e.put(0.2, i, i, j, k);
        e.put(0.2, i, j, i, k);
        e.put(0.2, i, j, j, k);
        e.put(0.2, j, j, i, k);
        e.put(0.2, i, j, k, k);
        e.put(0.2, i, k, j, k);
        for (int l = 1; l < k; ++l) {
          e.put(0.1, i, j, k, l);
          e.put(0.1, i, k, j, l);
          e.put(0.1, i, l, j, k);
        }
      }
    }
  }
  // one-electron integrals: negative diagonal, positive off-diagonal
  for (int i = 1; i <= n; ++i) {
    e.put(-1.0, i, i, 0, 0);
    for (int j = 1; j < i; ++j) e.put(1.0, i, j, 0, 0);
  }
  std::printf("%24.16e%4d%4d%4d%4d\n", 0.0, 0, 0, 0, 0);
  return 0;
}
