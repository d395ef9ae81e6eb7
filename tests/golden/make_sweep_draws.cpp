// Writes sweep_draws_20261202.txt: the first 32 (r_m, L, ln N) draws of the
// batched horn sweep from std::mt19937_64(20261202) with libstdc++'s
// uniform_real_distribution<double>, the fixture that pins the Python restatement
// in hornsculpt_amd/variants.py.
//   g++ -O2 -o /tmp/make_sweep_draws tests/golden/make_sweep_draws.cpp &&
//   /tmp/make_sweep_draws > tests/golden/sweep_draws_20261202.txt
#include <cmath>
#include <random>
#include <cstdio>
int main() {
    std::mt19937_64 g(20261202);
    std::uniform_real_distribution<double> rm(0.080, 0.200), len(0.150, 0.450), nt(std::log(4000.0), std::log(60000.0));
    for (int i = 0; i < 32; ++i) {
        double a = rm(g), b = len(g), c = nt(g);
        std::printf("%.17g %.17g %.17g\n", a, b, c);
    }
}
