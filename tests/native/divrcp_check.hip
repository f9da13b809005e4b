// Host-side check that div_rcp (division by a precomputed reciprocal with
// one fused-multiply-add correction) equals IEEE division bit for bit on the
// operands the stencils use: for every cell of the given axes, numerators
// x - g[s] and x - (g[m-1] + g[m]) / 2 at gridlines, their neighbouring
// doubles, midpoints and random points of the stencil's range, over the cell
// width (inv) and both pressure-node half-spacings (rc[0], rc[1]); plus random
// operand pairs over wide magnitude ranges.  Prints "checks mismatches".
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "../../velo2cycles_amd/csrc/v2c_search.hpp"

using namespace v2c;

static long checks = 0, bad = 0;

static void check(double a, double b, double rb)
{
    const double q = a / b, r = div_rcp(a, b, rb);
    checks++;
    if (!(q == r || (q != q && r != r))) {
        if (bad < 10) printf("# mismatch a=%a b=%a q=%a r=%a\n", a, b, q, r);
        bad++;
    }
}

int main(int argc, char **argv)
{
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int f = 1; f < argc; f++) {
        FILE *fp = fopen(argv[f], "rb");
        if (!fp) { printf("cannot open %s\n", argv[f]); return 2; }
        std::vector<double> g;
        double v;
        while (fread(&v, sizeof v, 1, fp) == 1) g.push_back(v);
        fclose(fp);
        const int n = (int)g.size();
        for (int s = 0; s + 1 < n; s++) {
            const Axis4 A = axis_cell(g.data(), n, s);
            std::vector<double> xs = {A.g[1], A.g[2], (A.g[1] + A.g[2]) / 2.0,
                                      std::nextafter(A.g[1], INFINITY), std::nextafter(A.g[2], -INFINITY)};
            for (int k = 0; k < 40; k++) xs.push_back(A.g[1] + (A.g[2] - A.g[1]) * U(rng));
            for (double x : xs) {
                check(x - A.g[1], A.g[2] - A.g[1], A.inv);
                for (int d = 0; d < 2; d++) {
                    if (s - 1 + d < 0 || s + 1 + d >= n) continue;
                    const double a0 = A.g[d], a1 = A.g[d + 1], a2 = A.g[d + 2];
                    check(x - (a0 + a1) / 2.0, (a2 - a0) / 2.0, A.rc[d]);
                }
            }
        }
    }
    for (long k = 0; k < 20000000; k++) {
        const double b = std::ldexp(1.0 + U(rng), (int)(U(rng) * 80) - 40);
        const double a = b * 2.0 * (U(rng) - 0.5) * (k % 7 == 0 ? 1e-9 : 1.0);
        check(a, b, 1.0 / b);
    }
    printf("%ld %ld\n", checks, bad);
    return bad ? 1 : 0;
}
