// Host-side check of the table-driven cell search (cell index and gridline
// record) against the reference bisection.  Reads axes from argv files (raw float64), probes gridlines, their
// neighbouring doubles, midpoints, random points, infinities and NaN; prints
// "probes mismatches" per axis.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "../../velo2cycles_amd/csrc/v2c_search.hpp"

using namespace v2c;

int main(int argc, char **argv)
{
    int bad_total = 0;
    for (int f = 1; f < argc; f++) {
        FILE *fp = fopen(argv[f], "rb");
        if (!fp) { printf("cannot open %s\n", argv[f]); return 2; }
        std::vector<double> g;
        double v;
        while (fread(&v, sizeof v, 1, fp) == 1) g.push_back(v);
        fclose(fp);
        const int n = (int)g.size();
        double b0, binv;
        std::vector<Axis4> bins, cells;
        make_tables(g.data(), n, bins, cells, b0, binv);
        std::vector<double> probes;
        for (int i = 0; i < n; i++) {
            probes.push_back(g[i]);
            probes.push_back(std::nextafter(g[i], INFINITY));
            probes.push_back(std::nextafter(g[i], -INFINITY));
            if (i + 1 < n) probes.push_back((g[i] + g[i + 1]) / 2.0);
        }
        std::mt19937_64 rng(12345 + f);
        const double span = g[n - 1] - g[0];
        std::uniform_real_distribution<double> U(g[0] - 0.1 * span, g[n - 1] + 0.1 * span);
        for (int k = 0; k < 200000; k++) probes.push_back(U(rng));
        for (size_t j = 0; j < bins.size(); j++) {
            const double xj = b0 + (double)j / binv;
            probes.push_back(xj);
            probes.push_back(std::nextafter(xj, INFINITY));
            probes.push_back(std::nextafter(xj, -INFINITY));
        }
        const double specials[] = {INFINITY, -INFINITY, NAN, -0.0, 0.0, 1e300, -1e300};
        for (double s : specials) probes.push_back(s);
        int bad = 0;
        for (double x : probes) {
            const Axis4 r = axis_find(g.data(), n, bins.data(), cells.data(), (int)bins.size(), b0, binv, x);
            const int a = r.s;
            const int b = host_bisect(g.data(), n, x);
            bool ok = a == b;
            // the record carries gridlines b-1 .. b+2 (NaN past either end)
            for (int k = 0; k < 4 && ok; k++) {
                const int i = b - 1 + k;
                ok = (i >= 0 && i < n) ? r.g[k] == g[i] : std::isnan(r.g[k]);
            }
            if (!ok) {
                if (bad < 5) printf("mismatch x=%.17g table=%d bisect=%d\n", x, a, b);
                bad++;
            }
        }
        printf("%zu %d\n", probes.size(), bad);
        bad_total += bad;
    }
    return bad_total ? 1 : 0;
}
