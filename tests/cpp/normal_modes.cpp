// Cal_A of the C++ API (simtrack_amd::detail::cal_a_4d) on one-turn
// matrices read from stdin (36 doubles each, row-major) until EOF; writes
// each 4x4 A (row-major, 16 doubles) to stdout.  Host arithmetic only.
#include <cstdio>

#include "simtrack_amd.hpp"

int main()
{
    double m[36];
    while (std::fread(m, sizeof(double), 36, stdin) == 36) {
        double a[4][4];
        try {
            simtrack_amd::detail::cal_a_4d(m, a);
        } catch (const simtrack_amd::Error &e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
        std::fwrite(&a[0][0], sizeof(double), 16, stdout);
    }
    return 0;
}
