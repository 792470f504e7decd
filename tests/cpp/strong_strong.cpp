// Strong-strong beam-beam through the C++ API: two beams of rows built from
// a closed formula (no random numbers), collided by BeamBeam4D_SoftGaussian,
// by BeamBeam6D_SoftGaussian with 5 x 4 equal-population slices (writing
// ./bunch1_Z.dat, ./bunch2_Z.dat in the working directory), with 4 x 4
// Gaussian slice boundaries, by BeamBeam4D_PIC and by BeamBeam6D_PIC with
// 3 x 2 slices.  Writes to stdout, as raw doubles, for each of the five
// collisions the two beams before and after it.
#include <cmath>
#include <cstdio>
#include <vector>

#include "simtrack_amd.hpp"

using namespace simtrack_amd;

static std::vector<double> beam(int n, double phase, double off)
{
    std::vector<double> x((size_t)n * 6);
    for (int i = 0; i < n; ++i) {
        double a = 0.37 * i + phase, r = 1e-4 * (0.3 + std::fmod(0.618 * i, 1.0));
        x[i * 6 + 0] = r * std::cos(a) + off;
        x[i * 6 + 1] = 2e-4 * r * std::sin(a) / 1e-4 * 0.5;
        x[i * 6 + 2] = 0.7 * r * std::sin(1.3 * a);
        x[i * 6 + 3] = 0.6e-4 * std::cos(2.1 * a);
        x[i * 6 + 4] = 0.05 * std::sin(0.731 * i + phase);
        x[i * 6 + 5] = 1e-4 * std::cos(0.413 * i);
    }
    return x;
}

int main()
{
    try {
        const int n1 = 3001, n2 = 2500;
        for (int form = 0; form < 5; ++form) {
            std::vector<double> a = beam(n1, 0.1, 1e-5), b = beam(n2, 0.7, -2e-5);
            std::fwrite(a.data(), sizeof(double), a.size(), stdout);
            std::fwrite(b.data(), sizeof(double), b.size(), stdout);
            if (form == 0) {
                BeamBeam4D_SoftGaussian(1.2e11, n1, a.data(), 0.9e11, n2, b.data(), 230., 240., 0.8, 1.1);
            } else if (form == 1) {
                BeamBeam6D_SoftGaussian(1.2e11, n1, a.data(), 0.9e11, n2, b.data(), 230., 240., 0.8, 1.1, 5, 4);
            } else if (form == 2) {
                double zb[5];
                Zboundary_slices_Gaussian(zb, 4, 0.04);
                BeamBeam6D_SoftGaussian(1.2e11, n1, a.data(), 0.9e11, n2, b.data(), 230., 240., 0.8, 1.1, 4, 4, zb,
                                        zb);
            } else if (form == 3) {
                BeamBeam4D_PIC(1.2e11, n1, a.data(), 0.9e11, n2, b.data(), 230., 240., 0.8, 1.1);
            } else {
                BeamBeam6D_PIC(1.2e11, n1, a.data(), 0.9e11, n2, b.data(), 230., 240., 0.8, 1.1, 3, 2);
            }
            std::fwrite(a.data(), sizeof(double), a.size(), stdout);
            std::fwrite(b.data(), sizeof(double), b.size(), stdout);
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
