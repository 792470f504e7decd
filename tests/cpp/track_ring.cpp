// Tracks a small FODO ring through the C++ API (include/simtrack_amd.hpp) on
// the GPU and writes the results as raw doubles/ints to stdout:
//   per-particle Track() calls, one batched Track(), batched Track_spin(),
//   and Track_DA1's file (path given as argv[1]).
#include <cmath>
#include <cstdio>
#include <vector>

#include "simtrack_amd.hpp"

using namespace simtrack_amd;

static void build_ring(Line &l)
{
    const int ncell = 8;
    const double ang = 2 * 3.14159265358979323 / (2 * ncell);
    for (int c = 0; c < ncell; ++c) {
        double kn[11] = {0.}, ks[11] = {0.};
        kn[1] = 0.0372;
        kn[2] = 0.01 * (c % 3 - 1);
        ks[3] = 0.5 * (c % 2);
        MULT *qf = new MULT("QF", 1.5, kn, ks);
        qf->SetP("Nint", 4);
        kn[1] = -0.0372;
        MULT *qd = new MULT("QD", 1.5, kn, ks);
        qd->SetP("Nint", 4);
        SEXT *sf = new SEXT("SF", 0.5, 0.12, 0.);
        sf->SetP("Nint", 4);
        l.Append(qf);
        l.Append(new DRIFT("D", 0.75));
        l.Append(sf);
        l.Append(new DRIFT("D", 0.75));
        l.Append(new SBEND("B", 32., ang, 0., 0.));
        l.Append(new DRIFT("D", 0.75));
        l.Append(new KICKER("K", 0.5, 1e-6, -1e-6));
        l.Append(new BPM("BPM", 0.));
        l.Append(qd);
        l.Append(new DRIFT("D", 2.0));
        l.Append(new SBEND("B", 32., ang, 0., 0.));
        l.Append(new DRIFT("D", 1.25));
    }
    l.Cell[3]->DX = 2e-4;
    l.Cell[8]->DPSI = 1e-3;
    for (Element *e : l.Cell) e->APx = e->APy = 0.05;
}

int main(int argc, char **argv)
{
    Set_RefPartEnergy(100000.0 / 938.27201, 938.27201, 1, 1);
    Line l;
    build_ring(l);
    const int np = 24, nturn = 200;
    std::vector<double> x0(np * 6, 0.), xs(np * 9, 0.);
    for (int p = 0; p < np; ++p) {
        double a = 2e-3 * (p + 1), t = 0.2 + 0.05 * p;
        x0[p * 6 + 0] = a * std::cos(t);
        x0[p * 6 + 2] = 0.6 * a * std::sin(t);
        x0[p * 6 + 5] = 2e-4;
        for (int k = 0; k < 6; ++k) xs[p * 9 + k] = x0[p * 6 + k];
        xs[p * 9 + 7] = 1.0;
    }
    /* reference API, one particle per call */
    std::vector<double> x1 = x0;
    std::vector<int> s1(np), t1(np), p1(np);
    for (int p = 0; p < np; ++p) {
        int st = 1, lt = 0, lp = 0;
        Track(l, &x1[p * 6], nturn, st, lt, lp);
        s1[p] = st;
        t1[p] = lt;
        p1[p] = lp;
    }
    /* batched */
    std::vector<double> x2 = x0;
    std::vector<int> s2(np, 1), t2(np, 0), p2(np, 0);
    Track(l, x2.data(), np, nturn, s2.data(), t2.data(), p2.data());
    std::vector<int> s3(np, 1), t3(np, 0), p3(np, 0);
    Track_spin(l, xs.data(), np, nturn, s3.data(), t3.data(), p3.data());
    if (argc > 1) Track_DA1(l, 50, 2e-4, 6e-4, 4e-4, argv[1]);
    std::fwrite(x1.data(), sizeof(double), x1.size(), stdout);
    std::fwrite(s1.data(), sizeof(int), np, stdout);
    std::fwrite(t1.data(), sizeof(int), np, stdout);
    std::fwrite(p1.data(), sizeof(int), np, stdout);
    std::fwrite(x2.data(), sizeof(double), x2.size(), stdout);
    std::fwrite(s2.data(), sizeof(int), np, stdout);
    std::fwrite(xs.data(), sizeof(double), xs.size(), stdout);
    std::fwrite(s3.data(), sizeof(int), np, stdout);
    return 0;
}
