// RHIC diagnostics through the C++ API, as the reference's first example
// driver sets the ring up: Au ions at gamma = 100, the MAD-X table (argv[1],
// uncompressed) read by Read_MADX_Lattice, the first RFCAV at 4 MV and
// h = 360.  Runs Cal_Orbit_Num, Cal_OneTurnMap and Track_tbt_ActionDiff
// (4 sigma, phase 0.7, emittances 1e-8 m, 64 turns x 2 steps, 16
// particles; ./Diff-output.dat in the working directory) and writes to
// stdout, as raw doubles: the last element's X[6], M[36], Beta1, Alfa1,
// Beta2, Alfa2, then the seed after the scan.  Then Cal_Twiss,
// Cal_Dispersion and Cal_Chrom, ./twiss0.dat by Print_Twiss, and to stdout
// Tune1, Tune2, Chromx1, Chromy1, Chromx2, Chromy2 and per element (n of
// each) Beta1, Alfa1, Beta2, Alfa2, Mu1, Mu2, r, Etax.
#include <cstdio>

#include "simtrack_amd.hpp"

using namespace simtrack_amd;

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: rhic_diagnostics MADX_TWISS_FILE\n");
        return 2;
    }
    try {
        Set_RefPartEnergy(100., 938.27201 * 197, 79, 197);
        Line l;
        Read_MADX_Lattice(l, argv[1]);
        for (Element *e : l.Cell)
            if (e->TYPE == "RFCAV") {
                e->SetP("VRF", 4.0);
                e->SetP("FRF", 78250.42279 * 360);
                break;
            }
        Track_tbt_ActionDiff(l, 4.0, 0.7, 1e-8, 1e-8, 64, 2, 16);
        const Element *last = l.Cell.back();
        std::fwrite(last->X, sizeof(double), 6, stdout);
        std::fwrite(last->M, sizeof(double), 36, stdout);
        const double tw[5] = {last->Beta1, last->Alfa1, last->Beta2, last->Alfa2, seed};
        std::fwrite(tw, sizeof(double), 5, stdout);
        Cal_Chrom(l);
        Cal_Dispersion(l);
        Print_Twiss(l, "./twiss0.dat");
        const double q[6] = {l.Tune1, l.Tune2, l.Chromx1, l.Chromy1, l.Chromx2, l.Chromy2};
        std::fwrite(q, sizeof(double), 6, stdout);
        for (int f = 0; f < 8; ++f)
            for (const Element *e : l.Cell) {
                const double v[8] = {e->Beta1, e->Alfa1, e->Beta2, e->Alfa2, e->Mu1, e->Mu2, e->r, e->Etax};
                std::fwrite(&v[f], sizeof(double), 1, stdout);
            }
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
