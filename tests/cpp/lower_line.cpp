// Builds a line of every element type with the C++ API (include/simtrack_amd.hpp),
// applies SetP edits and misalignments, and writes the lowered st_element
// records to stdout as raw bytes.  No GPU call is made.
#include <cmath>
#include <cstdio>

#include "simtrack_amd.hpp"

using namespace simtrack_amd;

/* Stand-in for a foreign Element class that exposes only what the
   reference's does: public fields and GetP (simtrack.h:379-395). */
struct Foreign {
    Element *e;
    std::string TYPE, NAME;
    double L, DX, DY, DS, DTHETA, DPHI, DPSI, APx, APy;
    double GetP(const char *n) { return e->GetP(n); }
};
struct ForeignLine {
    std::vector<Foreign *> Cell;
};

int main()
{
    double knl[11] = {0., 0.031, 0.2, -1.5, 0., 12., 0., 0., 0., 0., 0.};
    double knsl[11] = {0., 0.001, -0.05, 0., 0., 0., 0., 3.0, 0., 0., 0.};
    double z[6] = {0., 0., 0., 0., 0., 0.}, xo[6] = {1e-4, 0., -2e-4, 0., 0., 0.};
    double m66[36];
    for (int i = 0; i < 36; ++i) m66[i] = (i % 7 == 0) ? 1.0 : 1e-3 * i;
    double n1[3] = {0.3, 0.9, 0.1}, n2[3] = {1.0, 0.0, 0.2};
    Line l;
    l.Append(new DRIFT("D1", 1.3));
    l.Append(new MARKER("M0", 0.));
    l.Append(new BPM("B0", 0.));
    l.Append(new HBPM("HB", 0.2));
    l.Append(new VBPM("VB", 0.));
    l.Append(new SBEND("BE", 3.0, 0.02, 0.003, -0.002));
    l.Append(new GSBEND("GB", 2.2, 0.016, 0.0161, 0.001, 0.0));
    l.Append(new QUAD("QF", 1.1, 0.05, 0.002));
    l.Append(new SKEWQ("SQ", 0.4, 0.01));
    l.Append(new SEXT("SF", 0.5, 0.3, 0.02));
    l.Append(new OCT("OF", 0.3, 2.0, 0.5));
    l.Append(new MULT("MU", 0.9, knl, knsl));
    l.Append(new GMULT("GM", 1.5, -0.012, 0., 0., knl, knsl));
    l.Append(new SBENDMULT("SBM", 2.5, 0.018, 0.001, 0.001, knl, knsl));
    l.Append(new SMULT("SM", 1.2, 0.01, 0., 0., knl, knsl));
    l.Append(new GSBENDMULT("GBM", 1.9, 0.014, 0.0139, 0., 0., knl, knsl));
    l.Append(new KICKER("K", 0.4, 2e-5, -1e-5));
    l.Append(new HKICKER("HK", 0., 3e-5));
    l.Append(new VKICKER("VK", 0.2, -2e-5));
    l.Append(new RFCAV("RF", 0.6, 3.0, 28.15e6, 0.05));
    l.Append(new CRABRF("CR", 0.4, 2.0, 197e6, 0.1));
    l.Append(new CCMULT("CCM", 0., 25.8, 197e6, 0.0, 2.8e-3, -9.8e-7, 2.2e-3, -3.9e-5, 0.59, 2.3e-3, 0.1, 0.01));
    l.Append(new LBT("LBT", 0.0125));
    l.Append(new ILBT("ILBT", 0.0125));
    l.Append(new BEAMBEAM("BB6", 6, 1.72e11, -1.0, 0.02, 5, 20e-9, 0.45, 0., 1.3e-9, 0.056, 0.));
    l.Append(new CRBB("CRB", 1.5e11, 0.01, 1.0, 0.05, 5, 5e-9, 0.4, 0.02, 2.5e-9, 0.2, 0.));
    l.Append(new MATRIX("MAT", 0., z, xo, m66));
    l.Append(new ROTAT("ROT", 0.5, n1, 0.2));
    l.Append(new SNAKE("SNK", 0.1, n2, 1.5707963267948966));
    l.Append(new TRANS("TR", 2e-4, -1e-4, 0.05));
    l.Append(new SROTAT("SR", 0.013));
    l.Append(new YROTAT("YR", -0.007));
    l.Append(new XROTAT("XR", 0.009));
    l.Append(new COOLING("CO", 1e-4));
    l.Append(new LRBB("LR", 1.5e11, 1.0, 8e-4, -3e-4, 2e-4, 1e-4));
    l.Append(new ELENS("EL", 2.0, 1e11, 1.0, 3, 0.5, 3e-4, 2e-4));
    l.Append(new HELENS("HE", 2.0, 5e11, 1.0, 2, 0.5, 1e-2, 1e-6));
    l.Append(new HACDIP("HAC", 0.3, 2e-5, 0.31, 0.2, -3, 5));
    l.Append(new VACDIP("VAC", 0.0, 1e-5, 0.27, 0.0, -10, -2));
    l.Append(new HACMULT("HAM", 0.2, 2, 0.05, 100, 0.3));
    l.Append(new VACMULT("VAM", 0.0, 0, 1e-5, 50, 0.1));
    l.Append(new WIGGLER("WIG", 0.8, 8, 0.5, 2.0, 31.4, 0.1));
    /* edits through the reference's parameter names */
    l.Cell[7]->SetP("K1L", 0.06);
    l.Cell[7]->SetP("Nint", 7.9);
    l.Cell[11]->SetP("K4L", 0.5);        /* Norder stays 7 (set at construction) */
    l.Cell[24]->SetP("NSLICE", 3);
    l.Cell[5]->DX = 1e-4;
    l.Cell[5]->DPSI = 2e-3;
    l.Cell[9]->APx = 0.03;
    l.Cell[34]->SetP("SEPX", 9e-4);
    l.Cell[37]->SetP("TURNE", 6);
    l.Cell[39]->SetP("PHI0", 0.25);
    l.Cell[36]->DX = 3e-5;
    std::vector<st_element> r = records_from_line(l);
    if (l.Cell[11]->GetP("Norder") != 7 || l.Cell[7]->GetP("Nint") != 7) return 3;
    bool threw = false;
    try {
        l.Cell[0]->SetP("K1L", 1.0);
    } catch (const Error &) {
        threw = true;
    }
    if (!threw) return 4;
    /* the duck-typed path a reference-side driver takes (MATRIX excluded:
       the reference's MATRIX has no readable parameters) */
    std::vector<Foreign> fe;
    for (Element *e : l.Cell)
        if (e->TYPE != "MATRIX")
            fe.push_back({e, e->TYPE, e->NAME, e->L, e->DX, e->DY, e->DS, e->DTHETA, e->DPHI, e->DPSI, e->APx, e->APy});
    ForeignLine fl;
    for (Foreign &f : fe) fl.Cell.push_back(&f);
    std::vector<st_element> r2 = records_from_line(fl);
    size_t k = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        if (r[i].type == ST_MATRIX) continue;
        if (std::memcmp(&r[i], &r2[k++], sizeof(st_element)) != 0) {
            std::fprintf(stderr, "foreign lowering differs at element %zu\n", i);
            const unsigned char *a = (const unsigned char *)&r[i], *b = (const unsigned char *)&r2[k - 1];
            for (size_t o = 0; o < sizeof(st_element); ++o)
                if (a[o] != b[o]) std::fprintf(stderr, "  byte %zu\n", o);
            return 5;
        }
    }
    std::fwrite(r.data(), sizeof(st_element), r.size(), stdout);
    std::fprintf(stderr, "Length %.17g Ncell %ld\n", l.Length, l.Ncell);
    return 0;
}
