// Reads a lattice file with the C++ API's Read_MADX_Lattice or
// Read_BMAD_Lattice and writes the lowered st_element records to stdout as
// raw bytes, then GP.circumference as one double.  No GPU call is made.
#include <cstdio>
#include <cstring>

#include "simtrack_amd.hpp"

using namespace simtrack_amd;

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: read_lattice madx|bmad FILE\n");
        return 2;
    }
    Line l;
    try {
        if (!std::strcmp(argv[1], "madx"))
            Read_MADX_Lattice(l, argv[2]);
        else
            Read_BMAD_Lattice(l, argv[2]);
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::vector<st_element> rec = records_from_line(l);
    std::fwrite(rec.data(), sizeof(st_element), rec.size(), stdout);
    std::fwrite(&GP.circumference, sizeof(double), 1, stdout);
    return 0;
}
