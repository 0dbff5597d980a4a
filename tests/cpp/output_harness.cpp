// output_harness.cpp -- drives the background field-output writer
// (cupss_amd/csrc/host/output.cpp) without a device: writes a few 1D/2D/3D
// fields into the directory given as argv[1], then waits for them.
#include <cmath>
#include <cstdio>
#include <string>

#include "../../cupss_amd/csrc/host/output.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    int count = 0;
    for (int dim = 1; dim <= 3; dim++)
        for (int step = 0; step < 12; step++) {
            cupss_host::OutputJob j;
            j.dim = dim;
            j.sx = 8;
            j.sy = dim >= 2 ? 4 : 1;
            j.sz = dim == 3 ? 2 : 1;
            j.precision = step % 2 ? 6 : 3;
            j.name = "f" + std::to_string(dim);
            j.file = dir + "/" + j.name + ".csv." + std::to_string(step);
            j.values.resize((size_t)j.sx * j.sy * j.sz);
            for (size_t i = 0; i < j.values.size(); i++) j.values[i] = (float)std::sin(0.37 * i + step) * 3.25f;
            cupss_host::output_submit(std::move(j));
            count++;
        }
    cupss_host::output_flush();
    std::printf("%d\n", count);
    return 0;
}
