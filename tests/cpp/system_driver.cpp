// system_driver.cpp -- runs a system described by a reference system file
// through the C++ API: evolver + createFromFile (Fields / Parameters /
// Equations sections) + initializeFromFile for each field with an initial
// CSV + advanceTime with CSV output (writeOut cadence of the reference).
// Usage: system_driver SX SY SZ DX DT STEPS WRITE_EVERY SYSTEM_FILE [FIELD=IC.csv ...]
//        Noise seeds are fixed with evolver::setSeed(SEED) when CUPSS_SEED is set.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/cupss.h"

int main(int argc, char **argv)
{
    if (argc < 9) {
        std::cerr << "usage: system_driver SX SY SZ DX DT STEPS WRITE_EVERY SYSTEM_FILE [FIELD=IC.csv ...]"
                  << std::endl;
        return 2;
    }
    const int sx = std::atoi(argv[1]), sy = std::atoi(argv[2]), sz = std::atoi(argv[3]);
    const float dx = std::atof(argv[4]), dt = std::atof(argv[5]);
    const int steps = std::atoi(argv[6]), every = std::atoi(argv[7]);
    evolver *e;
    if (sz > 1)
        e = new evolver(RUN_GPU, sx, sy, sz, dx, dx, dx, dt, every);
    else if (sy > 1)
        e = new evolver(RUN_GPU, sx, sy, dx, dx, dt, every);
    else
        e = new evolver(RUN_GPU, sx, dx, dt, every);
    if (const char *seed = std::getenv("CUPSS_SEED")) e->setSeed(std::strtoull(seed, nullptr, 10));
    if (e->createFromFile(argv[8]) != 0) return 3;
    for (int a = 9; a < argc; a++) {
        std::string arg = argv[a];
        const size_t eq = arg.find('=');
        e->initializeFromFile(arg.substr(0, eq), arg.substr(eq + 1), 1, ',');
    }
    e->prepareProblem();
    for (int s = 0; s < steps; s++) e->advanceTime();
    e->writeOut();
    delete e;
    return 0;
}
