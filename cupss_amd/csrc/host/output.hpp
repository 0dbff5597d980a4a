// output.hpp -- background writer of field output files.
//
// field::writeToFile (src/field.cpp:350-402) formats every grid point of a
// field as a CSV line; for large grids that takes far longer than the steps
// between outputs.  Here the caller only copies the field's real values to
// host memory and checks them for NaN (the reference's exit condition, kept
// synchronous); formatting and writing happen on one background thread, in
// submission order, with the reference's byte format.  At most
// MAX_PENDING files wait at a time (the caller blocks beyond that), and the
// queue is drained when the last evolver is destroyed and at process exit.
// CUPSS_SYNC_OUTPUT=1 in the environment writes synchronously instead.
#pragma once
#include <string>
#include <vector>

namespace cupss_host {

struct OutputJob {
    std::string file;    // data/<name>.csv.<step>
    std::string name;    // column header
    int dim = 1;
    int sx = 1, sy = 1, sz = 1;
    int precision = 6;
    std::vector<float> values;  // x fastest
};

// queue a file (or write it now when output is synchronous)
void output_submit(OutputJob &&job);
// wait until every queued file is written
void output_flush();

}  // namespace cupss_host
