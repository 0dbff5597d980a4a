// cu_utils.cpp -- the two utility calls the reference's API exposes
// (inc/cupss/cu_utils.h: check_error, check_device), for the HIP runtime.
//
// check_error(e): a failed runtime call is fatal for the public API (the
//   reference's convention: message on stderr, exit status 1).  The message
//   names the error symbolically and numerically.
// check_device(): prints what a solver run will see -- HIP runtime and
//   driver versions and, per visible device, its name, architecture, compute
//   units, device memory and LDS per workgroup -- and exits with status 1
//   when no device is visible.  HIP encodes versions as
//   major * 10^7 + minor * 10^5 + patch.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>

#include "../../../include/cupss.h"

namespace {

std::string hip_version_string(int v)
{
    std::ostringstream os;
    os << v / 10000000 << "." << (v / 100000) % 100 << "." << v % 100000;
    return os.str();
}

[[noreturn]] void fatal(const std::string &what)
{
    std::cerr << what << std::endl;
    std::exit(1);
}

}  // namespace

void check_error(hipError_t err)
{
    if (err == hipSuccess) return;
    std::ostringstream os;
    os << "HIP runtime error " << static_cast<int>(err) << " (" << hipGetErrorName(err)
       << "): " << hipGetErrorString(err);
    fatal(os.str());
}

void check_device()
{
    int count = 0, runtime = 0, driver = 0;
    check_error(hipRuntimeGetVersion(&runtime));
    check_error(hipDriverGetVersion(&driver));
    const hipError_t e = hipGetDeviceCount(&count);
    if (e == hipErrorNoDevice || (e == hipSuccess && count == 0))
        fatal("no HIP device is visible to this process; a RUN_GPU system cannot run");
    check_error(e);

    std::ostringstream os;
    os << "HIP runtime " << hip_version_string(runtime) << ", driver " << hip_version_string(driver) << "\n";
    for (int d = 0; d < count; d++) {
        hipDeviceProp_t p;
        check_error(hipGetDeviceProperties(&p, d));
        char line[512];
        std::snprintf(line, sizeof line, "  device %d: %s [%s], %d compute units, %.1f GiB, %zu KiB LDS per workgroup\n",
                      d, p.name, p.gcnArchName, p.multiProcessorCount,
                      (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0), (size_t)p.sharedMemPerBlock / 1024);
        os << line;
    }
    std::cout << os.str() << std::flush;
}
