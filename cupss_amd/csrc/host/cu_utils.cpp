// cu_utils.cpp -- device error check and query on the HIP runtime
// (reference: src/cu_utils.cpp:4-37).
#include <cstdlib>
#include <iostream>

#include "../../../include/cupss.h"

void check_error(hipError_t err)
{
    if (err != hipSuccess) {
        std::cerr << "HIP Runtime Error" << std::endl;
        std::cerr << hipGetErrorString(err) << std::endl;
        std::exit(1);
    }
}

void check_device()
{
    int nDevices = 0, runtimeVersion = 0, driverVersion = 0;
    check_error(hipGetDeviceCount(&nDevices));
    check_error(hipRuntimeGetVersion(&runtimeVersion));
    check_error(hipDriverGetVersion(&driverVersion));
    if (nDevices == 0) {
        std::cerr << "No HIP devices found, exiting." << std::endl;
        std::exit(1);
    }
    std::cout << "Devices found:" << std::endl;
    for (int i = 0; i < nDevices; i++) {
        hipDeviceProp_t prop;
        check_error(hipGetDeviceProperties(&prop, i));
        std::cout << "Device " << i << ": " << prop.name << " (" << prop.gcnArchName << ")" << std::endl;
    }
    std::cout << "HIP Driver Version: " << driverVersion << std::endl;
    std::cout << "HIP Runtime Version: " << runtimeVersion << std::endl;
}
