// Host-side state of the immersed boundary stage of one lattice handle
// (shared by ibm_capi.cpp and checkpoint.cpp).
#pragma once
#include <vector>

#include "../../include/vlbm.h"
#include "ibm.h"

namespace vlbm {

struct IbmState {
    vlbm_ibm_desc d;
    IbmParams P;
    double r_max = 0;
    void* force[3] = {nullptr, nullptr, nullptr};
    void* macr[4] = {nullptr, nullptr, nullptr, nullptr};   // rho, u of the stage (constant force)
    std::vector<void*> allocs;
};

}  // namespace vlbm
