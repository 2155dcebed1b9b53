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
    void* stage = nullptr;     // StageRec per node: rho, u + (f - f0) / 2 rho of the stage
    std::vector<void*> allocs;
};

}  // namespace vlbm
