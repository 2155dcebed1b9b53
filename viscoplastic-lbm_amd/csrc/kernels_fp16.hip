// Step, auxiliary and IBM-tile kernels instantiated for _Float16 storage.
#include "kernels_inst.h"

namespace vlbm {

int launch_step_fp16(int q, int rheo, bool mask, bool direct, bool ibm, const void* p, void* stream) {
    return Inst<_Float16>::step(q, rheo, mask, direct, ibm, p, stream);
}

int launch_aux_fp16(int q, int rheo, bool mask, const void* p, void* stream) {
    return Inst<_Float16>::aux(q, rheo, mask, p, stream);
}

int launch_ibm_macr_fp16(int q, bool mask, const void* p, const uint32_t* list, const uint32_t* count,
                       uint32_t max_tiles, void* stream) {
    return Inst<_Float16>::ibm_macr(q, mask, p, list, count, max_tiles, stream);
}

}  // namespace vlbm
