// Step, auxiliary and IBM-tile kernels instantiated for bfloat16 storage.
#include "kernels_inst.h"

namespace vlbm {

int launch_step_bf16(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream) {
    return Inst<bf16>::step(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
}

int launch_aux_bf16(int q, int rheo, bool mask, int in_rep, const void* p, void* stream) {
    return Inst<bf16>::aux(q, rheo, mask, in_rep, p, stream);
}

int launch_ibm_macr_bf16(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream) {
    return Inst<bf16>::ibm_macr(q, mask, in_rep, p, list, count, max_tiles, stream);
}

}  // namespace vlbm
