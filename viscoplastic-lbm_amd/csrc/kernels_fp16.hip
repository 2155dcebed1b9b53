// Step, auxiliary and IBM-tile kernels instantiated for _Float16 storage.
#include "kernels_inst.h"

namespace vlbm {

int launch_step_fp16(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream) {
    return Inst<_Float16>::step(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
}

int launch_aux_fp16(int q, int rheo, bool mask, int in_rep, const void* p, void* stream) {
    return Inst<_Float16>::aux(q, rheo, mask, in_rep, p, stream);
}

int launch_ibm_macr_fp16(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream) {
    return Inst<_Float16>::ibm_macr(q, mask, in_rep, p, list, count, max_tiles, stream);
}

}  // namespace vlbm
