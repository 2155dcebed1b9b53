// Launch wrappers of the step / auxiliary / IBM-tile kernels for one storage
// type S; each kernels_fpXX.hip includes this with its S.
#pragma once
#include "kernels_impl.h"

namespace vlbm {

template <typename S>
struct Inst {
    using T = typename ComputeOf<S>::type;
    static dim3 grid_of(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

    template <int Q, int RHEO>
    static int step_q(bool mask, bool direct, bool ibm, const void* pv, hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        const dim3 grid = grid_of(p.n), blk(kBlock);
        if (ibm) {
            // the IBM force field applies to the first step after a checkpoint
            // load too (post-stream state read in place)
            if (mask) {
                if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, true, true, true>), grid, blk, 0, st, p);
                else hipLaunchKernelGGL((k_step<Q, S, RHEO, true, false, true>), grid, blk, 0, st, p);
            } else {
                if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, false, true, true>), grid, blk, 0, st, p);
                else hipLaunchKernelGGL((k_step<Q, S, RHEO, false, false, true>), grid, blk, 0, st, p);
            }
        } else if (mask) {
            if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, true, true>), grid, blk, 0, st, p);
            else hipLaunchKernelGGL((k_step<Q, S, RHEO, true, false>), grid, blk, 0, st, p);
        } else {
            if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, false, true>), grid, blk, 0, st, p);
            else hipLaunchKernelGGL((k_step<Q, S, RHEO, false, false>), grid, blk, 0, st, p);
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    template <int Q, int RHEO>
    static int aux_q(bool mask, const void* pv, hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        const dim3 grid = grid_of(p.n);
        if (mask) hipLaunchKernelGGL((k_aux<Q, S, RHEO, true>), grid, dim3(kBlock), 0, st, p);
        else hipLaunchKernelGGL((k_aux<Q, S, RHEO, false>), grid, dim3(kBlock), 0, st, p);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    template <int Q>
    static int ibm_macr_q(bool mask, const void* pv, const uint32_t* list, const uint32_t* count, uint32_t max_tiles,
                          hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        // grid-stride over the tile nodes; enough blocks to cover the tiles the
        // spheres can touch, at most 4096
        const uint32_t blocks = (max_tiles * (uint32_t)kTileNodes + kBlock - 1) / kBlock;
        const dim3 grid(blocks < 4096u ? (blocks ? blocks : 1u) : 4096u);
        if (mask) hipLaunchKernelGGL((k_ibm_macr<Q, S, true>), grid, dim3(kBlock), 0, st, p, list, count);
        else hipLaunchKernelGGL((k_ibm_macr<Q, S, false>), grid, dim3(kBlock), 0, st, p, list, count);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    static int step(int q, int rheo, bool mask, bool direct, bool ibm, const void* p, void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) {
            if (rheo == 0) return step_q<19, 0>(mask, direct, ibm, p, st);
            if (rheo == 1) return step_q<19, 1>(mask, direct, ibm, p, st);
            return step_q<19, 2>(mask, direct, ibm, p, st);
        }
        if (rheo == 0) return step_q<27, 0>(mask, direct, ibm, p, st);
        if (rheo == 1) return step_q<27, 1>(mask, direct, ibm, p, st);
        return step_q<27, 2>(mask, direct, ibm, p, st);
    }

    static int aux(int q, int rheo, bool mask, const void* p, void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) {
            if (rheo == 0) return aux_q<19, 0>(mask, p, st);
            if (rheo == 1) return aux_q<19, 1>(mask, p, st);
            return aux_q<19, 2>(mask, p, st);
        }
        if (rheo == 0) return aux_q<27, 0>(mask, p, st);
        if (rheo == 1) return aux_q<27, 1>(mask, p, st);
        return aux_q<27, 2>(mask, p, st);
    }

    static int ibm_macr(int q, bool mask, const void* p, const uint32_t* list, const uint32_t* count,
                        uint32_t max_tiles, void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) return ibm_macr_q<19>(mask, p, list, count, max_tiles, st);
        return ibm_macr_q<27>(mask, p, list, count, max_tiles, st);
    }
};

}  // namespace vlbm
