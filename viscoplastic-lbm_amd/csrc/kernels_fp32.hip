// Step and auxiliary kernels instantiated for float storage.
#include "kernels_impl.h"

namespace vlbm {

constexpr int kBlk = kBlock;
static inline dim3 grid_of(uint32_t n) { return dim3((n + kBlk - 1) / kBlk); }

template <int Q, int RHEO>
static int step_q(bool mask, bool direct, const void* pv, hipStream_t st) {
    using S = float;
    using T = typename ComputeOf<S>::type;
    const Params<T>& p = *(const Params<T>*)pv;
    const dim3 grid = grid_of(p.n);
    if (mask) {
        if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, true, true>), grid, dim3(kBlk), 0, st, p);
        else hipLaunchKernelGGL((k_step<Q, S, RHEO, true, false>), grid, dim3(kBlk), 0, st, p);
    } else {
        if (direct) hipLaunchKernelGGL((k_step<Q, S, RHEO, false, true>), grid, dim3(kBlk), 0, st, p);
        else hipLaunchKernelGGL((k_step<Q, S, RHEO, false, false>), grid, dim3(kBlk), 0, st, p);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int Q, int RHEO>
static int aux_q(bool mask, const void* pv, hipStream_t st) {
    using S = float;
    using T = typename ComputeOf<S>::type;
    const Params<T>& p = *(const Params<T>*)pv;
    const dim3 grid = grid_of(p.n);
    if (mask) hipLaunchKernelGGL((k_aux<Q, S, RHEO, true>), grid, dim3(kBlk), 0, st, p);
    else hipLaunchKernelGGL((k_aux<Q, S, RHEO, false>), grid, dim3(kBlk), 0, st, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_step_fp32(int q, int rheo, bool mask, bool direct, const void* p, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (q == 19) {
        if (rheo == 0) return step_q<19, 0>(mask, direct, p, st);
        if (rheo == 1) return step_q<19, 1>(mask, direct, p, st);
        return step_q<19, 2>(mask, direct, p, st);
    }
    if (rheo == 0) return step_q<27, 0>(mask, direct, p, st);
    if (rheo == 1) return step_q<27, 1>(mask, direct, p, st);
    return step_q<27, 2>(mask, direct, p, st);
}

int launch_aux_fp32(int q, int rheo, bool mask, const void* p, void* stream) {
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

}  // namespace vlbm
