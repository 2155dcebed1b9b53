// Launch wrappers of the step / auxiliary / IBM-tile kernels for one storage
// type S; each kernels_fpXX.hip includes this with its S. The coefficient
// representation (kRepMom) exists for float and double storage only.
#pragma once
#include <type_traits>

#include "kernels_impl.h"
#include "kernels_mr.h"

namespace vlbm {

template <typename S>
struct Inst {
    using T = typename ComputeOf<S>::type;
    static constexpr bool kHasMom = sizeof(S) > 2;
    static dim3 grid_of(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

    template <int Q, int RHEO, bool MASK, bool IBM>
    static void step_mom(bool direct, int in_rep, const Params<T>& p, hipStream_t st) {
        if constexpr (kHasMom) {
            if (!direct) {
                using SH = TileShape<T, Q, MASK>;
                const int zc = tile_zchunk(p.nx, p.ny, p.nz, SH::TY);
                const uint32_t tiles = (uint32_t)((p.nx + kTileX - 1) / kTileX) * (uint32_t)((p.ny + SH::TY - 1) / SH::TY) *
                                       (uint32_t)((p.nz + zc - 1) / zc);
                const uint32_t blocks = 8u * ((tiles + 7u) / 8u);     // XCD order of the tiles (see the kernel)
                hipLaunchKernelGGL((k_step_tile<Q, T, RHEO, MASK, IBM>), dim3(blocks), dim3(SH::NT), 0, st, p, zc);
            } else if (in_rep == kRepMom) {
                hipLaunchKernelGGL((k_step<Q, S, RHEO, MASK, true, IBM, kRepMom, kRepMom>), grid_of(p.n), dim3(kBlock),
                                   0, st, p);
            } else {
                hipLaunchKernelGGL((k_step<Q, S, RHEO, MASK, true, IBM, kRepPop, kRepMom>), grid_of(p.n), dim3(kBlock),
                                   0, st, p);
            }
        }
    }

    template <int Q, int RHEO, bool MASK, bool IBM>
    static void step_pop(bool direct, const Params<T>& p, hipStream_t st) {
        if (direct) {
            hipLaunchKernelGGL((k_step<Q, S, RHEO, MASK, true, IBM>), grid_of(p.n), dim3(kBlock), 0, st, p);
            return;
        }
        if constexpr (Q == 19) {
            // x shift through LDS when rows never cross a block edge (NX divides the block)
            if (kBlock % p.nx == 0) {
                hipLaunchKernelGGL((k_step<Q, S, RHEO, MASK, false, IBM, kRepPop, kRepPop, true>), grid_of(p.n),
                                   dim3(kBlock), 0, st, p);
                return;
            }
        }
        hipLaunchKernelGGL((k_step<Q, S, RHEO, MASK, false, IBM>), grid_of(p.n), dim3(kBlock), 0, st, p);
    }

    template <int Q, int RHEO>
    static int step_q(bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* pv, hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        if (out_rep == kRepMom) {
            if (!kHasMom) return -1;
            // the IBM force field applies to the first step after a checkpoint
            // load too (post-stream state read in place)
            if (ibm) {
                if (mask) step_mom<Q, RHEO, true, true>(direct, in_rep, p, st);
                else step_mom<Q, RHEO, false, true>(direct, in_rep, p, st);
            } else {
                if (mask) step_mom<Q, RHEO, true, false>(direct, in_rep, p, st);
                else step_mom<Q, RHEO, false, false>(direct, in_rep, p, st);
            }
        } else {
            if (in_rep != kRepPop) return -1;
            if (ibm) {
                if (mask) step_pop<Q, RHEO, true, true>(direct, p, st);
                else step_pop<Q, RHEO, false, true>(direct, p, st);
            } else {
                if (mask) step_pop<Q, RHEO, true, false>(direct, p, st);
                else step_pop<Q, RHEO, false, false>(direct, p, st);
            }
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    template <int Q, int RHEO>
    static int aux_q(bool mask, int in_rep, const void* pv, hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        const dim3 grid = grid_of(p.aux_count ? p.aux_count : p.n);
        if (in_rep == kRepMom) {
            if constexpr (kHasMom) {
                if (mask) hipLaunchKernelGGL((k_aux<Q, S, RHEO, true, kRepMom>), grid, dim3(kBlock), 0, st, p);
                else hipLaunchKernelGGL((k_aux<Q, S, RHEO, false, kRepMom>), grid, dim3(kBlock), 0, st, p);
            } else {
                return -1;
            }
        } else {
            if (mask) hipLaunchKernelGGL((k_aux<Q, S, RHEO, true>), grid, dim3(kBlock), 0, st, p);
            else hipLaunchKernelGGL((k_aux<Q, S, RHEO, false>), grid, dim3(kBlock), 0, st, p);
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    template <int Q>
    static int ibm_macr_q(bool mask, int in_rep, const void* pv, const uint32_t* list, const uint32_t* count,
                          uint32_t max_tiles, hipStream_t st) {
        const Params<T>& p = *(const Params<T>*)pv;
        // grid-stride over the tile nodes; enough blocks to cover the tiles the
        // spheres can touch, at most 4096
        const uint32_t blocks = (max_tiles * (uint32_t)kTileNodes + kBlock - 1) / kBlock;
        const dim3 grid(blocks < 4096u ? (blocks ? blocks : 1u) : 4096u);
        if (in_rep == kRepMom) {
            if constexpr (kHasMom) {
                if (mask) hipLaunchKernelGGL((k_ibm_macr<Q, S, true, kRepMom>), grid, dim3(kBlock), 0, st, p, list, count);
                else hipLaunchKernelGGL((k_ibm_macr<Q, S, false, kRepMom>), grid, dim3(kBlock), 0, st, p, list, count);
            } else {
                return -1;
            }
        } else {
            if (mask) hipLaunchKernelGGL((k_ibm_macr<Q, S, true>), grid, dim3(kBlock), 0, st, p, list, count);
            else hipLaunchKernelGGL((k_ibm_macr<Q, S, false>), grid, dim3(kBlock), 0, st, p, list, count);
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }

    static int step(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                    void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) {
            if (rheo == 0) return step_q<19, 0>(mask, direct, ibm, in_rep, out_rep, p, st);
            if (rheo == 1) return step_q<19, 1>(mask, direct, ibm, in_rep, out_rep, p, st);
            return step_q<19, 2>(mask, direct, ibm, in_rep, out_rep, p, st);
        }
        if (rheo == 0) return step_q<27, 0>(mask, direct, ibm, in_rep, out_rep, p, st);
        if (rheo == 1) return step_q<27, 1>(mask, direct, ibm, in_rep, out_rep, p, st);
        return step_q<27, 2>(mask, direct, ibm, in_rep, out_rep, p, st);
    }

    static int aux(int q, int rheo, bool mask, int in_rep, const void* p, void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) {
            if (rheo == 0) return aux_q<19, 0>(mask, in_rep, p, st);
            if (rheo == 1) return aux_q<19, 1>(mask, in_rep, p, st);
            return aux_q<19, 2>(mask, in_rep, p, st);
        }
        if (rheo == 0) return aux_q<27, 0>(mask, in_rep, p, st);
        if (rheo == 1) return aux_q<27, 1>(mask, in_rep, p, st);
        return aux_q<27, 2>(mask, in_rep, p, st);
    }

    static int ibm_macr(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                        uint32_t max_tiles, void* stream) {
        hipStream_t st = (hipStream_t)stream;
        if (q == 19) return ibm_macr_q<19>(mask, in_rep, p, list, count, max_tiles, st);
        return ibm_macr_q<27>(mask, in_rep, p, list, count, max_tiles, st);
    }
};

}  // namespace vlbm
