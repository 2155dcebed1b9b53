// Launch sequence of the immersed boundary stage (particles FP64; fluid-side
// fields in the lattice's compute type). The tile macroscopics kernel, which
// reads the populations, is launched from the per-precision kernel files.
#include "ibm_kernels.h"

namespace vlbm {

static inline dim3 blocks_for(uint32_t n) { return dim3(n ? (n + kIbmBlock - 1) / kIbmBlock : 1u); }

static inline dim3 wave_per_node_grid(uint32_t nodes) {
    return dim3((uint32_t)(((uint64_t)nodes * 64 + kInterpBlock - 1) / kInterpBlock));
}

template <typename T>
static int spread(const IbmParams& P, const IbmFluid<T>& fl, hipStream_t st) {
    hipLaunchKernelGGL((k_ibm_spread<T>), wave_per_node_grid((uint32_t)P.n_nodes), dim3(kInterpBlock), 0, st, P, fl);
    return 0;
}

template <>
int spread<float>(const IbmParams& P, const IbmFluid<float>& fl, hipStream_t st) {
    if (P.box_edge <= 0) {
        hipLaunchKernelGGL((k_ibm_spread<float>), wave_per_node_grid((uint32_t)P.n_nodes), dim3(kInterpBlock), 0,
                           st, P, fl);
        return 0;
    }
    const size_t lds = 3 * sizeof(int) * (size_t)P.box_edge * P.box_edge * P.box_edge;
    static size_t lds_set = 0;
    if (lds > lds_set) {
        if (hipFuncSetAttribute((const void*)k_ibm_spread_box, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -1;
        lds_set = lds;
    }
    hipLaunchKernelGGL(k_ibm_spread_box, dim3((uint32_t)P.n_spheres), dim3(kSphereBlock), lds, st, P, fl);
    return 0;
}

template <typename T>
static int stage_part(const IbmParams& P, const IbmFluid<T>& fl, int part, hipStream_t st) {
    const dim3 blk(kIbmBlock);
    const uint32_t ns = (uint32_t)P.n_spheres, nn = (uint32_t)P.n_nodes;
    switch (part) {
    case kIbmBegin:
        hipLaunchKernelGGL(k_ibm_begin, blocks_for(ns > P.n_tiles ? ns : P.n_tiles), blk, 0, st, P);
        if (nn) hipLaunchKernelGGL(k_ibm_mark, blocks_for(nn), blk, 0, st, P);
        hipLaunchKernelGGL(k_ibm_list_tiles, dim3((P.n_tiles + 1023) / 1024), dim3(1024), 0, st, P);
        break;
    case kIbmContacts:
        if (ns) {
            const dim3 waves((ns + kContactWaves - 1) / kContactWaves);
            hipLaunchKernelGGL(k_ibm_cell_sort, dim3(1), dim3(1024), 0, st, P);
            hipLaunchKernelGGL(k_ibm_pairs, waves, dim3(kContactBlock), 0, st, P);
            hipLaunchKernelGGL(k_ibm_contact_sums, waves, dim3(kContactBlock), 0, st, P);
        }
        break;
    case kIbmIteration:
        if (nn) {
            hipLaunchKernelGGL((k_ibm_interp<T>), dim3(xcd_chunk_grid(P.n_chunks)), dim3(kInterpBlock), 0, st, P, fl);
            hipLaunchKernelGGL(k_ibm_add_chunks, blocks_for(ns), blk, 0, st, P);   // with the velocity update
        } else if (ns) {
            hipLaunchKernelGGL(k_ibm_velocity, blocks_for(ns), blk, 0, st, P);
        }
        if (nn && spread<T>(P, fl, st)) return -1;
        break;
    case kIbmOutput: {
        const uint32_t blocks = (P.n_tiles * (uint32_t)kTileNodes + kIbmBlock - 1) / kIbmBlock;
        hipLaunchKernelGGL((k_ibm_output<T>), dim3(blocks < 4096u ? (blocks ? blocks : 1u) : 4096u), blk, 0, st, P, fl);
        break;
    }
    case kIbmMove:
        if (ns) hipLaunchKernelGGL(k_ibm_move, blocks_for(ns), blk, 0, st, P);
        if (nn) hipLaunchKernelGGL(k_ibm_move_nodes, blocks_for(nn), blk, 0, st, P);
        break;
    default:
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_ibm_stage(const IbmParams& P, int tbytes, const void* fluid, int part, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (tbytes == 8) return stage_part<double>(P, *(const IbmFluid<double>*)fluid, part, st);
    return stage_part<float>(P, *(const IbmFluid<float>*)fluid, part, st);
}

}  // namespace vlbm
