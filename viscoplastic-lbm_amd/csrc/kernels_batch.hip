// Kernels of the batched (multi-lattice) path: one persistent-grid step
// kernel over the flat work space of all admitted lattices, the per-job
// residual check with FP64 sums in a fixed order, and the rest-state fill.
#include "batch.h"
#include "kernels_impl.h"

namespace vlbm {

// Rheology of every resident job when they share one (the kernel is then
// specialised to it); kMixedRheo switches per work unit.
constexpr int kMixedRheo = 3;

template <int Q, typename S, int BLOCK, int RHEO>
__global__ __launch_bounds__(BLOCK) void k_batch_step(const BatchJob<typename ComputeOf<S>::type>* __restrict__ jobs,
                                                      const uint32_t* __restrict__ prefix,
                                                      const uint16_t* __restrict__ unit_job,
                                                      const uint32_t* __restrict__ n_jobs, int k) {
    using T = typename ComputeOf<S>::type;
    const uint32_t nj = *n_jobs;
    const uint32_t total = nj ? prefix[nj] : 0u;
    for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint32_t j = unit_job[w];
        const BatchJob<T>& J = jobs[j];
        const Params<T>& p = J.p[(J.phase + k) & 1];
        const uint32_t idx = (w - prefix[j]) * BLOCK + threadIdx.x;
        if (idx >= p.n) continue;
        if (RHEO != kMixedRheo) {
            step_node<Q, S, (RHEO == kMixedRheo ? 0 : RHEO), false, false, false>(p, idx);
        } else if (J.rheo == 0) {
            step_node<Q, S, 0, false, false, false>(p, idx);
        } else if (J.rheo == 1) {
            step_node<Q, S, 1, false, false, false>(p, idx);
        } else {
            step_node<Q, S, 2, false, false, false>(p, idx);
        }
    }
}

template <int Q, typename S, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_check(const BatchJob<typename ComputeOf<S>::type>* __restrict__ jobs,
                                                       const uint32_t* __restrict__ prefix,
                                                       const uint16_t* __restrict__ unit_job,
                                                       const uint32_t* __restrict__ n_jobs, double* partials) {
    using T = typename ComputeOf<S>::type;
    __shared__ double sh[3][BLOCK / 64];
    const uint32_t nj = *n_jobs;
    const uint32_t total = nj ? prefix[nj] : 0u;
    for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint32_t j = unit_job[w];
        const BatchJob<T>& J = jobs[j];
        if (!J.due) continue;                       // uniform per block
        const Params<T>& p = J.p[J.phase];          // current state = source of the next step
        const uint32_t idx = (w - prefix[j]) * BLOCK + threadIdx.x;
        double a = 0, b = 0, c = 0;
        if (idx < p.n) {
            const Node<T> nd = make_node(p, idx);
            const uint32_t word = node_word<false>(p, nd);
            const GlobalPC<Q, S, T, kRepPop> pc{p, nd};
            T g[Q];
            gather<Q, T>(g, pc);
            if (w_used(word) && w_scheme(word) != kNull) apply_bc<Q, T, false>(g, p, nd, word, pc);
            const Moments<Q, 0, T> m = moments<Q, 0, T>(g, p, (T)0, const_force(p));
            const double ux = (double)m.ux, uy = (double)m.uy, uz = (double)m.uz;
            const double dx = ux - (double)J.uold[0][idx], dy = uy - (double)J.uold[1][idx], dz = uz - (double)J.uold[2][idx];
            a = sqrt(dx * dx + dy * dy + dz * dz);
            b = sqrt(ux * ux + uy * uy + uz * uz);
            c = (double)m.rho;
            J.uold[0][idx] = m.ux; J.uold[1][idx] = m.uy; J.uold[2][idx] = m.uz;
        }
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_down(a, off, 64);
            b += __shfl_down(b, off, 64);
            c += __shfl_down(c, off, 64);
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) { sh[0][wave] = a; sh[1][wave] = b; sh[2][wave] = c; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int v = 0; v < BLOCK / 64; v++) { s0 += sh[0][v]; s1 += sh[1][v]; s2 += sh[2][v]; }
            partials[3 * w + 0] = s0;
            partials[3 * w + 1] = s1;
            partials[3 * w + 2] = s2;
        }
        __syncthreads();
    }
}

// one wave per job, fixed order over its work units
template <typename T>
__global__ __launch_bounds__(64) void k_batch_job_sums(const BatchJob<T>* __restrict__ jobs,
                                                       const uint32_t* __restrict__ prefix,
                                                       const double* __restrict__ partials, double* sums) {
    const uint32_t j = blockIdx.x;
    if (!jobs[j].due) return;
    double a = 0, b = 0, c = 0;
    for (uint32_t w = prefix[j] + threadIdx.x; w < prefix[j + 1]; w += 64) {
        a += partials[3 * w]; b += partials[3 * w + 1]; c += partials[3 * w + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    if (threadIdx.x == 0) { sums[3 * j] = a; sums[3 * j + 1] = b; sums[3 * j + 2] = c; }
}

template <int Q, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_fill_rest(S* lat, int64_t stride, uint32_t n, double rho0) {
    using T = typename ComputeOf<S>::type;
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n) return;
#pragma unroll
    for (int i = 0; i < Q; i++) lat[(int64_t)i * stride + idx] = (S)((T)weight<Q>(i) * ((T)rho0 - (T)1));
}

template <int Q, typename S, int BLOCK>
static int step_r(int rheo, uint32_t grid, const void* jobs, const uint32_t* prefix, const uint16_t* unit_job,
                  const uint32_t* nj, int k, hipStream_t st) {
    using T = typename ComputeOf<S>::type;
    const BatchJob<T>* J = (const BatchJob<T>*)jobs;
    const dim3 g(grid), b(BLOCK);
    if (rheo == 0) hipLaunchKernelGGL((k_batch_step<Q, S, BLOCK, 0>), g, b, 0, st, J, prefix, unit_job, nj, k);
    else if (rheo == 1) hipLaunchKernelGGL((k_batch_step<Q, S, BLOCK, 1>), g, b, 0, st, J, prefix, unit_job, nj, k);
    else if (rheo == 2) hipLaunchKernelGGL((k_batch_step<Q, S, BLOCK, 2>), g, b, 0, st, J, prefix, unit_job, nj, k);
    else hipLaunchKernelGGL((k_batch_step<Q, S, BLOCK, kMixedRheo>), g, b, 0, st, J, prefix, unit_job, nj, k);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int Q, typename S>
static int step_s(int block, int rheo, uint32_t grid, const void* jobs, const uint32_t* prefix,
                  const uint16_t* unit_job, const uint32_t* nj, int k, hipStream_t st) {
    if (block == 128) return step_r<Q, S, 128>(rheo, grid, jobs, prefix, unit_job, nj, k, st);
    return step_r<Q, S, 256>(rheo, grid, jobs, prefix, unit_job, nj, k, st);
}

template <int Q, typename S>
static int check_s(int block, uint32_t grid, const void* jobs, const uint32_t* prefix, const uint16_t* unit_job,
                   const uint32_t* nj, uint32_t nj_host, double* partials, double* sums, hipStream_t st) {
    using T = typename ComputeOf<S>::type;
    const BatchJob<T>* J = (const BatchJob<T>*)jobs;
    if (block == 128)
        hipLaunchKernelGGL((k_batch_check<Q, S, 128>), dim3(grid), dim3(128), 0, st, J, prefix, unit_job, nj, partials);
    else
        hipLaunchKernelGGL((k_batch_check<Q, S, 256>), dim3(grid), dim3(256), 0, st, J, prefix, unit_job, nj, partials);
    if (nj_host) hipLaunchKernelGGL((k_batch_job_sums<T>), dim3(nj_host), dim3(64), 0, st, J, prefix, partials, sums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_batch_step(int q, int prec, int block, int rheo, uint32_t grid, const void* jobs, const uint32_t* prefix,
                      const uint16_t* unit_job, const uint32_t* n_jobs, int k, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (block != 128 && block != 256) return -1;
    if (rheo < 0 || rheo > kMixedRheo) return -1;
#define VLBM_STEP(QQ, SS) step_s<QQ, SS>(block, rheo, grid, jobs, prefix, unit_job, n_jobs, k, st)
    if (q == 19) {
        if (prec == 8) return VLBM_STEP(19, double);
        if (prec == 4) return VLBM_STEP(19, float);
        return VLBM_STEP(19, _Float16);
    }
    if (prec == 8) return VLBM_STEP(27, double);
    if (prec == 4) return VLBM_STEP(27, float);
    return VLBM_STEP(27, _Float16);
#undef VLBM_STEP
}

int launch_batch_check(int q, int prec, int block, uint32_t grid, const void* jobs, const uint32_t* prefix,
                       const uint16_t* unit_job, const uint32_t* n_jobs, uint32_t n_jobs_host, double* partials,
                       double* sums, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    if (block != 128 && block != 256) return -1;
#define VLBM_CHECK(QQ, SS) check_s<QQ, SS>(block, grid, jobs, prefix, unit_job, n_jobs, n_jobs_host, partials, sums, st)
    if (q == 19) {
        if (prec == 8) return VLBM_CHECK(19, double);
        if (prec == 4) return VLBM_CHECK(19, float);
        return VLBM_CHECK(19, _Float16);
    }
    if (prec == 8) return VLBM_CHECK(27, double);
    if (prec == 4) return VLBM_CHECK(27, float);
    return VLBM_CHECK(27, _Float16);
#undef VLBM_CHECK
}

int launch_batch_fill_rest(int q, int prec, void* lat, int64_t stride, uint32_t n, double rho0, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((n + kBlock - 1) / kBlock);
#define VLBM_REST(QQ, SS) \
    hipLaunchKernelGGL((k_batch_fill_rest<QQ, SS>), grid, dim3(kBlock), 0, st, (SS*)lat, stride, n, rho0)
    if (q == 19) {
        if (prec == 8) VLBM_REST(19, double); else if (prec == 4) VLBM_REST(19, float); else VLBM_REST(19, _Float16);
    } else {
        if (prec == 8) VLBM_REST(27, double); else if (prec == 4) VLBM_REST(27, float); else VLBM_REST(27, _Float16);
    }
#undef VLBM_REST
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace vlbm
