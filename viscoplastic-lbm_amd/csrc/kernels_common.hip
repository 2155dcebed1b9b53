// Precision routing, initialisation, reduction and conversion kernels.
#include "kernels_impl.h"
#include "ibm.h"

namespace vlbm {

int launch_step_fp64(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream);
int launch_step_fp32(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream);
int launch_step_fp16(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream);
int launch_ibm_macr_fp64(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream);
int launch_ibm_macr_fp32(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream);
int launch_ibm_macr_fp16(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream);
int launch_aux_fp64(int q, int rheo, bool mask, int in_rep, const void* p, void* stream);
int launch_aux_fp32(int q, int rheo, bool mask, int in_rep, const void* p, void* stream);
int launch_aux_fp16(int q, int rheo, bool mask, int in_rep, const void* p, void* stream);
int launch_step_bf16(int q, int rheo, bool mask, bool direct, bool ibm, int in_rep, int out_rep, const void* p,
                     void* stream);
int launch_ibm_macr_bf16(int q, bool mask, int in_rep, const void* p, const uint32_t* list, const uint32_t* count,
                         uint32_t max_tiles, void* stream);
int launch_aux_bf16(int q, int rheo, bool mask, int in_rep, const void* p, void* stream);

__global__ void k_reduce_final(const double* partials, int nblocks, double* out3) {
    // one wave; fixed order -> deterministic
    double a = 0, b = 0, c = 0;
    for (int i = threadIdx.x; i < nblocks; i += 64) {
        a += partials[3 * i]; b += partials[3 * i + 1]; c += partials[3 * i + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    if (threadIdx.x == 0) { out3[0] = a; out3[1] = b; out3[2] = c; }
}

static inline dim3 grid_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

int launch_step(int q, int prec, int rheo, bool mask, bool direct, const void* p, void* stream, bool ibm, int in_rep,
                int out_rep) {
    if (prec == 8) return launch_step_fp64(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
    if (prec == 4) return launch_step_fp32(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
    if (in_rep != kRepPop || out_rep != kRepPop) return -1;
    if (prec == 3) return launch_step_bf16(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
    return launch_step_fp16(q, rheo, mask, direct, ibm, in_rep, out_rep, p, stream);
}

int launch_ibm_macr(int q, int prec, bool mask, const void* p, const void* /*fluid*/, const uint32_t* list,
                    const uint32_t* count, uint32_t max_tiles, void* stream, int in_rep) {
    if (prec == 8) return launch_ibm_macr_fp64(q, mask, in_rep, p, list, count, max_tiles, stream);
    if (prec == 4) return launch_ibm_macr_fp32(q, mask, in_rep, p, list, count, max_tiles, stream);
    if (prec == 3) return launch_ibm_macr_bf16(q, mask, in_rep, p, list, count, max_tiles, stream);
    return launch_ibm_macr_fp16(q, mask, in_rep, p, list, count, max_tiles, stream);
}

int launch_aux(int q, int prec, int rheo, bool mask, const void* p, void* stream, int in_rep) {
    if (prec == 8) return launch_aux_fp64(q, rheo, mask, in_rep, p, stream);
    if (prec == 4) return launch_aux_fp32(q, rheo, mask, in_rep, p, stream);
    if (prec == 3) return launch_aux_bf16(q, rheo, mask, in_rep, p, stream);
    return launch_aux_fp16(q, rheo, mask, in_rep, p, stream);
}

int launch_fill_feq(int q, int prec, void* state, int64_t stride, uint32_t n, const double* rho,
                    const double* ux, const double* uy, const double* uz, void* stream, int out_rep) {
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid = grid_for(n);
#define VLBM_FILL(QQ, SS, RR) hipLaunchKernelGGL((k_fill_feq<QQ, SS, RR>), grid, dim3(kBlock), 0, st, (SS*)state, stride, n, rho, ux, uy, uz)
    if (out_rep == kRepMom) {
        if (prec == 8) VLBM_FILL(19, double, kRepMom);
        else if (prec == 4) VLBM_FILL(19, float, kRepMom);
        else return -1;
    } else if (q == 19) {
        if (prec == 8) VLBM_FILL(19, double, kRepPop); else if (prec == 4) VLBM_FILL(19, float, kRepPop);
        else if (prec == 3) VLBM_FILL(19, bf16, kRepPop); else VLBM_FILL(19, _Float16, kRepPop);
    } else {
        if (prec == 8) VLBM_FILL(27, double, kRepPop); else if (prec == 4) VLBM_FILL(27, float, kRepPop);
        else if (prec == 3) VLBM_FILL(27, bf16, kRepPop); else VLBM_FILL(27, _Float16, kRepPop);
    }
#undef VLBM_FILL
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_reduce(int prec, const void* rho, const void* ux, const void* uy, const void* uz, void* ux_old,
                  void* uy_old, void* uz_old, uint32_t n, double* partials, double* out3, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const int nblocks = 1024;
    if (prec == 8)
        hipLaunchKernelGGL((k_reduce_partial<double>), dim3(nblocks), dim3(kBlock), 0, st, (const double*)rho,
                           (const double*)ux, (const double*)uy, (const double*)uz, (double*)ux_old,
                           (double*)uy_old, (double*)uz_old, n, partials);
    else
        hipLaunchKernelGGL((k_reduce_partial<float>), dim3(nblocks), dim3(kBlock), 0, st, (const float*)rho,
                           (const float*)ux, (const float*)uy, (const float*)uz, (float*)ux_old,
                           (float*)uy_old, (float*)uz_old, n, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, st, partials, nblocks, out3);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Plain streaming copy (16-byte loads and stores), the calibration kernel of
// the traffic counters.
__global__ __launch_bounds__(kBlock) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                       size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) dst[i] = src[i];
}

int launch_copy_probe(const void* src, void* dst, size_t bytes, void* stream) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(kBlock), 0, (hipStream_t)stream, (const float4*)src,
                       (float4*)dst, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_convert(int prec, void* buf, double* dbl, size_t count, int to_double, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)((count + kBlock - 1) / kBlock));
    if (prec == 8) hipLaunchKernelGGL((k_convert<double>), grid, dim3(kBlock), 0, st, (double*)buf, dbl, count, to_double);
    else if (prec == 4) hipLaunchKernelGGL((k_convert<float>), grid, dim3(kBlock), 0, st, (float*)buf, dbl, count, to_double);
    else if (prec == 3) hipLaunchKernelGGL((k_convert<bf16>), grid, dim3(kBlock), 0, st, (bf16*)buf, dbl, count, to_double);
    else hipLaunchKernelGGL((k_convert<_Float16>), grid, dim3(kBlock), 0, st, (_Float16*)buf, dbl, count, to_double);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace vlbm
