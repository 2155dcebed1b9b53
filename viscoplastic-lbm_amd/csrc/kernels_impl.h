// HIP kernels of the LBM time step for gfx950 (templates; instantiated per
// storage precision in kernels_fp64.hip, kernels_fp32.hip, kernels_fp16.hip).
//
// Design (replaces src/CUDA/lbm.cu:3-504 gpuMacrCollisionStream, lbm.cu:600-657
// gpuPopulationsTransfer, lbm.cu:579-598 gpuApplyBC and the pointer swap of
// src/CUDA/structs/populations.h:64-70 with ONE kernel per step):
//
//  * State = post-collision populations stored as deviations g_i = f_i - w_i,
//    SoA planes [i][z][y][x], two lattices A/B (read one, write the other).
//  * Pull streaming: node x gathers g_i from x - c_i with periodic wrap in all
//    three directions. This equals the reference's push + ghost-plane fold, so
//    the gathered values are the reference's post-stream populations.
//  * Boundary conditions are applied in registers right after the gather. They
//    need post-collision values at the node (bounce-back) or its tangential
//    neighbours (free slip), which are exactly what the source lattice holds.
//  * Then moments, apparent viscosity, regularized collision with Guo forcing,
//    and one coalesced store per population.
//
// Every element of the source lattice is read by exactly one node, so the step
// is a streaming kernel: 2*Q*sizeof(storage) bytes per node, HBM bound.

#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include "lattice.h"

namespace vlbm {

template <int Q> struct BBTable { uint32_t m[26]; };
template <int Q>
constexpr BBTable<Q> make_bb_table() {
    BBTable<Q> t{};
    for (int d = 0; d < 26; d++) t.m[d] = bb_mask<Q>(d);
    return t;
}
static __constant__ BBTable<19> c_bb19 = make_bb_table<19>();
static __constant__ BBTable<27> c_bb27 = make_bb_table<27>();

// ---- storage conversions ------------------------------------------------------
// bfloat16 storage: the upper 16 bits of a float (8 exponent bits, 7 stored
// mantissa bits); conversion from float rounds to nearest even.
struct bf16 {
    uint16_t bits;
    bf16() = default;
    __device__ explicit bf16(float f) {
        uint32_t u = __float_as_uint(f);
        if ((u & 0x7f800000u) == 0x7f800000u) {
            bits = (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));   // inf stays inf, NaN stays NaN
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);
            bits = (uint16_t)(u >> 16);
        }
    }
    __device__ explicit bf16(double d) : bf16((float)d) {}
    __device__ operator float() const { return __uint_as_float((uint32_t)bits << 16); }
};

template <typename S, typename T> __device__ __forceinline__ T ld(const S* p) { return (T)(*p); }
template <> __device__ __forceinline__ float ld<_Float16, float>(const _Float16* p) { return (float)(*p); }
template <typename S, typename T> __device__ __forceinline__ T ld_nt(const S* p) {
    return (T)__builtin_nontemporal_load(p);
}
template <> __device__ __forceinline__ float ld_nt<bf16, float>(const bf16* p) {
    return __uint_as_float((uint32_t)__builtin_nontemporal_load((const uint16_t*)p) << 16);
}

template <typename S, typename T> __device__ __forceinline__ void st_nt(S* p, T v) {
    __builtin_nontemporal_store((S)v, p);
}
template <> __device__ __forceinline__ void st_nt<bf16, float>(bf16* p, float v) {
    __builtin_nontemporal_store(bf16(v).bits, (uint16_t*)p);
}

// Stochastic rounding to 16-bit storage: v is rounded up in magnitude with
// probability equal to its distance past the representable value below it,
// in units of the spacing, so the stored value is v in expectation. The
// per-step forcing of a slow flow is smaller than half an FP16/BF16 spacing
// of the populations; rounding to nearest drops it every step, stochastic
// rounding keeps it on average. rnd: uniform random bits (the low 13 for
// FP16, 16 for BF16). Adding them below the kept bits of the float and
// truncating rounds the magnitude up exactly with that probability.
template <typename S>
__device__ __forceinline__ void st_nt_sr(S* p, float v, uint32_t rnd);
template <>
__device__ __forceinline__ void st_nt_sr<_Float16>(_Float16* p, float v, uint32_t rnd) {
    const float a = fabsf(v);
    // normal FP16 range: 13 float mantissa bits below the half's last bit;
    // below it (subnormal halves, spacing 6e-8) and above, to nearest
    if (a >= 6.103515625e-05f && a < 65504.0f) v = __uint_as_float((__float_as_uint(v) + (rnd & 0x1fffu)) & ~0x1fffu);
    __builtin_nontemporal_store((_Float16)v, p);
}
template <>
__device__ __forceinline__ void st_nt_sr<bf16>(bf16* p, float v, uint32_t rnd) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7f800000u) != 0x7f800000u) u += rnd & 0xffffu;   // inf and NaN as they are
    __builtin_nontemporal_store((uint16_t)(u >> 16), (uint16_t*)p);
}

// 32-bit integer mixer (the "lowbias32" finalizer found by C. Wellons' hash
// prospector): consecutive inputs give independent-looking outputs.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

constexpr int kBlock = 256;      // threads (nodes) per block of the node kernels

template <typename S> struct ComputeOf { using type = float; };
template <> struct ComputeOf<double> { using type = double; };

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    return (uint32_t)(((uint64_t)n * f.m) >> f.shift);
}

// Box-builder region of a node, in the reference builders' if-chain priority
// (corners, then edges, then faces; e.g. src/CUDA/boundaryConditionsBuilders/closedBoxBounceBack.cu:33-186).
__device__ __forceinline__ int box_region(int x, int y, int z, int nx, int ny, int nz) {
    const bool s = (y == 0), n = (y == ny - 1), w = (x == 0), e = (x == nx - 1);
    const bool b = (z == 0), f = (z == nz - 1);
    if (!(s | n | w | e | b | f)) return 26;
    if (s && w && b) return 0;
    if (s && w && f) return 1;
    if (s && e && b) return 2;
    if (s && e && f) return 3;
    if (n && w && b) return 4;
    if (n && w && f) return 5;
    if (n && e && b) return 6;
    if (n && e && f) return 7;
    if (s && w) return 8;
    if (s && e) return 9;
    if (n && w) return 10;
    if (n && e) return 11;
    if (s && b) return 12;
    if (s && f) return 13;
    if (n && b) return 14;
    if (n && f) return 15;
    if (w && b) return 16;
    if (w && f) return 17;
    if (e && b) return 18;
    if (e && f) return 19;
    if (s) return 20;
    if (n) return 21;
    if (w) return 22;
    if (e) return 23;
    if (b) return 24;
    return 25;
}

template <typename T>
struct Node {
    int x, y, z;
    int xm, xp, ym, yp, zm, zp;     // periodic neighbours
    uint32_t idx;
};

template <typename T>
__device__ __forceinline__ uint32_t row_off(const Params<T>& p, int ys, int zs) {
    return ((uint32_t)zs * (uint32_t)p.ny + (uint32_t)ys) * (uint32_t)p.nx;
}

// Source node of population i for the node (pull streaming with wrap).
template <typename T>
__device__ __forceinline__ uint32_t src_index(const Params<T>& p, const Node<T>& nd, int i) {
    const int xs = kCX[i] > 0 ? nd.xm : (kCX[i] < 0 ? nd.xp : nd.x);
    const int ys = kCY[i] > 0 ? nd.ym : (kCY[i] < 0 ? nd.yp : nd.y);
    const int zs = kCZ[i] > 0 ? nd.zm : (kCZ[i] < 0 ? nd.zp : nd.z);
    return row_off(p, ys, zs) + (uint32_t)xs;
}

template <typename S, typename T>
__device__ __forceinline__ T load_pop(const Params<T>& p, int i, uint32_t node) {
    const S* base = (const S*)p.src;
    return ld<S, T>(base + (int64_t)i * p.stride + node);
}

// ---------------------------------------------------------------------------
// Boundary conditions on gathered deviations g (post-stream values).
// ---------------------------------------------------------------------------

// Zou-He on a face (D3Q19), generic form: with outward normal n = s*e_a,
// T = sum of populations with c_a = 0, K = sum of outgoing ones (c.n > 0),
//   velocity: rho_w = (T + 2K) / (1 + u.n)
//   pressure: u.n   = (T + 2K) / rho_w - 1, tangential velocity 0
//   incoming i:  f_i = f_opp(i) + 6 w_i rho_w (c_i.u) - sum_t c_it N_t,
//   N_t = 1/2 sum_{c_a=0} f_j c_jt - rho_w u_t / 3.
// Reference-compatible lines (src/CUDA/boundaryConditionsSchemes/D3Q19_VelZouHe.cu,
// D3Q19_PresZouHe.cu), used unless zouhe_textbook is set:
//   * velocity BACK uses 1/(1+uz_w) (VelZouHe.cu:153) and u_x in place of u_y
//     for populations 11 and 18 (VelZouHe.cu:168-171);
//   * velocity WEST population 15 uses u_y in place of u_z (VelZouHe.cu:95);
//   * EAST (velocity and pressure) takes T + 2K from the WEST face's known set
//     (its own incoming populations) and adds to each incoming population its
//     own streamed value the WEST increment of its opposite
//     (VelZouHe.cu:104-119, PresZouHe.cu:103-119).
// Zou-He works on deviations unchanged: the weights of T + 2K sum to 1 and
// the tangential sums are antisymmetric.
template <typename T>
__device__ __forceinline__ void zou_he(T* g, int dir, bool pressure, T uxw, T uyw, T uzw, T rhow_in,
                                       bool compat) {
    if (dir > dB) return;
    const int a = (dir == dN || dir == dS) ? 1 : ((dir == dW || dir == dE) ? 0 : 2);
    const int s = (dir == dN || dir == dE || dir == dF) ? 1 : -1;
    auto ca = [&](int i) -> int { return a == 0 ? kCX[i] : (a == 1 ? kCY[i] : kCZ[i]); };
    // the reference's EAST lines sum the WEST face's known set (c_x = -1, the
    // EAST face's incoming populations) into T + 2K (VelZouHe.cu:104-108,
    // PresZouHe.cu:103-107)
    const int known = (compat && dir == dE) ? -s : s;
    T tsum = 0, ksum = 0, n_t[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const int c = ca(i);
        if (c == 0) {
            tsum += g[i];
            n_t[0] += (T)kCX[i] * g[i];
            n_t[1] += (T)kCY[i] * g[i];
            n_t[2] += (T)kCZ[i] * g[i];
        } else if (c == known) {
            ksum += g[i];
        }
    }
    const T tk = tsum + 2 * ksum;     // (T + 2K) - 1 in deviations
    T u[3] = {uxw, uyw, uzw};
    T rho_w;
    if (pressure) {
        rho_w = rhow_in;
        const T un = (tk + ((T)1 - rho_w)) / rho_w;   // u.n
        u[0] = u[1] = u[2] = 0;
        u[a] = (T)s * un;
    } else {
        T denom = (T)1 + (T)s * u[a];
        if (compat && dir == dB) denom = (T)1 + uzw;
        rho_w = ((T)1 + tk) / denom;
    }
    T nt[3];
#pragma unroll
    for (int t = 0; t < 3; t++) nt[t] = (T)0.5 * n_t[t] - (pressure ? (T)0 : rho_w * u[t] / (T)3);
    nt[a] = 0;
    const bool east_compat = compat && dir == dE;
    T out[19];
#pragma unroll
    for (int i = 1; i < 19; i++) out[i] = g[i];
#pragma unroll
    for (int i = 1; i < 19; i++) {
        if (ca(i) != -s) continue;                  // incoming populations only
        const int j = east_compat ? opp(i) : i;     // population whose formula is used
        T cu;
        if (compat && !pressure && dir == dB && (j == 11 || j == 18))
            cu = (T)kCZ[j] * u[2] + (T)kCY[j] * u[0];
        else if (compat && !pressure && (dir == dW || dir == dE) && j == 15)
            cu = (T)kCX[j] * u[0] + (T)kCZ[j] * u[1];
        else
            cu = (T)kCX[j] * u[0] + (T)kCY[j] * u[1] + (T)kCZ[j] * u[2];
        const T corr = (T)kCX[j] * nt[0] + (T)kCY[j] * nt[1] + (T)kCZ[j] * nt[2];
        const T inc = (T)(6.0 * weight<19>(j)) * rho_w * cu - corr;
        out[i] = (east_compat ? g[i] : g[opp(i)]) + inc;
    }
#pragma unroll
    for (int i = 1; i < 19; i++) g[i] = out[i];
}

// Velocity bounce-back (src/CUDA/boundaryConditionsSchemes/D3Q19_VelBounceBack.cu:27-149):
// f_i = f_opp(i) + 6 rho w_i (c_i . u_w), everything post-stream, rho summed
// over all post-stream populations of the node.
template <int Q, typename T>
__device__ __forceinline__ void vel_bounce_back(T* g, int dir, T uxw, T uyw, T uzw) {
    if (dir > dB) return;
    T dr = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) dr += g[i];
    const T rho = (T)1 + dr;
    const unsigned faces = kDirFaces[dir];
    T out[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) out[i] = g[i];
#pragma unroll
    for (int i = 1; i < Q; i++) {
        if (!incoming(i, faces)) continue;
        const T cu = (T)kCX[i] * uxw + (T)kCY[i] * uyw + (T)kCZ[i] * uzw;
        out[i] = g[opp(i)] + (T)(6.0 * weight<Q>(i)) * rho * cu;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) g[i] = out[i];
}

// Interpolated bounce-back (Bouzidi, linear) against the duct circle(s)
// (src/CUDA/boundaryConditionsSchemes/interpolatedBounceBack.cu:6-431).
// q is the normalised distance node -> wall along the D2Q9 projection of the
// incoming population, evaluated with the reference's distPoints2D (first
// difference cast to float, src/CUDA/globalFunctions.h:158-162).
template <typename T>
__device__ __forceinline__ T dist2d_ref(T x1, T y1, T x2, T y2) {
    return (T)sqrt((T)((float)(x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2)));
}
template <typename T>
__device__ T wall_q(int kind, T x1, T y1, T x2, T y2, T R) {
    if (kind == 0) {           // along x
        T y = y2, x = sqrt(R * R - y * y);
        if (x2 < 0) x = -x;
        return dist2d_ref(x, y, x1, y1);
    } else if (kind == 1) {    // along y
        T x = x2, y = sqrt(R * R - x * x);
        if (y2 < 0) y = -y;
        return dist2d_ref(x, y, x1, y1);
    } else if (kind == 2) {    // along the (1,1) diagonal
        const T b = y1 - x1, dr = sqrt(2 * R * R - b * b);
        T x = (-b + dr) * (T)0.5;
        if ((x1 > x2 && (x > x1 || x < x2)) || (x1 < x2 && (x < x1 || x > x2))) x = (-b - dr) * (T)0.5;
        return dist2d_ref(x, x + b, x1, y1) / (T)1.41421356237309504880;
    } else {                   // along the (1,-1) diagonal
        const T b = y1 + x1, dr = sqrt(2 * R * R - b * b);
        T x = (b + dr) * (T)0.5;
        if ((x1 > x2 && (x > x1 || x < x2)) || (x1 < x2 && (x < x1 || x > x2))) x = (b - dr) * (T)0.5;
        return dist2d_ref(x, -x + b, x1, y1) / (T)1.41421356237309504880;
    }
}

template <int Q, typename T, class PC>
__device__ void interp_bounce_back(T* g, const Params<T>& p, const Node<T>& nd, unsigned unknown, bool inside,
                                   PC pc) {
    // D2Q9 projections in the reference's bit order (nodeTypeMap.h:100-107)
    const int dxs[8] = {1, 0, -1, 0, 1, -1, -1, 1};
    const int dys[8] = {0, 1, 0, -1, 1, 1, -1, -1};
    const int kinds[8] = {0, 1, 0, 1, 2, 3, 2, 3};
    const T xN = (T)nd.x - (T)(p.nx - 1) / (T)2, yN = (T)nd.y - (T)(p.ny - 1) / (T)2;
    const T rr = sqrt(xN * xN + yN * yN);
    const T c = xN / rr, s = yN / rr;
    T radius, uxw, uyw, uzw;
    if (inside) {
        radius = p.duct_r; uxw = -p.w_in * radius * s; uyw = p.w_in * radius * c; uzw = p.u_in;
    } else {
        const T wo = p.w_out * ((T)nd.z / (T)p.nz);
        radius = p.duct_R; uxw = -wo * radius * s; uyw = wo * radius * c; uzw = p.u_out;
    }
    for (int b = 0; b < 8; b++) {
        if (!(unknown & (1u << b))) continue;
        const T qd = wall_q<T>(kinds[b], xN, yN, xN - (T)dxs[b], yN - (T)dys[b], radius);
#pragma unroll
        for (int i = 1; i < Q; i++) {
            if (kCX[i] != dxs[b] || kCY[i] != dys[b]) continue;
            const int j = opp(i);
            const T cu = (T)kCX[i] * uxw + (T)kCY[i] * uyw + (T)kCZ[i] * uzw;
            const T df = (T)weight<Q>(i) * cu;
            const T fj = pc(j, 0, 0, 0);
            if (qd > (T)0.5) {
                const T fi = pc(i, 0, 0, 0);
                g[i] = ((T)0.5 / qd) * fj + ((qd - (T)0.5) / qd) * fi + (T)3 * df / qd;
            } else {
                // the reference reads x + c_i unwrapped; wrapping keeps degenerate
                // (1-2 node wide) lattices in bounds and is the same node otherwise
                T fa = pc(j, kCX[i], kCY[i], kCZ[i]);
                // rheometer: the reference refills an unused plane's populations
                // toward the plates from its extra z plane (the rest state) in
                // every step's stream, so only the first boundary pass sees the
                // initial values there
                if (p.plates && !p.first_bc) {
                    const int za = (int)nd.z + kCZ[i];
                    if (za == 0 || za == p.nz - 1) fa = (T)0;
                }
                g[i] = 2 * qd * fj + (1 - 2 * qd) * fa + (T)6 * df;
            }
        }
    }
}


// ---------------------------------------------------------------------------
// Per-node pieces
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Node<T> make_node(const Params<T>& p, uint32_t idx) {
    Node<T> nd;
    const uint32_t yz = fdiv(idx, p.div_nx);
    nd.x = (int)(idx - yz * (uint32_t)p.nx);
    nd.z = (int)fdiv(idx, p.div_nxy);
    nd.y = (int)(yz - (uint32_t)nd.z * (uint32_t)p.ny);
    nd.idx = idx;
    nd.xm = nd.x == 0 ? p.nx - 1 : nd.x - 1;
    nd.xp = nd.x == p.nx - 1 ? 0 : nd.x + 1;
    nd.ym = nd.y == 0 ? p.ny - 1 : nd.y - 1;
    nd.yp = nd.y == p.ny - 1 ? 0 : nd.y + 1;
    nd.zm = nd.z == 0 ? p.nz - 1 : nd.z - 1;
    nd.zp = nd.z == p.nz - 1 ? 0 : nd.z + 1;
    return nd;
}

template <bool MASK, typename T>
__device__ __forceinline__ uint32_t node_word(const Params<T>& p, const Node<T>& nd) {
    if (MASK) return p.mask[nd.idx];
    return p.region_word[box_region(nd.x, nd.y, nd.z, p.nx, p.ny, p.nz)];
}

// ---------------------------------------------------------------------------
// State representations. kRepPop: post-collision deviations g*_i, Q planes.
// kRepMom: the regularized post-collision state as 10 coefficients per node,
//   g*_i = w_i (A + B.c_i + sum_a c_ia^2 G_aa + sum_{a<b} c_ia c_ib G_ab),
// planes in the order A, Bx, By, Bz, Gxx, Gyy, Gzz, Gxy, Gxz, Gyz. The
// regularized collision (and the equilibrium) is a quadratic polynomial in
// c_i, so the coefficients hold the post-collision populations exactly.
// ---------------------------------------------------------------------------

// g*_i of the node from its 10 coefficients at c (element stride st).
template <int Q, typename T, typename P>
__device__ __forceinline__ T recon(int i, const P* c, int64_t st) {
    const int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
    T v = (T)c[0];
    if (cx) v += (T)cx * (T)c[1 * st] + (T)c[4 * st];
    if (cy) v += (T)cy * (T)c[2 * st] + (T)c[5 * st];
    if (cz) v += (T)cz * (T)c[3 * st] + (T)c[6 * st];
    if (cx * cy) v += (T)(cx * cy) * (T)c[7 * st];
    if (cx * cz) v += (T)(cx * cz) * (T)c[8 * st];
    if (cy * cz) v += (T)(cy * cz) * (T)c[9 * st];
    return (T)weight<Q>(i) * v;
}

// Node at offset (dx, dy, dz) in {-1, 0, 1}^3 from nd, wrapped.
template <typename T>
__device__ __forceinline__ uint32_t node_at(const Params<T>& p, const Node<T>& nd, int dx, int dy, int dz) {
    const int x = dx > 0 ? nd.xp : (dx < 0 ? nd.xm : nd.x);
    const int y = dy > 0 ? nd.yp : (dy < 0 ? nd.ym : nd.y);
    const int z = dz > 0 ? nd.zp : (dz < 0 ? nd.zm : nd.z);
    return row_off(p, y, z) + (uint32_t)x;
}

// Post-collision population i of the node at offset (dx, dy, dz), read from
// the source lattice in global memory.
template <int Q, typename S, typename T, int IN>
struct GlobalPC {
    const Params<T>& p;
    Node<T> nd;
    __device__ __forceinline__ T operator()(int i, int dx, int dy, int dz) const {
        const uint32_t node = node_at(p, nd, dx, dy, dz);
        if (IN == kRepPop) return load_pop<S, T>(p, i, node);
        return recon<Q, T, S>(i, (const S*)p.src + node, p.stride);
    }
    // coefficient states: the node's own 10 coefficients, read once for the
    // several populations a bounce-back rebuilds from them
    static constexpr bool kCoefs = IN == kRepMom;
    __device__ __forceinline__ void own(T* c) const {
#pragma unroll
        for (int k = 0; k < kCoef; k++) c[k] = (T)((const S*)p.src)[(int64_t)k * p.stride + nd.idx];
    }
    // Population i of this node in the idle lattice (its post-collision value
    // of the step before the last); read before the step overwrites it.
    __device__ __forceinline__ T held(int i) const {
        return ld<S, T>((const S*)p.held + (int64_t)i * p.stride + nd.idx);
    }
};

// The D3Q19 population pairs (c_x = +1, c_x = -1): slot of population i
// among them (-1 for c_x = 0).
__host__ __device__ constexpr int xpair_slot(int i) {
    return (i == 1 || i == 2) ? 0 : (i == 7 || i == 8) ? 1 : (i == 9 || i == 10) ? 2
         : (i == 13 || i == 14) ? 3 : (i == 15 || i == 16) ? 4 : -1;
}

// Gather (pull streaming) of the post-stream deviations: g_i(x) = g*_i(x - c_i).
template <int Q, typename T, class PC>
__device__ __forceinline__ void gather(T* g, PC pc) {
#pragma unroll
    for (int i = 0; i < Q; i++) g[i] = pc(i, -kCX[i], -kCY[i], -kCZ[i]);
}

// Pull streaming of populations with the x shift done in LDS. Each thread of
// the block (kBlock consecutive nodes) loads, for every population i, the
// value at its own x of the source row (y - c_iy, z - c_iz), so the block
// reads whole aligned row segments (non-temporal: each value is read once per
// step). The populations with c_ix = +-1 go through LDS: after one barrier a
// node takes the value loaded by the thread of its source node x - c_ix
// (periodic wrap included) when that node is in the block, which is every
// node when rows are whole blocks; the others load it after the barrier.
// Every thread of the block calls this; threads past the lattice pass a valid
// node and their values are never taken. Compute-free pull copies of a 256^3
// D3Q19 FP32 lattice: 39.0 GLUPS this way, 37.6 with per-thread shifted loads.
// The launcher uses it for D3Q19 when NX divides kBlock (no node then loads
// after the barrier): 256^3 D3Q19 FP32 Bingham plates 39.5 GLUPS against
// 37.6 with gather(). Row ends inside blocks (185x119x81: 25.5 against 29.1)
// and D3Q27 (18 LDS rows, 22.6 against 25.0) are faster with gather().
template <int Q> struct XShifted {
    static constexpr int count() {
        int c = 0;
        for (int i = 0; i < Q; i++) c += kCX[i] != 0;
        return c;
    }
    static constexpr int slot(int i) {          // LDS row of population i (c_ix != 0)
        int c = 0;
        for (int k = 0; k < i; k++) c += kCX[k] != 0;
        return c;
    }
};

template <int Q, typename S, typename T>
__device__ __forceinline__ void gather_rows(T* g, const Params<T>& p, const Node<T>& nd, T* sm) {
    const int t = (int)threadIdx.x;
    const uint32_t row = row_off(p, nd.y, nd.z) - blockIdx.x * (uint32_t)kBlock;
    const uint32_t t_left = row + (uint32_t)nd.xm, t_right = row + (uint32_t)nd.xp;   // threads of x -+ 1
    const bool left_in = t_left < (uint32_t)kBlock, right_in = t_right < (uint32_t)kBlock;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int ys = kCY[i] > 0 ? nd.ym : (kCY[i] < 0 ? nd.yp : nd.y);
        const int zs = kCZ[i] > 0 ? nd.zm : (kCZ[i] < 0 ? nd.zp : nd.z);
        g[i] = ld_nt<S, T>((const S*)p.src + (int64_t)i * p.stride + row_off(p, ys, zs) + nd.x);
    }
#pragma unroll
    for (int i = 0; i < Q; i++)
        if (kCX[i] != 0) sm[XShifted<Q>::slot(i) * kBlock + t] = g[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < Q; i++) {
        if (kCX[i] == 0) continue;
        if (kCX[i] > 0 ? left_in : right_in) {
            g[i] = sm[XShifted<Q>::slot(i) * kBlock + (int)(kCX[i] > 0 ? t_left : t_right)];
        } else {
            const int ys = kCY[i] > 0 ? nd.ym : (kCY[i] < 0 ? nd.yp : nd.y);
            const int zs = kCZ[i] > 0 ? nd.zm : (kCZ[i] < 0 ? nd.zp : nd.z);
            g[i] = ld<S, T>((const S*)p.src + (int64_t)i * p.stride + row_off(p, ys, zs) +
                            (kCX[i] > 0 ? nd.xm : nd.xp));
        }
    }
}

// Post-stream state read in place (after an initialisation or a load).
template <int Q, typename T, class PC>
__device__ __forceinline__ void load_direct(T* g, PC pc) {
#pragma unroll
    for (int i = 0; i < Q; i++) g[i] = pc(i, 0, 0, 0);
}

// gpuSchSpecial of the rheometer (parallelPlatesRheometer.cu:139-198), on
// deviations (opposite populations have equal weights, so the reference's
// lines carry over): velocity bounce-back from the node's own post-stream
// values, the FRONT plate (z = NZ-2) turning at the outer rotation about the
// duct axis, the BACK plate (z = 1) at rest; the wall density is the sum of
// all 19 post-stream populations (the five from the unused plane are the
// held values of plate_held).
template <int Q, typename T>
__device__ __forceinline__ void rheometer_plate(T* g, const Params<T>& p, const Node<T>& nd, bool front) {
    const T xN = (T)((double)nd.x - (p.nx - 1) / 2.0), yN = (T)((double)nd.y - (p.ny - 1) / 2.0);
    const T rr = sqrt(xN * xN + yN * yN);
    const T c = xN / (rr), s = yN / (rr);
    const T wr = front ? p.w_out : (T)0;
    T rho_w = 0;
#pragma unroll
    for (int i = 0; i < 19; i++) rho_w = rho_w + (g[i] + (T)weight<19>(i));
    const T W1 = (T)weight<19>(1), W2 = (T)weight<19>(7);
    const T ux_w = -wr * rr * s, uy_w = wr * rr * c, uz_w = 0;
    if (front) {
        g[6] = g[5] - 6 * rho_w * W1 * (uz_w);
        g[10] = g[9] - 6 * rho_w * W2 * (uz_w + ux_w);
        g[12] = g[11] - 6 * rho_w * W2 * (uz_w + uy_w);
        g[15] = g[16] - 6 * rho_w * W2 * (uz_w - ux_w);
        g[17] = g[18] - 6 * rho_w * W2 * (uz_w - uy_w);
    } else {
        g[5] = g[6] - 6 * rho_w * W1 * (-uz_w);
        g[9] = g[10] - 6 * rho_w * W2 * (-uz_w - ux_w);
        g[11] = g[12] - 6 * rho_w * W2 * (-uz_w - uy_w);
        g[16] = g[15] - 6 * rho_w * W2 * (-uz_w + ux_w);
        g[18] = g[17] - 6 * rho_w * W2 * (-uz_w + uy_w);
    }
}

// Rheometer plate nodes (z = 1, NZ-2; every one is a boundary node): no node
// on the unused planes z = 0, NZ-1 collides or streams, and the reference's
// stream step writes each boundary node's post-collision populations back
// into the buffer it read (gpuMacrCollisionStream with the post-collision
// save), so a population whose source is an unused plane arrives as the
// node's own post-collision value of the step before. In the pull scheme
// that value is the node's slot in the idle lattice.
template <int Q, typename T, class PC>
__device__ __forceinline__ void plate_held(T* g, const Params<T>& p, const Node<T>& nd, PC pc) {
#pragma unroll
    for (int i = 1; i < Q; i++) {
        if (kCZ[i] == 0) continue;
        const int zs = (int)nd.z - kCZ[i];
        if (zs == 0 || zs == p.nz - 1) g[i] = pc.held(i);
    }
}

// Boundary condition of the node's word on the gathered populations
// (dispatch of src/CUDA/boundaryConditionsHandler.cu:25-79); pc gives the
// post-collision populations of the node and its neighbours. MASK: the words
// come from a per-node map (duct and rheometer builders). Box builders' words
// never carry interpolated bounce-back or the rheometer's special scheme, so
// their kernels are compiled without that code.
template <int Q, typename T, bool MASK, class PC>
__device__ __forceinline__ void apply_bc(T* g, const Params<T>& p, const Node<T>& nd, uint32_t word, PC pc) {
    const int scheme = w_scheme(word);
    const int dir = w_dir(word);
    if constexpr (MASK) {
        if (p.plates) plate_held<Q, T>(g, p, nd, pc);
    }
    if (scheme == kBounceBack) {
        // convex edges/corners carry no bounce-back (boundaryConditionsHandler.cu:158-256)
        if (dir >= dNW && w_geom(word)) return;
        if (dir > dSEB) return;
        const uint32_t m = (Q == 19) ? c_bb19.m[dir] : c_bb27.m[dir];
        if constexpr (PC::kCoefs && Q == 19) {
            T c[kCoef];
            pc.own(c);
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (m & (1u << i)) g[i] = recon<Q, T, T>(opp(i), c, 1);
        } else {
#pragma unroll
            for (int i = 1; i < Q; i++)
                if (m & (1u << i)) g[i] = pc(opp(i), 0, 0, 0);
        }
    } else if (scheme == kFreeSlip) {
        // specular reflection: incoming i comes from the mirror population m
        // of the node displaced by minus the tangential part of c_i
        // (src/CUDA/boundaryConditionsSchemes/freeSlip.cu:26-168)
        if (dir > dB) return;
        const unsigned faces = kDirFaces[dir];
#pragma unroll
        for (int i = 1; i < Q; i++) {
            if (!incoming(i, faces)) continue;
            // the source node is displaced by minus the tangential part of c_i
            int mx = kCX[i], my = kCY[i], mz = kCZ[i];
            int ox = -mx, oy = -my, oz = -mz;
            if (dir == dN || dir == dS) {
                my = -my;
                oy = 0;
            } else if (dir == dW || dir == dE) {
                mx = -mx;
                ox = 0;
            } else {
                mz = -mz;
                oz = 0;
            }
            const int m = find_pop(mx, my, mz);
            g[i] = pc(m, ox, oy, oz);
        }
    } else if (scheme == kVelBB) {
        if (Q == 19)
            vel_bounce_back<Q, T>(g, dir, p.ux_bc[w_uxi(word)], p.uy_bc[w_uyi(word)], p.uz_bc[w_uzi(word)]);
    } else if (scheme == kVelZouHe || scheme == kPresZouHe) {
        if (Q == 19)
            zou_he<T>(g, dir, scheme == kPresZouHe, p.ux_bc[w_uxi(word)], p.uy_bc[w_uyi(word)],
                      p.uz_bc[w_uzi(word)], p.rho_bc[w_rhoi(word)], !p.zouhe_textbook);
    } else if constexpr (MASK) {
        if (scheme == kInterpBB) {
            interp_bounce_back<Q, T>(g, p, nd, w_interp(word), dir == dS, pc);
        } else if (scheme == kSpecial) {
            if (Q == 19 && (dir == dF || dir == dB)) rheometer_plate<Q, T>(g, p, nd, dir == dF);
        }
    }
}

// Moments and apparent relaxation frequency of the (post-BC) deviations.
template <int Q, int RHEO, typename T>
struct Moments {
    T dr, rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz, omega;
    T tau_t;     // LES eddy relaxation time (0 without LES)
};

// Body force acting on a node: the constant force, or the per-node field of
// the immersed boundary stage.
template <typename T>
struct Force {
    T x, y, z;
};

template <typename T>
__device__ __forceinline__ Force<T> const_force(const Params<T>& p) { return Force<T>{p.fx, p.fy, p.fz}; }

// Tile of the IBM stage holding the node.
template <typename T>
__device__ __forceinline__ uint32_t tile_of(const Params<T>& p, int x, int y, int z) {
    return tile_index(x, y, z, p.tiles_x, p.tiles_y);
}

// omega_old: power-law state of the node (ignored otherwise).
template <int Q, int RHEO, typename T>
__device__ __forceinline__ Moments<Q, RHEO, T> moments(const T* g, const Params<T>& p, T omega_old,
                                                       const Force<T>& F) {
    Moments<Q, RHEO, T> m;
    T dr = 0, jx = 0, jy = 0, jz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const T v = g[i];
        dr += v;
        if (kCX[i]) { jx += (T)kCX[i] * v; sxx += v; }
        if (kCY[i]) { jy += (T)kCY[i] * v; syy += v; }
        if (kCZ[i]) { jz += (T)kCZ[i] * v; szz += v; }
        if (kCX[i] * kCY[i]) sxy += (T)(kCX[i] * kCY[i]) * v;
        if (kCX[i] * kCZ[i]) sxz += (T)(kCX[i] * kCZ[i]) * v;
        if (kCY[i] * kCZ[i]) syz += (T)(kCY[i] * kCZ[i]) * v;
    }
    const T rho = (T)1 + dr;
    const T inv = (T)1 / rho;
    const T ux = (jx + (T)0.5 * F.x) * inv;
    const T uy = (jy + (T)0.5 * F.y) * inv;
    const T uz = (jz + (T)0.5 * F.z) * inv;
    // Pi^neq = sum f c c - rho (u u + I/3) = sum g c c - rho u u - (rho-1)/3 I
    const T dr3 = dr / (T)3;
    m.pxx = sxx - rho * ux * ux - dr3;
    m.pyy = syy - rho * uy * uy - dr3;
    m.pzz = szz - rho * uz * uz - dr3;
    m.pxy = sxy - rho * ux * uy;
    m.pxz = sxz - rho * ux * uz;
    m.pyz = syz - rho * uy * uz;
    m.dr = dr; m.rho = rho; m.ux = ux; m.uy = uy; m.uz = uz;
    m.tau_t = 0;
    T om = p.omega;
    if (RHEO != 0 || p.les_c != (T)0) {
        // |Pi^neq + sym(u F)| (src/CUDA/lbm.cu:192-208)
        const T axx = m.pxx + ux * F.x, ayy = m.pyy + uy * F.y, azz = m.pzz + uz * F.z;
        const T axy = m.pxy + (ux * F.y + uy * F.x) * (T)0.5;
        const T axz = m.pxz + (ux * F.z + uz * F.x) * (T)0.5;
        const T ayz = m.pyz + (uy * F.z + uz * F.y) * (T)0.5;
        const T smag = sqrt((T)0.5 * (axx * axx + ayy * ayy + azz * azz + 2 * (axy * axy + axz * axz + ayz * ayz)));
        if (RHEO == 0) {
            // implicit Smagorinsky closure (LES_MODEL, src/CUDA/lbm.cu:226-242):
            // tau_t = (sqrt(tau^2 + C |S|) - tau) / 2, omega = 1 / (tau + tau_t)
            const T tau_t = (T)0.5 * sqrt(p.tau * p.tau + p.les_c * smag) - (T)0.5 * p.tau;
            m.tau_t = tau_t;
            om = (T)1 / (p.tau + tau_t);
        } else if (RHEO == 1) {
            // Bingham: omega = omega_p max(0, 1 - tau_y/|S|) (src/CUDA/NNF/nnf.h:83);
            // a zero yield stress is the Newtonian plastic fluid.
            if (p.s_y == (T)0) {
                om = p.omega_p;
            } else {
                const T t = (T)1 - p.s_y / smag;
                om = p.omega_p * ((T)0 > t ? (T)0 : t);
            }
        } else {
            // power law with the previous step's omega (src/CUDA/NNF/nnf.h:38-53)
            T eta = ((T)1 / omega_old - (T)0.5) / (T)3;
            const T gdot = ((T)1 - (T)0.5 * omega_old) * smag / eta;
            eta = (gdot <= p.gamma_0) ? p.k_cons : p.k_cons * pow(gdot, p.n_minus_1);
            om = (T)1 / ((T)0.5 + (T)3 * eta);
        }
        if (RHEO != 0 && p.clamp) {
            // tau in [tau_min, tau_max]  <=>  omega in [omega_lo, omega_hi]
            if (!(om <= p.omega_hi)) om = p.omega_hi;
            else if (om < p.omega_lo) om = p.omega_lo;
        }
    }
    m.omega = om;
    return m;
}

// Regularized collision with Guo forcing, in deviation form:
//   g*_i = w_i [ (rho-1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u) ]
//        + (1-omega) 4.5 w_i (Q_i : Pi^neq - c_i.F/3)
//        + (1-omega/2) w_i (3 c.F + 9 (c.u)(c.F) - 3 u.F)
// with Q_i = c_i c_i - I/3 (src/CUDA/lbm.cu:263-461).
template <int Q, int RHEO, typename T>
__device__ __forceinline__ void collide(T* g, const Moments<Q, RHEO, T>& m, const Force<T>& F) {
    const T om = m.omega;
    const T a = (T)4.5 * ((T)1 - om);
    const T b = (T)1 - (T)0.5 * om;
    const T uf = m.ux * F.x + m.uy * F.y + m.uz * F.z;
    const T usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
    const T tr = m.pxx + m.pyy + m.pzz;
    const T k0 = m.dr - (T)1.5 * m.rho * usq - a * tr / (T)3 - (T)3 * b * uf;
    const T r3 = (T)3 * m.rho, r45 = (T)4.5 * m.rho;
    const T a3 = a / (T)3, b3 = (T)3 * b, b9 = (T)9 * b;
    const T pxy2 = 2 * m.pxy, pxz2 = 2 * m.pxz, pyz2 = 2 * m.pyz;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        const T cu = (T)cx * m.ux + (T)cy * m.uy + (T)cz * m.uz;
        const T cf = (T)cx * F.x + (T)cy * F.y + (T)cz * F.z;
        T cpc = 0;
        if (cx) cpc += m.pxx;
        if (cy) cpc += m.pyy;
        if (cz) cpc += m.pzz;
        if (cx * cy) cpc += (T)(cx * cy) * pxy2;
        if (cx * cz) cpc += (T)(cx * cz) * pxz2;
        if (cy * cz) cpc += (T)(cy * cz) * pyz2;
        const T v = k0 + cu * (r3 + r45 * cu + b9 * cf) + a * cpc + (b3 - a3) * cf;
        g[i] = (T)weight<Q>(i) * v;
    }
}

// BGK collision with Guo forcing in deviation form
// (src/CUDA/collisionSchemes/recursiveCollisionForceNoReg.txt, with the
// apparent omega of the rheology):
//   g*_i = (1-omega) g_i + omega w_i [ (rho-1) + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u) ]
//        + (1-omega/2) w_i (3 c.F + 9 (c.u)(c.F) - 3 u.F)
template <int Q, int RHEO, typename T>
__device__ __forceinline__ void collide_bgk(T* g, const Moments<Q, RHEO, T>& m, const Force<T>& F) {
    const T om = m.omega;
    const T keep = (T)1 - om;
    const T b = (T)1 - (T)0.5 * om;
    const T uf = m.ux * F.x + m.uy * F.y + m.uz * F.z;
    const T usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
    const T e0 = m.dr - (T)1.5 * m.rho * usq;
    const T f0 = -(T)3 * uf;
    const T r3 = (T)3 * m.rho, r45 = (T)4.5 * m.rho;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int cx = kCX[i], cy = kCY[i], cz = kCZ[i];
        const T cu = (T)cx * m.ux + (T)cy * m.uy + (T)cz * m.uz;
        const T cf = (T)cx * F.x + (T)cy * F.y + (T)cz * F.z;
        const T eq = e0 + cu * (r3 + r45 * cu);
        const T fo = f0 + (T)3 * cf + (T)9 * cu * cf;
        g[i] = keep * g[i] + (T)weight<Q>(i) * (om * eq + b * fo);
    }
}

// Regularized collision in coefficient form (kRepMom): the same g*_i as
// collide(), as A + B.c + c^T G c with
//   A    = (rho-1) - 1.5 rho u.u - a tr(Pi)/3 - 3 b u.F
//   B_a  = 3 rho u_a + (3 b - a/3) F_a
//   G_aa = 4.5 rho u_a^2 + 9 b u_a F_a + a Pi_aa
//   G_ab = 9 rho u_a u_b + 9 b (u_a F_b + u_b F_a) + 2 a Pi_ab      (a < b)
// with a = 4.5 (1 - omega), b = 1 - omega/2.
template <int Q, int RHEO, typename T>
__device__ __forceinline__ void coeffs_collide(T* c, const Moments<Q, RHEO, T>& m, const Force<T>& F) {
    const T om = m.omega;
    const T a = (T)4.5 * ((T)1 - om);
    const T b = (T)1 - (T)0.5 * om;
    const T uf = m.ux * F.x + m.uy * F.y + m.uz * F.z;
    const T usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
    const T tr = m.pxx + m.pyy + m.pzz;
    const T r3 = (T)3 * m.rho, r45 = (T)4.5 * m.rho, r9 = (T)9 * m.rho;
    const T bf = (T)3 * b - a / (T)3, b9 = (T)9 * b, a2 = (T)2 * a;
    c[0] = m.dr - (T)1.5 * m.rho * usq - a * tr / (T)3 - (T)3 * b * uf;
    c[1] = r3 * m.ux + bf * F.x;
    c[2] = r3 * m.uy + bf * F.y;
    c[3] = r3 * m.uz + bf * F.z;
    c[4] = r45 * m.ux * m.ux + b9 * m.ux * F.x + a * m.pxx;
    c[5] = r45 * m.uy * m.uy + b9 * m.uy * F.y + a * m.pyy;
    c[6] = r45 * m.uz * m.uz + b9 * m.uz * F.z + a * m.pzz;
    c[7] = r9 * m.ux * m.uy + b9 * (m.ux * F.y + m.uy * F.x) + a2 * m.pxy;
    c[8] = r9 * m.ux * m.uz + b9 * (m.ux * F.z + m.uz * F.x) + a2 * m.pxz;
    c[9] = r9 * m.uy * m.uz + b9 * (m.uy * F.z + m.uz * F.y) + a2 * m.pyz;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Moments, apparent omega, saved macroscopics and collision of one node whose
// post-stream, post-BC deviations are g; the post-collision state goes to
// dst in representation OUT.
template <int Q, typename S, int RHEO, bool IBM, int OUT>
__device__ __forceinline__ void collide_store(const Params<typename ComputeOf<S>::type>& p, uint32_t idx,
                                              const Node<typename ComputeOf<S>::type>& nd,
                                              typename ComputeOf<S>::type* g) {
    using T = typename ComputeOf<S>::type;
    Force<T> F = const_force(p);
    if (IBM && p.tile_flag[tile_of(p, nd.x, nd.y, nd.z)]) F = Force<T>{p.ibm_fx[idx], p.ibm_fy[idx], p.ibm_fz[idx]};
    const T om_old = (RHEO == 2) ? p.omega_state[idx] : (T)0;
    const Moments<Q, RHEO, T> m = moments<Q, RHEO, T>(g, p, om_old, F);
    if (RHEO == 2) p.omega_state[idx] = m.omega;
    // LES_EXPORT_VISC_TURBULENT: every step, as the reference (lbm.cu:237-239)
    if (RHEO == 0 && p.visc_turb) p.visc_turb[idx] = m.tau_t / (T)3;
    if (p.save) {
        p.m_rho[idx] = m.rho; p.m_ux[idx] = m.ux; p.m_uy[idx] = m.uy; p.m_uz[idx] = m.uz;
        if (RHEO != 0) p.m_omega[idx] = m.omega;
    }
    S* dst = (S*)p.dst;
    if (OUT == kRepMom) {
        T c[kCoef];
        coeffs_collide<Q, RHEO, T>(c, m, F);
#pragma unroll
        for (int k = 0; k < kCoef; k++) st_nt<S, T>(dst + (int64_t)k * p.stride + idx, c[k]);
        return;
    }
    if (p.bgk) collide_bgk<Q, RHEO, T>(g, m, F);
    else collide<Q, RHEO, T>(g, m, F);
    if constexpr (sizeof(S) == 2) {
        if (p.sr) {
            // random bits per node and step, 16 per population
            const uint32_t base = mix32(idx ^ p.sr_seed);
#pragma unroll
            for (int i = 0; i < Q; i += 2) {
                const uint32_t r = mix32(base + (uint32_t)i * 0x9e3779b9u);
                st_nt_sr<S>(dst + (int64_t)i * p.stride + idx, g[i], r);
                if (i + 1 < Q) st_nt_sr<S>(dst + (int64_t)(i + 1) * p.stride + idx, g[i + 1], r >> 16);
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < Q; i++) st_nt<S, T>(dst + (int64_t)i * p.stride + idx, g[i]);
}

// One node of one time step. DIRECT: the source holds the post-stream state
// (first step after init / load), read in place without streaming or BC.
// IBM: the force of nodes in flagged tiles comes from the IBM force field
// (src/CUDA/lbm.cu:45-58 reads macr.f per node in IBM builds). IN / OUT:
// representation of the source and destination lattices.
template <int Q, typename S, int RHEO, bool MASK, bool DIRECT, bool IBM, int IN = kRepPop, int OUT = kRepPop>
__device__ __forceinline__ void step_node(const Params<typename ComputeOf<S>::type>& p, uint32_t idx) {
    using T = typename ComputeOf<S>::type;
    const Node<T> nd = make_node(p, idx);
    const uint32_t word = node_word<MASK>(p, nd);
    if (!w_used(word)) return;
    const GlobalPC<Q, S, T, IN> pc{p, nd};
    T g[Q];
    if (DIRECT) {
        load_direct<Q, T>(g, pc);
    } else {
        gather<Q, T>(g, pc);
        if (w_scheme(word) != kNull) apply_bc<Q, T, MASK>(g, p, nd, word, pc);
    }
    collide_store<Q, S, RHEO, IBM, OUT>(p, idx, nd, g);
}

// One pulled step of a population node (gather_rows); every thread of the
// block calls it, in = false past the lattice (idx is then any valid node).
template <int Q, typename S, int RHEO, bool MASK, bool IBM, int OUT>
__device__ __forceinline__ void step_node_rows(const Params<typename ComputeOf<S>::type>& p, uint32_t idx, bool in,
                                               typename ComputeOf<S>::type* sm) {
    using T = typename ComputeOf<S>::type;
    const Node<T> nd = make_node(p, idx);
    const uint32_t word = node_word<MASK>(p, nd);
    T g[Q];
    if constexpr (!MASK && Q == 19) {
        // the x-end nodes (x = 0, NX-1) load, before the row pull, the own
        // post-collision populations a bounce-back on the W (c_x = -1 ones)
        // or E face (c_x = +1 ones) takes: in the row pull they would be a
        // second round of loads after the barrier, for 2 nodes of every row,
        // which the whole block waits for. With NX = 1 the node is both ends:
        // it prefetches the W face's values and loads the E face's in the BC.
        const GlobalPC<Q, S, T, kRepPop> pc{p, nd};
        const bool xw = nd.x == 0, xe = nd.x == p.nx - 1 && !xw;
        T v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0;
        if (xw) { v0 = pc(2, 0, 0, 0); v1 = pc(8, 0, 0, 0); v2 = pc(10, 0, 0, 0); v3 = pc(14, 0, 0, 0); v4 = pc(16, 0, 0, 0); }
        if (xe) { v0 = pc(1, 0, 0, 0); v1 = pc(7, 0, 0, 0); v2 = pc(9, 0, 0, 0); v3 = pc(13, 0, 0, 0); v4 = pc(15, 0, 0, 0); }
        gather_rows<Q, S, T>(g, p, nd, sm);
        if (!in || !w_used(word)) return;
        const int scheme = w_scheme(word), dir = w_dir(word);
        if (scheme == kBounceBack) {
            // convex edges/corners carry no bounce-back (boundaryConditionsHandler.cu:158-256)
            if (!(dir >= dNW && w_geom(word)) && dir <= dSEB) {
                const uint32_t m = c_bb19.m[dir];
#pragma unroll
                for (int i = 1; i < Q; i++) {
                    if (!(m & (1u << i))) continue;
                    const int j = opp(i), slot = xpair_slot(j);
                    const bool pre = slot >= 0 && (kCX[j] < 0 ? xw : xe);
                    const T vs = slot == 0 ? v0 : slot == 1 ? v1 : slot == 2 ? v2 : slot == 3 ? v3 : v4;
                    g[i] = pre ? vs : pc(j, 0, 0, 0);
                }
            }
        } else if (scheme != kNull) {
            apply_bc<Q, T, MASK>(g, p, nd, word, pc);
        }
    } else {
        gather_rows<Q, S, T>(g, p, nd, sm);
        if (!in || !w_used(word)) return;
        if (w_scheme(word) != kNull) apply_bc<Q, T, MASK>(g, p, nd, word, GlobalPC<Q, S, T, kRepPop>{p, nd});
    }
    collide_store<Q, S, RHEO, IBM, OUT>(p, idx, nd, g);
}

// One time step over the lattice, one node per thread; ROWS: pull through
// gather_rows (populations, not direct).
template <int Q, typename S, int RHEO, bool MASK, bool DIRECT, bool IBM = false, int IN = kRepPop, int OUT = kRepPop,
          bool ROWS = false>
__global__ __launch_bounds__(kBlock) void k_step(const Params<typename ComputeOf<S>::type> p) {
    static_assert(!ROWS || (!DIRECT && IN == kRepPop), "the row pull reads populations of the previous step");
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if constexpr (ROWS) {
        __shared__ typename ComputeOf<S>::type sm[XShifted<Q>::count() * kBlock];
        const bool in = idx < p.n;                            // the grid holds no block wholly past the lattice
        step_node_rows<Q, S, RHEO, MASK, IBM, OUT>(p, in ? idx : p.n - 1, in, sm);
    } else {
        if (idx >= p.n) return;
        step_node<Q, S, RHEO, MASK, DIRECT, IBM, IN, OUT>(p, idx);
    }
}

// Auxiliary pass over the current state (representation IN) without
// advancing it:
//   kAuxMoments:     macroscopics the next step would compute
//   kAuxMaterialize: post-stream, post-BC populations into dst (Q planes)
//   kAuxGhost:       the reference's extra z plane after a step (doubles)
// Nodes that do not collide (outside a duct wall, the rheometer's planes
// z = 0 and NZ-1) hold what the reference's push streaming leaves in them:
// the post-collision value of a colliding source node x - c_i, the extra
// plane's rest state where the stream wraps in z from a non-colliding node,
// and otherwise their initial (or loaded) value.
// The pass covers nodes [aux_base, aux_base + aux_count) (every node when
// aux_count is 0); kAuxGhost and the double output of kAuxMaterialize
// (aux_f64) write planes of aux_count values.
template <int Q, typename S, int RHEO, bool MASK, int IN = kRepPop>
__global__ __launch_bounds__(kBlock) void k_aux(const Params<typename ComputeOf<S>::type> p) {
    using T = typename ComputeOf<S>::type;
    const uint32_t plane = (uint32_t)p.nx * (uint32_t)p.ny;
    const uint32_t count = p.aux_count ? p.aux_count : (p.aux_mode == kAuxGhost ? plane : p.n);
    const uint32_t local = blockIdx.x * kBlock + threadIdx.x;
    if (local >= count) return;
    const uint32_t idx = p.aux_base + local;
    if (p.aux_mode == kAuxGhost) {
        // plane z = NZ of the reference: the populations pushed across the z
        // ends before the fold (c_z > 0 from z = NZ-1, c_z < 0 from z = 0),
        // the rest state w_i where the source does not collide or c_z = 0
        double* out = (double*)p.dst;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            double v = 0;
            if (kCZ[i] != 0) {
                const Node<T> nd = make_node(p, idx + (kCZ[i] > 0 ? (uint32_t)(p.nz - 1) * plane : 0u));
                const GlobalPC<Q, S, T, IN> pc{p, nd};
                const Node<T> src = make_node(p, node_at(p, nd, -kCX[i], -kCY[i], 0));
                if (w_used(node_word<MASK>(p, src))) v = (double)pc(i, -kCX[i], -kCY[i], 0);
            }
            out[(int64_t)i * count + local] = v;
        }
        return;
    }
    const Node<T> nd = make_node(p, idx);
    if (p.aux_mode == kAuxMomentsOutsideTiles && p.tile_flag[tile_of(p, nd.x, nd.y, nd.z)]) return;
    const uint32_t word = node_word<MASK>(p, nd);
    const GlobalPC<Q, S, T, IN> pc{p, nd};
    T g[Q];
    if (p.direct || !w_used(word)) {
        load_direct<Q, T>(g, pc);
        if (MASK && !p.direct && p.aux_mode == kAuxMaterialize) {
#pragma unroll
            for (int i = 0; i < Q; i++) {
                const Node<T> src = make_node(p, node_at(p, nd, -kCX[i], -kCY[i], -kCZ[i]));
                if (w_used(node_word<MASK>(p, src))) g[i] = pc(i, -kCX[i], -kCY[i], -kCZ[i]);
                else if ((nd.z == 0 && kCZ[i] > 0) || (nd.z == p.nz - 1 && kCZ[i] < 0)) g[i] = (T)0;
            }
        }
    } else {
        gather<Q, T>(g, pc);
        if (w_scheme(word) != kNull) apply_bc<Q, T, MASK>(g, p, nd, word, pc);
    }
    if (p.aux_mode == kAuxMaterialize) {
        if (p.aux_f64) {
            // rounded to the storage type first, as the population planes hold them
            double* out = (double*)p.dst;
#pragma unroll
            for (int i = 0; i < Q; i++) out[(int64_t)i * count + local] = (double)(T)(S)g[i];
        } else {
            S* dst = (S*)p.dst;
#pragma unroll
            for (int i = 0; i < Q; i++) dst[(int64_t)i * p.stride + idx] = (S)g[i];
        }
        return;
    }
    if (!w_used(word)) return;
    const T om_old = (RHEO == 2) ? p.omega_state[idx] : (T)0;
    const Moments<Q, RHEO, T> m = moments<Q, RHEO, T>(g, p, om_old, const_force(p));
    p.m_rho[idx] = m.rho; p.m_ux[idx] = m.ux; p.m_uy[idx] = m.uy; p.m_uz[idx] = m.uz;
    if (RHEO != 0) p.m_omega[idx] = m.omega;
}

// IBM macroscopics on the flagged tiles (gpuUpdateMacrIBM, src/CUDA/IBM/ibm.cu:544-648):
// rho and u of the post-step populations with the constant force, as the
// node's stage record; the force field of those nodes restarts from the
// constant force. Grid-stride over the tile list (its length is read on the
// device).
template <int Q, typename S, bool MASK, int IN = kRepPop>
__global__ __launch_bounds__(kBlock) void k_ibm_macr(const Params<typename ComputeOf<S>::type> p,
                                                     const uint32_t* tile_list, const uint32_t* tile_count) {
    using T = typename ComputeOf<S>::type;
    const uint32_t total = *tile_count * (uint32_t)kTileNodes;
    T* ffx = (T*)p.ibm_fx;
    T* ffy = (T*)p.ibm_fy;
    T* ffz = (T*)p.ibm_fz;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < total; e += gridDim.x * kBlock) {
        int x, y, z;
        tile_node(tile_list[e / kTileNodes], (int)(e % kTileNodes), p.tiles_x, p.tiles_y, x, y, z);
        if (x >= p.nx || y >= p.ny || z >= p.nz) continue;
        const uint32_t idx = ((uint32_t)z * (uint32_t)p.ny + (uint32_t)y) * (uint32_t)p.nx + (uint32_t)x;
        const Node<T> nd = make_node(p, idx);
        const uint32_t word = node_word<MASK>(p, nd);
        const GlobalPC<Q, S, T, IN> pc{p, nd};
        T g[Q];
        if (p.direct || !w_used(word)) {
            load_direct<Q, T>(g, pc);
        } else {
            gather<Q, T>(g, pc);
            if (w_scheme(word) != kNull) apply_bc<Q, T, MASK>(g, p, nd, word, pc);
        }
        const Moments<Q, 0, T> m = moments<Q, 0, T>(g, p, (T)0, const_force(p));
        p.ibm_stage[idx] = StageRec<T>{m.rho, m.ux, m.uy, m.uz};
        ffx[idx] = p.fx; ffy[idx] = p.fy; ffz[idx] = p.fz;
    }
}

// Equilibrium deviations from (rho, u) (gpu_f_eq, src/CUDA/globalFunctions.h:50-56),
// as populations or as coefficients (A = rho-1 - 1.5 rho u.u, B = 3 rho u,
// G_aa = 4.5 rho u_a^2, G_ab = 9 rho u_a u_b).
template <int Q, typename S, int OUT = kRepPop>
__global__ __launch_bounds__(kBlock) void k_fill_feq(S* state, int64_t stride, uint32_t n, const double* rho,
                                                     const double* ux, const double* uy, const double* uz) {
    using T = typename ComputeOf<S>::type;
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n) return;
    const T r = (T)rho[idx], vx = (T)ux[idx], vy = (T)uy[idx], vz = (T)uz[idx];
    const T usq = vx * vx + vy * vy + vz * vz;
    if (OUT == kRepMom) {
        const T c[kCoef] = {(r - (T)1) - (T)1.5 * r * usq, (T)3 * r * vx, (T)3 * r * vy, (T)3 * r * vz,
                            (T)4.5 * r * vx * vx, (T)4.5 * r * vy * vy, (T)4.5 * r * vz * vz,
                            (T)9 * r * vx * vy, (T)9 * r * vx * vz, (T)9 * r * vy * vz};
#pragma unroll
        for (int k = 0; k < kCoef; k++) state[(int64_t)k * stride + idx] = (S)c[k];
        return;
    }
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const T cu = (T)kCX[i] * vx + (T)kCY[i] * vy + (T)kCZ[i] * vz;
        const T w = (T)weight<Q>(i);
        // w (rho (1 - 1.5 u.u + 3 c.u + 4.5 (c.u)^2) - 1)
        const T dev = w * ((r - (T)1) + r * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq));
        state[(int64_t)i * stride + idx] = (S)dev;
    }
}

// Residual / mean-density partial sums with FP64 accumulation: per block
// {sum |u - u_old|, sum |u|, sum rho}; u_old <- u.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce_partial(const T* rho, const T* ux, const T* uy, const T* uz,
                                                           T* oux, T* ouy, T* ouz, uint32_t n, double* partials) {
    __shared__ double sh[3][kBlock / 64];
    double a = 0, b = 0, c = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double vx = ux[i], vy = uy[i], vz = uz[i];
        const double dx = vx - (double)oux[i], dy = vy - (double)ouy[i], dz = vz - (double)ouz[i];
        a += sqrt(dx * dx + dy * dy + dz * dz);
        b += sqrt(vx * vx + vy * vy + vz * vz);
        c += (double)rho[i];
        oux[i] = ux[i]; ouy[i] = uy[i]; ouz[i] = uz[i];
    }
    // wave64 tree
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
        for (int w = 0; w < kBlock / 64; w++) { s0 += sh[0][w]; s1 += sh[1][w]; s2 += sh[2][w]; }
        partials[3 * blockIdx.x + 0] = s0;
        partials[3 * blockIdx.x + 1] = s1;
        partials[3 * blockIdx.x + 2] = s2;
    }
}

// Element-wise conversion between a storage-typed buffer and doubles.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_convert(S* buf, double* dbl, size_t n, int to_double) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (to_double) dbl[i] = (double)buf[i];
    else buf[i] = (S)dbl[i];
}

}  // namespace vlbm
