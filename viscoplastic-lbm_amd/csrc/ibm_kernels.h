// HIP kernels of the immersed boundary stage (rigid spheres), gfx950.
//
// One stage per time step, after the LBM step (src/CUDA/main.cu:369-389):
//   begin     old values, cell index of each sphere, node forces reset; the
//             16x2x2 tiles touched by the interpolation stencils are appended to
//             a tile list (first touch wins an atomic flag)
//   macr      rho, u of the post-step populations with the constant force on
//             the listed tiles only; force field reset to the constant there
//   contacts  uniform cell list (one-workgroup bitonic sort of (cell, sphere)
//             in LDS), Hertz-Mindlin sphere-sphere contacts by the lower-index
//             sphere and force / torque sums per sphere in a fixed order, one
//             wave per sphere; walls; velocity update; the spreading plan
//             (overlapping boxes, stray nodes)
//   iteration (x max_iterations) chunks of 64 nodes of one sphere:
//             interpolation and force of each node, FP64 sums of the node
//             forces / moments (fixed-order tree); sphere sums and velocity
//             update; spreading of the node forces straight into u (F / 2 rho)
//             and the force field, through an LDS box around each sphere for
//             FP32 fluids (one workgroup per sphere)
//   move      sphere and node movement
// Everything outside the flagged tiles is left alone: the reference's full-
// lattice passes (gpuUpdateMacrIBM, gpuEulerSumIBMAuxsReset over the bordered
// lattice) become passes over the tiles the spheres touch.
#pragma once
#include <hip/hip_runtime.h>
#include "ibm.h"
#include "lattice.h"

namespace vlbm {

// Peskin's 4-point kernel (src/CUDA/IBM/ibmGlobalFunctions.h:28-48).
__device__ __forceinline__ double peskin4(double x) {
    const double a = fabs(x);
    if (a <= 1) return (1.0 / 8.0) * (3.0 - 2.0 * a + sqrt(1 + 4 * a - 4 * a * a));
    if (a > 1.0 && a <= 2.0) return (1.0 / 8.0) * (5.0 - 2.0 * a - sqrt(-7.0 + 12.0 * a - 4.0 * a * a));
    return 0.0;
}

// vector_length / vector_normalize (src/CUDA/globalFunctions.cu:81-100): the
// reference evaluates them with sqrtf / rsqrtf also in FP64 builds.
__device__ __forceinline__ double vec_len(const IbmParams& P, const double v[3]) {
    const double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (P.ref_float) return (double)sqrtf((float)s);
    return sqrt(s);
}
__device__ __forceinline__ void vec_norm(const IbmParams& P, const double v[3], double o[3]) {
    const double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double inv = P.ref_float ? (double)(1.0f / sqrtf((float)s)) : 1.0 / sqrt(s);
    if (isnan(inv) || isinf(inv)) { o[0] = o[1] = o[2] = 0; return; }
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

__device__ __forceinline__ int axis_n(const IbmParams& P, int a) { return a == 0 ? P.nx : (a == 1 ? P.ny : P.nz); }

__device__ __forceinline__ double wrap_pos(const IbmParams& P, int a, double v) {
    const double n = (double)axis_n(P, a);
    return fmod(v + n, n);
}

// ---------------------------------------------------------------------------
// stencil of a Lagrangian node (gpuForceInterpolationSpread, ibm.cu:194-330):
// base int(x) - 1, four nodes per axis, clipped at walls, wrapped if periodic.
// ---------------------------------------------------------------------------
struct Stencil {
    int base[3], mn[3], mx[3];
    bool valid;
};

__device__ __forceinline__ Stencil make_stencil(const IbmParams& P, const double pos[3]) {
    Stencil s;
    for (int a = 0; a < 3; a++) {
        const double pa = (pos[a] > -8.0 && pos[a] < (double)axis_n(P, a) + 8.0) ? pos[a] : -8.0;
        s.base[a] = (int)pa - 2 + 1;
        const int n = axis_n(P, a);
        if (P.periodic[a]) {
            s.mn[a] = 0;
            s.mx[a] = 3;
        } else {
            s.mx[a] = ((s.base[a] + 3) < n) ? 3 : (n - 1 - s.base[a]);
            s.mn[a] = (s.base[a] >= 0) ? 0 : -s.base[a];
        }
    }
    s.valid = !(s.mx[0] <= 0 || s.mx[1] <= 0 || s.mx[2] <= 0 || s.mn[0] >= 4 || s.mn[1] >= 4 || s.mn[2] >= 4);
    // nodes far outside the lattice (or not finite) touch nothing; this also
    // keeps the integer stencil arithmetic in range
    for (int a = 0; a < 3; a++)
        if (!(pos[a] > -4.0 && pos[a] < (double)axis_n(P, a) + 4.0)) s.valid = false;
    return s;
}

// v mod n in [0, n). Stencil and box coordinates lie within one period of
// the range (valid stencils start in [-4, n + 3]), so one add or subtract
// wraps them; values farther out (axes shorter than the stencil's reach)
// take the integer modulo.
__device__ __forceinline__ int wrap_axis(int v, int n) {
    v = v < 0 ? v + n : (v >= n ? v - n : v);
    if ((unsigned)v >= (unsigned)n) v = ((v % n) + n) % n;
    return v;
}

__device__ __forceinline__ int stencil_coord(const IbmParams& P, const Stencil& s, int a, int j) {
    const int c = s.base[a] + j;
    return P.periodic[a] ? wrap_axis(c, axis_n(P, a)) : c;
}

// ---------------------------------------------------------------------------
// begin: old values (gpuUpdateParticleOldValues, ibm.cu:995-1042), cell index,
// tile flags cleared
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIbmBlock) void k_ibm_begin(IbmParams P) {
    const uint32_t gid = blockIdx.x * kIbmBlock + threadIdx.x;
    const uint32_t stride = gridDim.x * kIbmBlock;
    for (uint32_t t = gid; t < P.n_tiles; t += stride) P.tile_flag[t] = 0;
    if (gid == 0) { *P.tile_count = 0; *P.box_stray = 0; }
    if (gid >= (uint32_t)P.n_spheres) return;
    IbmSphere& c = P.sph[gid];
    int cc[3];
    for (int a = 0; a < 3; a++) {
        c.dP[a] = 0; c.dL[a] = 0;
        c.pos_old[a] = c.pos[a]; c.vel_old[a] = c.vel[a]; c.w_old[a] = c.w[a];
        c.f_old[a] = c.f[a]; c.f[a] = 0; c.M[a] = 0;
        int k = (int)floor(c.pos[a] / P.cell_w[a]);
        k = k < 0 ? 0 : (k >= P.cells[a] ? P.cells[a] - 1 : k);
        cc[a] = k;
    }
    c.cell = cc[0] + P.cells[0] * (cc[1] + P.cells[1] * cc[2]);
}

// node forces reset (gpuResetNodesForces, ibm.cu:650-664) and tiles flagged
__global__ __launch_bounds__(kIbmBlock) void k_ibm_mark(IbmParams P) {
    const uint32_t i = blockIdx.x * kIbmBlock + threadIdx.x;
    if (i >= (uint32_t)P.n_nodes) return;
    P.pfx[i] = 0; P.pfy[i] = 0; P.pfz[i] = 0;
    P.pdx[i] = 0; P.pdy[i] = 0; P.pdz[i] = 0;
    P.pmx[i] = 0; P.pmy[i] = 0; P.pmz[i] = 0;
    const double pos[3] = {P.px[i], P.py[i], P.pz[i]};
    const Stencil s = make_stencil(P, pos);
    if (!s.valid) return;
    int tl[3][4], nt[3];
    for (int a = 0; a < 3; a++) {
        nt[a] = 0;
        for (int j = s.mn[a]; j <= s.mx[a]; j++) {
            const int t = stencil_coord(P, s, a, j) >> tile_shift(a);
            bool seen = false;
            for (int k = 0; k < nt[a]; k++) seen |= (tl[a][k] == t);
            if (!seen) tl[a][nt[a]++] = t;
        }
    }
    for (int c = 0; c < nt[2]; c++)
        for (int b = 0; b < nt[1]; b++)
            for (int a = 0; a < nt[0]; a++)
                P.tile_flag[tl[0][a] + P.tiles[0] * (tl[1][b] + P.tiles[1] * tl[2][c])] = 1u;
}

// flagged tiles -> list: a block-wide scan of 1024 flags, one counter atomic
// per block (list order differs from run to run; every use of the list is
// per node, so results do not)
__global__ __launch_bounds__(1024) void k_ibm_list_tiles(IbmParams P) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t base;
    const uint32_t t = blockIdx.x * 1024 + threadIdx.x;
    const bool on = t < P.n_tiles && P.tile_flag[t] != 0u;
    const uint64_t mask = __ballot(on);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < 16; w++) { const uint32_t v = wsum[w]; wsum[w] = tot; tot += v; }
        base = tot ? atomicAdd(P.tile_count, tot) : 0u;
    }
    __syncthreads();
    if (on) P.tile_list[base + wsum[wave] + before] = t;
}

// ---------------------------------------------------------------------------
// cell list
// ---------------------------------------------------------------------------
// The distinct cells whose spheres can touch a sphere of the given cell, as
// one index range per axis: axis a visits lo[a] + k (mod the cell count when
// periodic) for k < cnt[a]. Kept as ranges, not a list: an indexed list of
// 27 cells would live in private (scratch) memory, each lane of a contact
// wave computes its one cell from the ranges.
struct CellRange {
    int lo[3], cnt[3];
};

__device__ __forceinline__ CellRange neighbour_range(const IbmParams& P, int cell) {
    const int cc[3] = {cell % P.cells[0], (cell / P.cells[0]) % P.cells[1], cell / (P.cells[0] * P.cells[1])};
    CellRange r;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int nc = P.cells[a];
        if (P.periodic[a] && nc <= 3) {
            r.lo[a] = 0;
            r.cnt[a] = nc;
        } else if (P.periodic[a]) {
            r.lo[a] = cc[a] - 1;
            r.cnt[a] = 3;
        } else {
            const int lo = cc[a] > 0 ? cc[a] - 1 : 0, hi = cc[a] + 1 < nc ? cc[a] + 1 : nc - 1;
            r.lo[a] = lo;
            r.cnt[a] = hi - lo + 1;
        }
    }
    return r;
}

// cell (k0, k1, k2) of a range, k_a < cnt[a]
__device__ __forceinline__ uint64_t range_cell(const IbmParams& P, const CellRange& r, int k0, int k1, int k2) {
    int k[3] = {r.lo[0] + k0, r.lo[1] + k1, r.lo[2] + k2};
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (P.periodic[a]) k[a] = (k[a] + P.cells[a]) % P.cells[a];
    return (uint64_t)(uint32_t)(k[0] + P.cells[0] * (k[1] + P.cells[1] * k[2]));
}

__device__ __forceinline__ int lower_bound_key(const uint64_t* keys, int n, uint64_t k) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// contacts (src/CUDA/IBM/collision/ibmCollision.cu)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wall_slot(const double n[3]) {   // calculateWallIndex
    return 7 + (1 - (int)n[0]) - 2 * (1 - (int)n[1]) - 3 * (1 - (int)n[2]);
}
__device__ __forceinline__ void end_collision(IbmSphere& c, int idx, int step) {
    if (idx >= 0 && idx <= 6) {
        c.last[idx] = -1;
    } else if (idx >= kFirstPartnerSlot && idx < kMaxColl) {
        if (step - c.last[idx] > 1) {
            c.tang[idx][0] = c.tang[idx][1] = c.tang[idx][2] = 0;
            c.last[idx] = -1;
            c.partner[idx] = -1;
        }
    }
}
__device__ __forceinline__ int start_collision(IbmSphere& c, int partner, bool wall, const double n[3], int step) {
    if (wall) {
        const int idx = wall_slot(n);
        c.tang[idx][0] = c.tang[idx][1] = c.tang[idx][2] = 0;
        c.last[idx] = step;
        return idx;
    }
    for (int i = kFirstPartnerSlot; i < kMaxColl; i++)
        if (c.last[i] == -1) {
            c.partner[i] = partner;
            c.tang[i][0] = c.tang[i][1] = c.tang[i][2] = 0;
            c.last[i] = step;
            return i;
        }
    return -1;
}
__device__ __forceinline__ void update_tang(IbmSphere& c, int idx, const double d[3], int step, double out[3]) {
    for (int a = 0; a < 3; a++) { c.tang[idx][a] += d[a]; out[a] = c.tang[idx][a]; }
    c.last[idx] = step;
}
__device__ __forceinline__ int slot_of_partner(const IbmSphere& c, int partner, int step) {
    for (int i = kFirstPartnerSlot; i < kMaxColl; i++)
        if (c.partner[i] == partner && step - c.last[i] <= 1) return i;
    return -1;
}

// sphereSphereGap (ibmCollision.cu:1061-1088)
__device__ __forceinline__ double sphere_gap(const IbmParams& P, const IbmSphere& a, const IbmSphere& b) {
    double dl[3];
    for (int k = 0; k < 3; k++) {
        dl[k] = a.pos[k] - b.pos[k];
        if (P.periodic[k]) {
            const double L = axis_n(P, k);
            if (dl[k] > L / 2.0) dl[k] -= L;
            if (dl[k] < -L / 2.0) dl[k] += L;
        }
    }
    const double s2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
    const double dist = P.ref_float ? (double)sqrtf((float)s2) : sqrt(s2);
    return dist - (a.radius + b.radius);
}

// sphereSphereCollision (ibmCollision.cu:1719-1913) seen from sphere i (the
// lower index, whose history is updated). Returns false without contact.
__device__ bool pair_contact(const IbmParams& P, IbmSphere& ci, const IbmSphere& cj, int row, int step,
                             IbmContact& out) {
    const double r_i = ci.radius, r_j = cj.radius;
    const double m_i = ci.volume * ci.density, m_j = cj.volume * cj.density;
    double dp[3];
    for (int a = 0; a < 3; a++) {
        double d = ci.pos[a] - cj.pos[a];
        if (P.periodic[a]) {
            const double L = axis_n(P, a);
            if (fabs(ci.pos[a] - cj.pos[a]) > L / 2.0) d = ci.pos[a] < cj.pos[a] ? (ci.pos[a] + L - cj.pos[a]) : (ci.pos[a] - L - cj.pos[a]);
        }
        dp[a] = d;
    }
    const double md = sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
    if (md > r_i + r_j) return false;
    const double n[3] = {dp[0] / md, dp[1] / md, dp[2] / md};
    const double disp = r_i + r_j - md;
    const double G[3] = {ci.vel[0] - cj.vel[0], ci.vel[1] - cj.vel[1], ci.vel[2] - cj.vel[2]};
    const double er = 1.0 / ((r_i + r_j) / (r_i * r_j));
    const double em = 1.0 / ((m_i + m_j) / (m_i * m_j));
    const double kn = P.pp_kn * sqrt(er);
    const double kt = P.pp_kt * sqrt(er) * sqrt(fabs(disp));
    const double dn = P.pp_damp * sqrt(em * kn), dt = P.pp_damp * sqrt(em * kt);
    const double f_kn = -kn * sqrt(fabs(disp * disp * disp));
    const double Gn = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
    const double p25 = pow(fabs(disp), 0.25);
    double fn[3];
    for (int a = 0; a < 3; a++) fn[a] = f_kn * n[a] - dn * Gn * n[a] * p25;
    const double f_n = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
    const double* wi = ci.w;
    const double* wj = cj.w;
    double Gct[3];
    Gct[0] = G[0] + r_i * (wi[1] * n[2] - wi[2] * n[1]) + r_j * (wj[1] * n[2] - wj[2] * n[1]) - Gn * n[0];
    Gct[1] = G[1] + r_i * (wi[2] * n[0] - wi[0] * n[2]) + r_j * (wj[2] * n[0] - wj[0] * n[2]) - Gn * n[1];
    Gct[2] = G[2] + r_i * (wi[0] * n[1] - wi[1] * n[0]) + r_j * (wj[0] * n[1] - wj[1] * n[0]) - Gn * n[2];
    double mag = sqrt(Gct[0] * Gct[0] + Gct[1] * Gct[1] + Gct[2] * Gct[2]);
    double t[3] = {0, 0, 0};
    if (mag != 0)
        for (int a = 0; a < 3; a++) t[a] = Gct[a] / mag;
    double td[3];
    int ti = slot_of_partner(ci, row, step);
    if (ti == -1) {
        ti = start_collision(ci, row, false, n, step);
        for (int a = 0; a < 3; a++) td[a] = Gct[a];
    } else if (step - ci.last[ti] > 1) {
        end_collision(ci, ti, step);
        ti = start_collision(ci, row, false, n, step);
        for (int a = 0; a < 3; a++) td[a] = Gct[a];
    } else {
        update_tang(ci, ti, G, step, td);   // the reference accumulates G here
    }
    double ft[3];
    for (int a = 0; a < 3; a++) ft[a] = -kt * td[a] - dt * Gct[a] * pow(fabs(td[a]), 0.25);
    mag = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
    if (mag > P.pp_mu * fabs(f_n)) {
        const double mg[3] = {-Gct[0], -Gct[1], -Gct[2]};
        if (ti >= 0) update_tang(ci, ti, mg, step, td);   // no free slot: the reference writes slot -1
        for (int a = 0; a < 3; a++) ft[a] = -P.pp_mu * f_n * t[a];
    }
    out.j = row;
    out.pad = 0;
    for (int a = 0; a < 3; a++) out.fd[a] = fn[a] + ft[a];
    out.mi[0] = r_i * (n[1] * ft[2] - n[2] * ft[1]);
    out.mi[1] = r_i * (n[2] * ft[0] - n[0] * ft[2]);
    out.mi[2] = r_i * (n[0] * ft[1] - n[1] * ft[0]);
    out.mj[0] = r_j * (n[1] * ft[2] - n[2] * ft[1]);
    out.mj[1] = r_j * (n[2] * ft[0] - n[0] * ft[2]);
    out.mj[2] = r_j * (n[0] * ft[1] - n[1] * ft[0]);
    return true;
}

// sphereWallCollision (ibmCollision.cu:394-508), accumulating into f / M
__device__ void wall_contact(const IbmParams& P, IbmSphere& c, const double wall_normal[3], double disp, int step,
                             double f[3], double M[3]) {
    const double m_i = c.volume * c.density, r_i = c.radius;
    const double n[3] = {-wall_normal[0], -wall_normal[1], -wall_normal[2]};
    const double* G = c.vel;
    const double* w_i = c.w;
    const double kn = P.pw_kn * sqrt(fabs(r_i));
    const double kt = P.pw_kt * sqrt(r_i) * sqrt(fabs(disp));
    const double dn = P.pw_damp * sqrt(m_i * kn), dt = P.pw_damp * sqrt(m_i * kt);
    const double f_kn = -kn * sqrt(fabs(disp * disp * disp));
    const double Gn = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
    const double p25 = pow(fabs(disp), 0.25);
    double fn[3];
    for (int a = 0; a < 3; a++) fn[a] = f_kn * n[a] - dn * Gn * n[a] * p25;
    const double f_n = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
    double Gct[3];
    Gct[0] = G[0] + r_i * (w_i[1] * n[2] - w_i[2] * n[1]) - Gn * n[0];
    Gct[1] = G[1] + r_i * (w_i[2] * n[0] - w_i[0] * n[2]) - Gn * n[1];
    Gct[2] = G[2] + r_i * (w_i[0] * n[1] - w_i[1] * n[0]) - Gn * n[2];
    double mag = sqrt(Gct[0] * Gct[0] + Gct[1] * Gct[1] + Gct[2] * Gct[2]);
    double t[3] = {0, 0, 0};
    if (mag != 0)
        for (int a = 0; a < 3; a++) t[a] = Gct[a] / mag;
    const int ti = wall_slot(n);
    double td[3];
    if (step - c.last[ti] > 1) {
        end_collision(c, ti, step);
        start_collision(c, ti, true, n, step);
        for (int a = 0; a < 3; a++) td[a] = Gct[a];
    } else {
        update_tang(c, ti, Gct, step, td);
    }
    double ft[3];
    for (int a = 0; a < 3; a++) ft[a] = -kt * td[a] - dt * Gct[a] * pow(fabs(td[a]), 0.25);
    mag = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
    const double fnab = P.ref_float ? (double)fabsf((float)f_n) : fabs(f_n);
    if (mag > P.pw_mu * fnab) {
        const double mg[3] = {-Gct[0], -Gct[1], -Gct[2]};
        update_tang(c, ti, mg, step, td);
        for (int a = 0; a < 3; a++) ft[a] = -P.pw_mu * f_n * t[a];
    }
    for (int a = 0; a < 3; a++) f[a] += fn[a] + ft[a];
    M[0] += r_i * (n[1] * ft[2] - n[2] * ft[1]);
    M[1] += r_i * (n[2] * ft[0] - n[0] * ft[2]);
    M[2] += r_i * (n[0] * ft[1] - n[1] * ft[0]);
}

// checkCollisionWallsSphere (ibmCollision.cu:188-251): walls at 0 and N-1
__device__ void wall_contacts(const IbmParams& P, IbmSphere& c, double f[3], double M[3]) {
    for (int a = 0; a < 3; a++) {
        if (P.periodic[a]) continue;
        double nrm[3] = {0, 0, 0};
        nrm[a] = 1;
        double dist = c.pos[a] * 1.0 - 0.0;
        if (dist < c.radius) wall_contact(P, c, nrm, c.radius - dist, P.step, f, M);
        nrm[a] = -1;
        dist = (double)(axis_n(P, a) - 1) + (-c.pos[a]);
        if (dist < c.radius) wall_contact(P, c, nrm, c.radius - dist, P.step, f, M);
    }
}

// Origin of a sphere's spreading box on axis a (the box is box_edge nodes
// wide, centred on the sphere).
__device__ __forceinline__ int box_origin(const IbmParams& P, const IbmSphere& c, int a) {
    return (int)floor(c.pos[a]) - (P.box_edge / 2 - 3) - 2;
}

// Whether lattice coordinate q (unwrapped, axis a) lies in the box of origin o.
__device__ __forceinline__ bool in_box_axis(const IbmParams& P, int a, int q, int o) {
    int d = q - o;
    if (P.periodic[a]) d = wrap_axis(d, axis_n(P, a));
    return d >= 0 && d < P.box_edge;
}

// Spreading plan of sphere s (by one wave; positions are fixed until the
// stage moves the spheres): the spheres whose boxes overlap this one's
// (at most kBoxNb listed, else the count is -1), and box_stray set when a
// node lies farther than box_reach from its centre on some axis. Nodes within
// box_reach put every stencil point in their own sphere's box, so with no
// stray node a box node outside every other box receives values from its own
// sphere's workgroup only.
__device__ void box_plan(const IbmParams& P, int s, int lane) {
    const IbmSphere& c = P.sph[s];
    bool out = false;
    for (int k = lane; k < c.n_nodes; k += 64) {
        const int i = c.first_node + k;
        const double p[3] = {P.px[i], P.py[i], P.pz[i]};
        for (int a = 0; a < 3; a++) {
            double d = p[a] - c.pos[a];
            if (P.periodic[a]) {
                const double n = (double)axis_n(P, a);
                if (d > n / 2) d -= n;
                if (d < -n / 2) d += n;
            }
            if (!(fabs(d) <= P.box_reach)) out = true;
        }
    }
    if (__ballot(out) && lane == 0) atomicOr(P.box_stray, 1);
    const int o[3] = {box_origin(P, c, 0), box_origin(P, c, 1), box_origin(P, c, 2)};
    int* list = P.box_nb + s * (kBoxNb + 1);
    int cnt = 0;
    for (int j0 = 0; j0 < P.n_spheres; j0 += 64) {
        const int j = j0 + lane;
        bool ov = false;
        if (j < P.n_spheres && j != s) {
            const IbmSphere& cj = P.sph[j];
            ov = true;
            for (int a = 0; a < 3; a++) {
                const int oj = box_origin(P, cj, a);
                // two boxes of one width intersect when either origin lies in the other box
                if (!in_box_axis(P, a, oj, o[a]) && !in_box_axis(P, a, o[a], oj)) ov = false;
            }
        }
        const uint64_t m = __ballot(ov);
        const int at = cnt + __popcll(m & ((1ull << lane) - 1ull));
        if (ov && at < kBoxNb) list[1 + at] = j;
        cnt += __popcll(m);
    }
    if (lane == 0) list[0] = cnt > kBoxNb ? -1 : cnt;
}

__device__ void sphere_velocity(const IbmParams& P, int s);

// The contact kernels run one wave per sphere. Each lane of the wave takes
// one cell of the sphere's neighbour range (at most 27) and walks that cell's
// run of the sorted keys; what the lanes find is handed to lane 0 by
// shuffles, in rounds the whole wave takes together, and lane 0 keeps it in a
// list sorted by sphere index in LDS. Lane 0 then does the order-dependent
// work (history updates, sums) exactly as one thread per sphere would.
constexpr int kContactWaves = 4;                  // waves (spheres) per workgroup
constexpr int kContactBlock = 64 * kContactWaves;

// The sorted keys into LDS, by every thread of the workgroup.
__device__ __forceinline__ void load_cell_keys(const IbmParams& P, uint64_t* sk) {
    for (int k = threadIdx.x; k < P.n_spheres; k += kContactBlock) sk[k] = P.cell_keys[k];
    __syncthreads();
}

// This lane's run of keys [k, kend): cell `lane` of the range around `cell`.
__device__ __forceinline__ void lane_cell_run(const IbmParams& P, const uint64_t* sk, int cell, int lane, int& k,
                                              int& kend) {
    const CellRange nb = neighbour_range(P, cell);
    k = kend = 0;
    if (lane < nb.cnt[0] * nb.cnt[1] * nb.cnt[2]) {
        const uint64_t c = range_cell(P, nb, lane % nb.cnt[0], (lane / nb.cnt[0]) % nb.cnt[1],
                                      lane / (nb.cnt[0] * nb.cnt[1]));
        k = lower_bound_key(sk, P.n_spheres, c << 32);
        kend = lower_bound_key(sk, P.n_spheres, (c + 1) << 32);
    }
}

// Inserts v (and w beside it, when lw is given) into the list of n values
// kept ascending, at most kMaxContacts long: when full, the largest value is
// dropped and `over` set.
__device__ __forceinline__ void insert_sorted(int* lv, int* lw, int& n, int v, int w, bool& over) {
    if (n == kMaxContacts) {
        over = true;
        if (v > lv[n - 1]) return;
        n--;
    }
    int at = n++;
    while (at > 0 && lv[at - 1] > v) {
        lv[at] = lv[at - 1];
        if (lw) lw[at] = lw[at - 1];
        at--;
    }
    lv[at] = v;
    if (lw) lw[at] = w;
}

// Cell sort: one workgroup, bitonic sort of the (cell, sphere) keys in LDS,
// written to cell_keys.
__global__ __launch_bounds__(1024) void k_ibm_cell_sort(IbmParams P) {
    __shared__ uint64_t key[kMaxSortKeys];
    const int n = P.n_spheres;
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += 1024)
        key[i] = i < n ? (((uint64_t)(uint32_t)P.sph[i].cell << 32) | (uint32_t)i) : ~0ull;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += 1024) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool up = (i & k) == 0;
                    const uint64_t a = key[i], b = key[ixj];
                    if ((a > b) == up) { key[i] = b; key[ixj] = a; }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n; i += 1024) P.cell_keys[i] = key[i];
}

// Sphere-sphere contacts of sphere i with the higher-index spheres of the
// neighbouring cells, in ascending partner order (the reference's pair order
// for that sphere's history), and the sphere's spreading plan; one wave per
// sphere.
__global__ __launch_bounds__(kContactBlock) void k_ibm_pairs(IbmParams P) {
    __shared__ uint64_t sk[kMaxSortKeys];
    __shared__ int cand[kContactWaves][kMaxContacts];
    load_cell_keys(P, sk);
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int i = (int)blockIdx.x * kContactWaves + wave;
    if (i >= P.n_spheres) return;
    if (P.box_edge > 0 && P.box_plain) box_plan(P, i, lane);
    IbmSphere& ci = P.sph[i];
    int k, kend;
    lane_cell_run(P, sk, ci.cell, lane, k, kend);
    int nc = 0;
    bool over = false;
    while (__ballot(k < kend)) {
        int j = -1;
        if (k < kend) {
            const int jj = (int)(sk[k++] & 0xffffffffu);
            if (jj > i) {
                const IbmSphere& cj = P.sph[jj];
                if ((ci.movable || cj.movable) && sphere_gap(P, ci, cj) < 0) j = jj;
            }
        }
        for (uint64_t m = __ballot(j >= 0); m; m &= m - 1) {
            const int jb = __shfl(j, (int)__builtin_ctzll(m), 64);
            if (lane == 0) insert_sorted(cand[wave], nullptr, nc, jb, 0, over);
        }
    }
    if (lane != 0) return;
    if (over) atomicOr(P.error, (int)kIbmErrContacts);
    int rec = 0;
    for (int c = 0; c < nc; c++) {
        const int j = cand[wave][c];
        IbmContact ct;
        if (pair_contact(P, ci, P.sph[j], j, P.step, ct)) P.contacts[i * kMaxContacts + rec++] = ct;
    }
    P.n_contacts[i] = rec;
}

// Force / torque of the contacts in the reference's accumulation order for
// sphere p: as the higher index (ascending partner), as the lower index
// (ascending partner), then the walls; then the sphere's velocity update.
// One wave per sphere.
__global__ __launch_bounds__(kContactBlock) void k_ibm_contact_sums(IbmParams P) {
    __shared__ uint64_t sk[kMaxSortKeys];
    __shared__ int from_i[kContactWaves][kMaxContacts], from_k[kContactWaves][kMaxContacts];
    load_cell_keys(P, sk);
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int p = (int)blockIdx.x * kContactWaves + wave;
    if (p >= P.n_spheres) return;
    IbmSphere& cp = P.sph[p];
    int k, kend;
    lane_cell_run(P, sk, cp.cell, lane, k, kend);
    int nr = 0;
    bool over = false;
    while (__ballot(k < kend)) {
        int fi = -1, fk = -1;
        if (k < kend) {
            const int i = (int)(sk[k++] & 0xffffffffu);
            if (i < p) {
                const int nci = P.n_contacts[i];
                for (int r = 0; r < nci; r++)
                    if (P.contacts[i * kMaxContacts + r].j == p) { fi = i; fk = r; break; }
            }
        }
        for (uint64_t m = __ballot(fi >= 0); m; m &= m - 1) {
            const int b = (int)__builtin_ctzll(m);
            const int ib = __shfl(fi, b, 64), kb = __shfl(fk, b, 64);
            if (lane == 0) insert_sorted(from_i[wave], from_k[wave], nr, ib, kb, over);
        }
    }
    if (lane != 0) return;
    if (over) atomicOr(P.error, (int)kIbmErrPartners);
    double f[3] = {cp.f[0], cp.f[1], cp.f[2]};
    double M[3] = {cp.M[0], cp.M[1], cp.M[2]};
    for (int r = 0; r < nr; r++) {
        const IbmContact& ct = P.contacts[from_i[wave][r] * kMaxContacts + from_k[wave][r]];
        for (int a = 0; a < 3; a++) { f[a] += ct.fd[a]; M[a] += ct.mj[a]; }
    }
    const int own = P.n_contacts[p];
    for (int r = 0; r < own; r++) {
        const IbmContact& ct = P.contacts[p * kMaxContacts + r];
        for (int a = 0; a < 3; a++) { f[a] += -ct.fd[a]; M[a] += ct.mi[a]; }
    }
    if (cp.movable) wall_contacts(P, cp, f, M);
    for (int a = 0; a < 3; a++) { cp.f[a] = f[a]; cp.M[a] = M[a]; }
    sphere_velocity(P, p);
}

// gpuUpdateParticleCenterVelocityAndRotation (ibm.cu:755-897)
__device__ void sphere_velocity(const IbmParams& P, int s) {
    IbmSphere& c = P.sph[s];
    if (!c.movable) return;
    const double al = P.alpha;
    const double inv_volume = 1 / c.volume;
    for (int a = 0; a < 3; a++)
        c.vel[a] = c.vel_old[a] + (((c.f_old[a] * (1.0 - al) + c.f[a] * al) + c.dP[a]) * inv_volume +
                                   (c.density - P.fluid_density) * P.g[a]) / (c.density);
    const double xx = c.I[0], yy = c.I[1], zz = c.I[2], xy = c.I[3], xz = c.I[4], yz = c.I[5];
    const double inv_det_neg = 1.0 / (zz * xy * xy + yy * xz * xz + xx * yz * yz - xx * yy * zz - 2 * xy * xz * yz);
    double wA[3], LM[3];
    for (int a = 0; a < 3; a++) {
        wA[a] = (c.w_old[a] * (1.0 - al) + c.w_old[a] * al);
        LM[a] = c.dL[a] + (c.M[a] * al + c.M_old[a] * (1.0 - al));
    }
    const double Iwx = xx * wA[0] + xy * wA[1] + xz * wA[2];
    const double Iwy = xy * wA[0] + yy * wA[1] + yz * wA[2];
    const double Iwz = xz * wA[0] + yz * wA[1] + zz * wA[2];
    const double Ax = LM[0] + wA[2] * Iwy - wA[1] * Iwz;
    const double Ay = LM[1] + wA[0] * Iwz - wA[2] * Iwx;
    const double Az = LM[2] + wA[1] * Iwx - wA[0] * Iwy;
    double wN[3];
    wN[0] = c.w_old[0] + ((yz * yz - yy * zz) * Ax - (xy * yz - xz * yy) * Az - (xz * yz - xy * zz) * Ay) * inv_det_neg;
    wN[1] = c.w_old[1] + ((xz * xz - xx * zz) * Ay - (xy * xz - xx * yz) * Az - (xz * yz - xy * zz) * Ax) * inv_det_neg;
    wN[2] = c.w_old[2] + ((xy * xy - xx * yy) * Az - (xy * xz - xx * yz) * Ay - (xy * yz - xz * yy) * Ax) * inv_det_neg;
    // inertia rotated by the average rotation vector: R I R^T
    const double ang = vec_len(P, wA);
    double ax[3];
    vec_norm(P, wA, ax);
    const double h = ang / 2.0;
    const double qw = cos(h), qx = ax[0] * sin(h), qy = ax[1] * sin(h), qz = ax[2] * sin(h);
    double R[3][3];
    R[0][0] = 1 - 2 * (qy * qy + qz * qz); R[0][1] = 2 * (qx * qy - qw * qz); R[0][2] = 2 * (qx * qz + qw * qy);
    R[1][0] = 2 * (qx * qy + qw * qz); R[1][1] = 1 - 2 * (qx * qx + qz * qz); R[1][2] = 2 * (qy * qz - qw * qx);
    R[2][0] = 2 * (qx * qz - qw * qy); R[2][1] = 2 * (qy * qz + qw * qx); R[2][2] = 1 - 2 * (qx * qx + qy * qy);
    const double M3[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
    double Tm[3][3], N3[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += R[i][k] * M3[k][j];
            Tm[i][j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += Tm[i][k] * R[j][k];
            N3[i][j] = acc;
        }
    for (int a = 0; a < 3; a++) c.w[a] = wN[a];
    c.I[0] = N3[0][0]; c.I[1] = N3[1][1]; c.I[2] = N3[2][2];
    c.I[3] = N3[0][1]; c.I[4] = N3[0][2]; c.I[5] = N3[1][2];
}

__global__ __launch_bounds__(kIbmBlock) void k_ibm_velocity(IbmParams P) {
    const int s = blockIdx.x * kIbmBlock + threadIdx.x;
    if (s < P.n_spheres) sphere_velocity(P, s);
}

// ---------------------------------------------------------------------------
// one iteration (ibm.cu:194-542 and the Euler sum, particleEulerNodesUpdate.cu:287-330)
//
// The reference's single kernel interpolates, computes each node's force and
// spreads it into auxiliary fields that a later full-lattice pass adds to u
// and f. Here the iteration is two kernels: (1) interpolation and node force
// for every node of every sphere, (2) spreading of those forces straight into
// u (as F / 2 rho) and the force field. All interpolation of an iteration
// reads u before any spreading of that iteration, as in the reference.
// ---------------------------------------------------------------------------

// Stencil weights and wrapped coordinates of a node.
struct NodeStencil {
    Stencil s;
    double sv[3][4];
    int cx[4], cy[4], cz[4];
};

__device__ __forceinline__ bool node_stencil(const IbmParams& P, const double pos[3], NodeStencil& ns) {
    ns.s = make_stencil(P, pos);
    if (!ns.s.valid) return false;
    for (int a = 0; a < 3; a++)
        for (int j = 0; j < 4; j++)
            ns.sv[a][j] = (j >= ns.s.mn[a] && j <= ns.s.mx[a]) ? peskin4(ns.s.base[a] + j - pos[a]) : 0.0;
    for (int j = 0; j < 4; j++) {
        ns.cx[j] = stencil_coord(P, ns.s, 0, j);
        ns.cy[j] = stencil_coord(P, ns.s, 1, j);
        ns.cz[j] = stencil_coord(P, ns.s, 2, j);
    }
    return true;
}

// One wave per Lagrangian node, one lane per point of its 4^3 stencil
// (lane = x + 4 y + 16 z), so the loads and stores of a wave fall on 16 rows
// of 4 neighbouring nodes. Lanes 0..11 evaluate the 12 kernel weights and
// the others read theirs with shuffles.
constexpr int kInterpBlock = 256;
constexpr int kSphereBlock = 1024;
constexpr int kRowLanesShare = 16;   // lanes per node in the row-parallel kernels

struct PointOf {
    int xi, yj, zk;
    bool in;          // the point survives the wall clipping
    double a;         // its weight (phi_z phi_y) phi_x, 0 when clipped
};

__device__ __forceinline__ PointOf stencil_point(const Stencil& st, const double pos[3], int lane) {
    PointOf pt;
    pt.xi = lane & 3;
    pt.yj = (lane >> 2) & 3;
    pt.zk = lane >> 4;
    double wl = 0.0;
    if (lane < 12) {
        const int a = lane >> 2, j = lane & 3;
        if (j >= st.mn[a] && j <= st.mx[a]) wl = peskin4(st.base[a] + j - pos[a]);
    }
    const double wx = __shfl(wl, pt.xi, 64), wy = __shfl(wl, 4 + pt.yj, 64), wz = __shfl(wl, 8 + pt.zk, 64);
    pt.in = pt.xi >= st.mn[0] && pt.xi <= st.mx[0] && pt.yj >= st.mn[1] && pt.yj <= st.mx[1] &&
            pt.zk >= st.mn[2] && pt.zk <= st.mx[2];
    pt.a = pt.in ? (wz * wy) * wx : 0.0;
    return pt;
}

// 16-lane groups (one node per group, lane r = one (z, y) row): lanes 0..11
// of the group evaluate the node's 12 one-dimensional kernel weights once and
// every lane reads its row weight phi_z phi_y and the four x weights with
// shuffles inside the group. Called by all lanes (uniform control flow).
struct RowWeights {
    double zy;         // phi_z(zk) phi_y(yj) of the lane's row
    double x[4];       // phi_x of the four x points
};

__device__ __forceinline__ RowWeights row_weights(int bx, int by, int bz, double px, double py, double pz, bool ok,
                                                  int r) {
    double wl = 0.0;
    if (ok && r < 12) {
        const int a = r >> 2, j = r & 3;
        // the same expression as peskin4(st.base[a] + j - pos[a])
        const int b = a == 0 ? bx : (a == 1 ? by : bz);
        const double pa = a == 0 ? px : (a == 1 ? py : pz);
        wl = peskin4(b + j - pa);
    }
    const int g = (int)(threadIdx.x & 63) & ~(kRowLanesShare - 1);
    RowWeights w;
    w.zy = __shfl(wl, g + 8 + (r >> 2), 64) * __shfl(wl, g + 4 + (r & 3), 64);
#pragma unroll
    for (int xi = 0; xi < 4; xi++) w.x[xi] = __shfl(wl, g + xi, 64);
    return w;
}

// (1) one workgroup per chunk of up to 64 nodes of one sphere, 16 lanes per
// node (one per (z, y) row of the stencil): interpolated rho and u (the stage
// records' u + (f - f0) / 2 rho, the velocity after the iterations so far,
// one vector load per point) over the lane's row, the 16 row sums combined by
// a butterfly; the node's force
// dF = 2 rho dA thickness (u - u_node) and its moment about the centre; the
// chunk's force and moment sums in a fixed order (strided per-thread partial
// sums, wave64 tree), added to the sphere by k_ibm_add_chunks.
constexpr int kRowLanes = 16;

// Workgroups go to the 8 XCDs of the device in turn (blockIdx mod 8), each
// XCD with its own L2. Block b takes chunk (b mod 8) ceil(n / 8) + b / 8, so
// consecutive chunks (one sphere's) run on one XCD and share its L2 copy of
// the sphere's stage records. The grid is 8 ceil(n / 8) blocks.
__device__ __forceinline__ int xcd_chunk(int n_chunks) {
    const int per = (n_chunks + 7) / 8;
    return (int)(blockIdx.x % 8u) * per + (int)(blockIdx.x / 8u);
}
inline uint32_t xcd_chunk_grid(int n_chunks) { return 8u * (uint32_t)((n_chunks + 7) / 8); }

template <typename T>
__global__ __launch_bounds__(kInterpBlock) void k_ibm_interp(IbmParams P, IbmFluid<T> fl) {
    __shared__ double red[6][kInterpBlock / 64];
    const int chunk = xcd_chunk(P.n_chunks);
    if (chunk >= P.n_chunks) return;
    const IbmSphere& c = P.sph[P.chunk_sphere[chunk]];
    const int first = P.chunk_first[chunk];
    const int r = threadIdx.x & (kRowLanes - 1);
    const int zk = r >> 2, yj = r & 3;
    const int items = P.chunk_count[chunk] * kRowLanes;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    // the sphere's centre and motion in registers (the node stores below
    // could alias them in memory, so they would be reloaded every round)
    const double cpos[3] = {c.pos[0], c.pos[1], c.pos[2]};
    const bool movable = c.movable;
    const double cvel[3] = {c.vel[0], c.vel[1], c.vel[2]}, cw[3] = {c.w[0], c.w[1], c.w[2]};
    // position, area and accumulated force of the node of item k (the
    // chunk's first node past the end); the next item's are loaded while the
    // current one is interpolated
    auto fetch = [&](int k, double (&q)[7]) {
        const int i = first + (k < items ? k / kRowLanes : 0);
        q[0] = P.px[i]; q[1] = P.py[i]; q[2] = P.pz[i];
        q[3] = P.pS[i];
        q[4] = P.pfx[i]; q[5] = P.pfy[i]; q[6] = P.pfz[i];
    };
    double nxt[7] = {0, 0, 0, 0, 0, 0, 0};
    if (items > 0) fetch((int)threadIdx.x, nxt);
    for (int b0 = 0; b0 < items; b0 += kInterpBlock) {      // uniform trip count for the shuffles
        const int k = b0 + (int)threadIdx.x;
        const bool act = k < items;
        const int i = first + k / kRowLanes;
        double pv[4] = {0, 0, 0, 0};
        const double pos[3] = {nxt[0], nxt[1], nxt[2]};
        const double area = nxt[3], pf[3] = {nxt[4], nxt[5], nxt[6]};
        if (b0 + kInterpBlock < items) fetch(k + kInterpBlock, nxt);
        Stencil st;
        st.valid = false;
        if (act) st = make_stencil(P, pos);
        const RowWeights rw = row_weights(st.base[0], st.base[1], st.base[2], pos[0], pos[1], pos[2], act && st.valid, r);
        if (act) {
            if (st.valid && zk >= st.mn[2] && zk <= st.mx[2] && yj >= st.mn[1] && yj <= st.mx[1]) {
                const double a1 = rw.zy;
                const size_t row = ((size_t)stencil_coord(P, st, 2, zk) * P.ny + stencil_coord(P, st, 1, yj)) * P.nx;
                // all four x points at once (weight 0 and an in-range node for
                // points the wall clips), so the loads issue together
                double wx[4];
                size_t kx[4];
#pragma unroll
                for (int xi = 0; xi < 4; xi++) {
                    const bool in = xi >= st.mn[0] && xi <= st.mx[0];
                    wx[xi] = in ? a1 * rw.x[xi] : 0.0;
                    kx[xi] = row + stencil_coord(P, st, 0, in ? xi : st.mn[0]);
                }
                StageRec<T> v[4];
#pragma unroll
                for (int xi = 0; xi < 4; xi++) v[xi] = fl.stage[kx[xi]];
#pragma unroll
                for (int xi = 0; xi < 4; xi++) {
                    pv[0] += (double)v[xi].rho * wx[xi];
                    pv[1] += (double)v[xi].ux * wx[xi];
                    pv[2] += (double)v[xi].uy * wx[xi];
                    pv[3] += (double)v[xi].uz * wx[xi];
                }
            }
        }
        for (int off = 1; off < kRowLanes; off <<= 1)
            for (int q = 0; q < 4; q++) pv[q] += __shfl_xor(pv[q], off, 64);
        if (!act || r != 0) continue;
        if (!st.valid) {
            P.pdx[i] = 0; P.pdy[i] = 0; P.pdz[i] = 0;
            P.pmx[i] = 0; P.pmy[i] = 0; P.pmz[i] = 0;
            continue;
        }
        double d[3] = {pos[0] - cpos[0], pos[1] - cpos[1], pos[2] - cpos[2]};
        for (int a = 0; a < 3; a++) {
            if (!P.periodic[a]) continue;
            const double n = axis_n(P, a);
            if (fabs(d[a]) > n / 2.0) d[a] = (d[a] < 0) ? (pos[a] + n) - cpos[a] : (pos[a] - n) - cpos[a];
        }
        double uc[3] = {0, 0, 0};
        if (movable) {
            uc[0] = cvel[0] + (cw[1] * d[2] - cw[2] * d[1]);
            uc[1] = cvel[1] + (cw[2] * d[0] - cw[0] * d[2]);
            uc[2] = cvel[2] + (cw[0] * d[1] - cw[1] * d[0]);
        }
        const double aux = 2 * pv[0] * area * P.thickness;
        const double dF[3] = {aux * (pv[1] - uc[0]), aux * (pv[2] - uc[1]), aux * (pv[3] - uc[2])};
        const double m3[3] = {d[1] * dF[2] - d[2] * dF[1], d[2] * dF[0] - d[0] * dF[2], d[0] * dF[1] - d[1] * dF[0]};
        P.pfx[i] = pf[0] + dF[0]; P.pfy[i] = pf[1] + dF[1]; P.pfz[i] = pf[2] + dF[2];
        P.pdx[i] = dF[0]; P.pdy[i] = dF[1]; P.pdz[i] = dF[2];
        P.pmx[i] = m3[0]; P.pmy[i] = m3[1]; P.pmz[i] = m3[2];
        acc[0] += dF[0]; acc[1] += dF[1]; acc[2] += dF[2];
        acc[3] += m3[0]; acc[4] += m3[1]; acc[5] += m3[2];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 6; a++) acc[a] += __shfl_down(acc[a], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 6; a++) red[a][wave] = acc[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[6] = {0, 0, 0, 0, 0, 0};
        for (int w = 0; w < kInterpBlock / 64; w++)
            for (int a = 0; a < 6; a++) t[a] += red[a][w];
        for (int a = 0; a < 6; a++) P.chunk_sums[6 * chunk + a] = t[a];
    }
}

// chunk sums -> sphere force and torque, chunks in order; velocity update
__global__ __launch_bounds__(kIbmBlock) void k_ibm_add_chunks(IbmParams P) {
    const int s = blockIdx.x * kIbmBlock + threadIdx.x;
    if (s >= P.n_spheres) return;
    IbmSphere& c = P.sph[s];
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int k = P.chunk0[s]; k < P.chunk0[s + 1]; k++)
        for (int a = 0; a < 6; a++) t[a] += P.chunk_sums[6 * k + a];
    for (int a = 0; a < 3; a++) { c.f[a] += t[a]; c.M[a] += t[3 + a]; }
    sphere_velocity(P, s);
}

// Adds the spread force v to node kk of density rho: to the force field and,
// as v / 2 rho, to the node's stage velocity (velAux of spreadForce,
// ibm.cu:489-509).
template <typename T>
__device__ __forceinline__ void add_spread(const IbmFluid<T>& fl, size_t kk, T rho, T vx, T vy, T vz) {
    unsafeAtomicAdd(fl.fx + kk, vx);
    unsafeAtomicAdd(fl.fy + kk, vy);
    unsafeAtomicAdd(fl.fz + kk, vz);
    StageRec<T>* r = fl.stage + kk;
    const T h = (T)0.5 / rho;
    unsafeAtomicAdd(&r->ux, vx * h);
    unsafeAtomicAdd(&r->uy, vy * h);
    unsafeAtomicAdd(&r->uz, vz * h);
}

template <typename T>
__device__ __forceinline__ void add_spread(const IbmFluid<T>& fl, size_t kk, T vx, T vy, T vz) {
    add_spread<T>(fl, kk, fl.stage[kk].rho, vx, vy, vz);
}

// u + v h with the product rounded before the sum, as an atomic add of the
// rounded product gives it (no fused multiply-add)
__device__ __forceinline__ float add_rounded_product(float u, float v, float h) {
#pragma clang fp contract(off)
    return u + v * h;
}

// (2a) spreading into the force field with global atomics, one wave per node
// (FP64 fluids, and FP32 when the LDS box does not fit).
template <typename T>
__global__ __launch_bounds__(kInterpBlock) void k_ibm_spread(IbmParams P, IbmFluid<T> fl) {
    const int lane = threadIdx.x & 63;
    const int i = (int)((blockIdx.x * kInterpBlock + threadIdx.x) >> 6);
    if (i >= P.n_nodes) return;
    const double pos[3] = {P.px[i], P.py[i], P.pz[i]};
    const Stencil st = make_stencil(P, pos);
    if (!st.valid) return;
    const PointOf pt = stencil_point(st, pos, lane);
    if (!pt.in) return;
    const size_t kk = ((size_t)stencil_coord(P, st, 2, pt.zk) * P.ny + stencil_coord(P, st, 1, pt.yj)) * P.nx +
                      stencil_coord(P, st, 0, pt.xi);
    add_spread<T>(fl, kk, (T)(-P.pdx[i] * pt.a), (T)(-P.pdy[i] * pt.a), (T)(-P.pdz[i] * pt.a));
}

// (2b) FP32 fluids: one workgroup per sphere, 16 lanes per node (one stencil
// row each), collects the spread force in an LDS box of edge^3 nodes around
// the sphere and adds the box to the force field and the stage velocities
// with one global atomic per box node, field and component. Points outside
// the box (nodes that left the sphere's radius) go straight to global memory.
//
// The box holds fixed-point integers (LDS integer atomics; float LDS atomics
// cost several times more): component a of a contribution -dF_a w is stored
// as rint(-dF_a w s_a) with s_a = 2^30 / (sum over the sphere's nodes of
// |dF_a| / 8). A box node takes at most one contribution per node, each with
// weight w <= 1/8 (the 4-point kernel is at most 1/2 per axis), so every box
// sum stays within 2^30 plus the rounding, inside int32. The sums are exact
// and independent of the order of the atomics; the rounding of a
// contribution is at most 2^-31 of the sphere's force sum. A sphere whose
// force sum is not finite spreads through the global float atomics, so that
// the fluid sees the same NaN or infinity as the sphere.
__device__ __forceinline__ int box_fixed(double v, double s) { return (int)rint(v * s); }

__global__ __launch_bounds__(kSphereBlock) void k_ibm_spread_box(IbmParams P, IbmFluid<float> fl) {
    extern __shared__ int box[];                   // [3][edge^3]
    __shared__ double red[3][kSphereBlock / 64];
    __shared__ double scale[3];
    __shared__ int finite[3];
    __shared__ int nb_o[kBoxNb][3];                // origins of the overlapping boxes
    const IbmSphere& c = P.sph[blockIdx.x];
    const int* plan = P.box_nb + blockIdx.x * (kBoxNb + 1);
    const int n_nb = P.box_plain && *P.box_stray == 0 ? plan[0] : -1;   // -1: atomics everywhere
    if ((int)threadIdx.x < n_nb) {
        const IbmSphere& cj = P.sph[plan[1 + threadIdx.x]];
        for (int a = 0; a < 3; a++) nb_o[threadIdx.x][a] = box_origin(P, cj, a);
    }
    const int edge = P.box_edge;
    const int e3 = edge * edge * edge;
    for (int e = threadIdx.x; e < 3 * e3; e += kSphereBlock) box[e] = 0;
    const int half = edge / 2 - 3;                 // ceil(radius)
    int o[3];
    for (int a = 0; a < 3; a++) o[a] = (int)floor(c.pos[a]) - half - 2;
    const int first = c.first_node;
    const int items = c.n_nodes * kRowLanes;
    {
        double sa[3] = {0, 0, 0};
        for (int k = threadIdx.x; k < c.n_nodes; k += kSphereBlock) {
            sa[0] += fabs(P.pdx[first + k]); sa[1] += fabs(P.pdy[first + k]); sa[2] += fabs(P.pdz[first + k]);
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int a = 0; a < 3; a++) sa[a] += __shfl_down(sa[a], off, 64);
        if ((threadIdx.x & 63) == 0)
            for (int a = 0; a < 3; a++) red[a][threadIdx.x >> 6] = sa[a];
        __syncthreads();
        if (threadIdx.x < 3) {
            double t = 0;
            for (int w = 0; w < kSphereBlock / 64; w++) t += red[threadIdx.x][w];
            scale[threadIdx.x] = t > 0 ? 1073741824.0 / (0.125 * t) : 1.0;
            finite[threadIdx.x] = isfinite(t) ? 1 : 0;
        }
    }
    const int r = threadIdx.x & (kRowLanes - 1);
    const int zk = r >> 2, yj = r & 3;
    // position and force of the node of item k (the sphere's first node past
    // the end); the next item's are loaded while the current one is spread
    auto fetch = [&](int k, double (&q)[6]) {
        const int i = first + (k < items ? k / kRowLanes : 0);
        q[0] = P.px[i]; q[1] = P.py[i]; q[2] = P.pz[i];
        q[3] = P.pdx[i]; q[4] = P.pdy[i]; q[5] = P.pdz[i];
    };
    double nxt[6] = {0, 0, 0, 0, 0, 0};
    if (items > 0) fetch((int)threadIdx.x, nxt);
    __syncthreads();
    const double sx = scale[0], sy = scale[1], sz = scale[2];
    const bool use_box = finite[0] && finite[1] && finite[2];
    for (int b0 = 0; b0 < items; b0 += kSphereBlock) {      // uniform trip count for the shuffles
        const int k = b0 + (int)threadIdx.x;
        const bool act = k < items;
        double q[6];
#pragma unroll
        for (int j = 0; j < 6; j++) q[j] = nxt[j];
        if (b0 + kSphereBlock < items) fetch(k + kSphereBlock, nxt);
        const double pos[3] = {q[0], q[1], q[2]};
        Stencil st;
        st.valid = false;
        if (act) st = make_stencil(P, pos);
        const RowWeights rw = row_weights(st.base[0], st.base[1], st.base[2], pos[0], pos[1], pos[2], act && st.valid, r);
        if (!act || !st.valid || zk < st.mn[2] || zk > st.mx[2] || yj < st.mn[1] || yj > st.mx[1]) continue;
        const double a1 = rw.zy;
        const int cz = stencil_coord(P, st, 2, zk), cy = stencil_coord(P, st, 1, yj);
        int bz = st.base[2] + zk - o[2], by = st.base[1] + yj - o[1];
        if (P.periodic[2]) bz = wrap_axis(bz, P.nz);
        if (P.periodic[1]) by = wrap_axis(by, P.ny);
#pragma unroll
        for (int xi = 0; xi < 4; xi++) {
            if (xi < st.mn[0] || xi > st.mx[0]) continue;
            const double a = a1 * rw.x[xi];
            const float vx = (float)(-q[3] * a), vy = (float)(-q[4] * a), vz = (float)(-q[5] * a);
            int bx = st.base[0] + xi - o[0];
            if (P.periodic[0]) bx = wrap_axis(bx, P.nx);
            if (use_box && bx >= 0 && bx < edge && by >= 0 && by < edge && bz >= 0 && bz < edge) {
                const int e = (bz * edge + by) * edge + bx;
                atomicAdd(box + e, box_fixed(-q[3] * a, sx));
                atomicAdd(box + e3 + e, box_fixed(-q[4] * a, sy));
                atomicAdd(box + 2 * e3 + e, box_fixed(-q[5] * a, sz));
            } else {
                const size_t kk = ((size_t)cz * P.ny + cy) * P.nx + stencil_coord(P, st, 0, xi);
                add_spread<float>(fl, kk, vx, vy, vz);
            }
        }
    }
    __syncthreads();
    // flush in rounds of kFlushBatch box nodes per thread. A node that no
    // other sphere's box covers (and, with use_box, no other workgroup
    // writes: see box_plan) takes plain loads and stores, the others
    // global atomics; both add the same float values, so a node's result is
    // the same either way. The loads of a round are issued before its
    // atomics and stores, so each waits for the previous round's only
    // (memory operations retire in order).
    constexpr int kFlushBatch = 8;
    const bool plain_ok = use_box && n_nb >= 0;
    for (int e0 = 0; e0 < e3; e0 += kFlushBatch * kSphereBlock) {
        size_t kk[kFlushBatch];
        StageRec<float> rec[kFlushBatch];
        float f[kFlushBatch][3];
        bool on[kFlushBatch], plain[kFlushBatch];
#pragma unroll
        for (int j = 0; j < kFlushBatch; j++) {
            const int e = e0 + j * kSphereBlock + (int)threadIdx.x;
            on[j] = e < e3 && (box[e] != 0 || box[e3 + e] != 0 || box[2 * e3 + e] != 0);
            int q[3] = {o[0] + e % edge, o[1] + (e / edge) % edge, o[2] + e / (edge * edge)};
            plain[j] = plain_ok;
            for (int k = 0; k < kBoxNb; k++)
                if (k < n_nb && in_box_axis(P, 0, q[0], nb_o[k][0]) && in_box_axis(P, 1, q[1], nb_o[k][1]) &&
                    in_box_axis(P, 2, q[2], nb_o[k][2]))
                    plain[j] = false;
            for (int a = 0; a < 3; a++) {
                const int n = axis_n(P, a);
                if (P.periodic[a]) q[a] = wrap_axis(q[a], n);
                else if (q[a] < 0 || q[a] >= n) on[j] = false;
            }
            kk[j] = on[j] ? ((size_t)q[2] * P.ny + q[1]) * P.nx + q[0] : 0;
            rec[j] = on[j] ? fl.stage[kk[j]] : StageRec<float>{1.f, 0.f, 0.f, 0.f};
            if (on[j] && plain[j]) {
                f[j][0] = fl.fx[kk[j]]; f[j][1] = fl.fy[kk[j]]; f[j][2] = fl.fz[kk[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < kFlushBatch; j++) {
            const int e = e0 + j * kSphereBlock + (int)threadIdx.x;
            if (!on[j]) continue;
            const float vx = (float)((double)box[e] / sx), vy = (float)((double)box[e3 + e] / sy),
                        vz = (float)((double)box[2 * e3 + e] / sz);
            if (plain[j]) {
                const float h = 0.5f / rec[j].rho;
                fl.fx[kk[j]] = f[j][0] + vx;
                fl.fy[kk[j]] = f[j][1] + vy;
                fl.fz[kk[j]] = f[j][2] + vz;
                fl.stage[kk[j]] = StageRec<float>{rec[j].rho, add_rounded_product(rec[j].ux, vx, h),
                                                  add_rounded_product(rec[j].uy, vy, h),
                                                  add_rounded_product(rec[j].uz, vz, h)};
            } else {
                add_spread<float>(fl, kk[j], rec[j].rho, vx, vy, vz);
            }
        }
    }
}

// saved macroscopics on the listed tiles: rho and the velocity after the
// stage, u + (f - f0) / 2 rho (the reference's u after its Euler sums), from
// the stage records
template <typename T>
__global__ __launch_bounds__(kIbmBlock) void k_ibm_output(IbmParams P, IbmFluid<T> fl) {
    const uint32_t total = *P.tile_count * (uint32_t)kTileNodes;
    for (uint32_t e = blockIdx.x * kIbmBlock + threadIdx.x; e < total; e += gridDim.x * kIbmBlock) {
        int x, y, z;
        tile_node(P.tile_list[e / kTileNodes], (int)(e % kTileNodes), P.tiles[0], P.tiles[1], x, y, z);
        if (x >= P.nx || y >= P.ny || z >= P.nz) continue;
        const size_t k = ((size_t)z * P.ny + y) * P.nx + x;
        const StageRec<T> r = fl.stage[k];
        fl.out_rho[k] = r.rho;
        fl.out_ux[k] = r.ux;
        fl.out_uy[k] = r.uy;
        fl.out_uz[k] = r.uz;
    }
}

// ---------------------------------------------------------------------------
// movement (gpuParticleMovement ibm.cu:899-993, gpuParticleNodeMovement
// ibm.cu:1044-1197)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIbmBlock) void k_ibm_move(IbmParams P) {
    const int s = blockIdx.x * kIbmBlock + threadIdx.x;
    if (s >= P.n_spheres) return;
    IbmSphere& c = P.sph[s];
    if (!c.movable) return;
    const double al = P.alpha;
    for (int a = 0; a < 3; a++) {
        const double dp = (c.vel[a] * al + c.vel_old[a] * (1.0 - al));
        c.pos[a] = P.periodic[a] ? wrap_pos(P, a, c.pos[a] + dp) : c.pos[a] + dp;
    }
    for (int a = 0; a < 3; a++) {
        c.w_avg[a] = (c.w[a] * al + c.w_old[a] * (1.0 - al));
        c.w_pos[a] += c.w_avg[a];
    }
    // semi-axis points rotate about the centre with no guard on |w_avg| = 0
    // (ibm.cu:958-991): a step without rotation leaves them NaN, as there
    const double wn = sqrt(c.w_avg[0] * c.w_avg[0] + c.w_avg[1] * c.w_avg[1] + c.w_avg[2] * c.w_avg[2]);
    const double q0 = cos(0.5 * wn);
    const double qi = (c.w_avg[0] / wn) * sin(0.5 * wn);
    const double qj = (c.w_avg[1] / wn) * sin(0.5 * wn);
    const double qk = (c.w_avg[2] / wn) * sin(0.5 * wn);
    const double t = (q0 * q0) - 0.5;
    for (int k = 0; k < 3; k++) {
        const double v[3] = {c.axis[k][0] - c.pos_old[0], c.axis[k][1] - c.pos_old[1], c.axis[k][2] - c.pos_old[2]};
        c.axis[k][0] = c.pos[0] + 2 * ((t + (qi * qi)) * v[0] + ((qi * qj) - (q0 * qk)) * v[1] + ((qi * qk) + (q0 * qj)) * v[2]);
        c.axis[k][1] = c.pos[1] + 2 * (((qi * qj) + (q0 * qk)) * v[0] + (t + (qj * qj)) * v[1] + ((qj * qk) - (q0 * qi)) * v[2]);
        c.axis[k][2] = c.pos[2] + 2 * (((qi * qj) - (q0 * qj)) * v[0] + ((qj * qk) + (q0 * qi)) * v[1] + (t + (qk * qk)) * v[2]);
    }
}

// the rotation keeps the reference's (qi*qj) - (q0*qj) entry (ibm.cu:1174)
__global__ __launch_bounds__(kIbmBlock) void k_ibm_move_nodes(IbmParams P) {
    const uint32_t i = blockIdx.x * kIbmBlock + threadIdx.x;
    if (i >= (uint32_t)P.n_nodes) return;
    const IbmSphere& c = P.sph[P.pid[i]];
    if (!c.movable) return;
    double p3[3] = {P.px[i], P.py[i], P.pz[i]};
    const double wn = sqrt(c.w_avg[0] * c.w_avg[0] + c.w_avg[1] * c.w_avg[1] + c.w_avg[2] * c.w_avg[2]);
    if (wn <= 1e-8) {
        for (int a = 0; a < 3; a++) {
            double dd = c.pos[a] - c.pos_old[a];
            if (P.periodic[a]) {
                const double n = axis_n(P, a);
                if (fabs(dd) > n / 2.0) dd = (c.pos[a] < c.pos_old[a]) ? (c.pos[a] + n) - c.pos_old[a] : (c.pos[a] - n) - c.pos_old[a];
                p3[a] = wrap_pos(P, a, p3[a] + dd);
            } else {
                p3[a] += dd;
            }
        }
    } else {
        double v[3];
        for (int a = 0; a < 3; a++) {
            v[a] = p3[a] - c.pos_old[a];
            if (P.periodic[a]) {
                const double n = axis_n(P, a);
                if (fabs(v[a]) > n / 2.0) {
                    if (p3[a] < c.pos_old[a]) p3[a] += n;
                    else p3[a] -= n;
                }
                v[a] = p3[a] - c.pos_old[a];
            }
        }
        const double q0 = cos(0.5 * wn);
        const double qi = (c.w_avg[0] / wn) * sin(0.5 * wn);
        const double qj = (c.w_avg[1] / wn) * sin(0.5 * wn);
        const double qk = (c.w_avg[2] / wn) * sin(0.5 * wn);
        const double t = (q0 * q0) - 0.5;
        const double np3[3] = {
            c.pos[0] + 2 * ((t + (qi * qi)) * v[0] + ((qi * qj) - (q0 * qk)) * v[1] + ((qi * qk) + (q0 * qj)) * v[2]),
            c.pos[1] + 2 * (((qi * qj) + (q0 * qk)) * v[0] + (t + (qj * qj)) * v[1] + ((qj * qk) - (q0 * qi)) * v[2]),
            c.pos[2] + 2 * (((qi * qj) - (q0 * qj)) * v[0] + ((qj * qk) + (q0 * qi)) * v[1] + (t + (qk * qk)) * v[2])};
        for (int a = 0; a < 3; a++) p3[a] = P.periodic[a] ? wrap_pos(P, a, np3[a]) : np3[a];
    }
    P.px[i] = p3[0]; P.py[i] = p3[1]; P.pz[i] = p3[2];
}

}  // namespace vlbm
