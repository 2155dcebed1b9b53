// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Velocity sets follow src/CUDA/velocitySets/D3Q19.h:31-46 and D3Q27.h:39-56
// (same population numbering: D3Q19 is the first 19 entries of D3Q27;
// opposite pairs are (2k-1, 2k)). Node-type words follow
// src/CUDA/structs/nodeTypeMap.h:15-108 bit for bit.
#pragma once

#include <stdint.h>
#include <stddef.h>

namespace vlbm {

// ---- velocity set -------------------------------------------------------
constexpr int8_t kCX[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1};
constexpr int8_t kCY[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1, 1,-1};
constexpr int8_t kCZ[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1, 1,-1};

__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

template <int Q>
__host__ __device__ constexpr double weight(int i) {
    return Q == 19 ? (i == 0 ? 1.0 / 3 : (i < 7 ? 1.0 / 18 : 1.0 / 36))
                   : (i == 0 ? 8.0 / 27 : (i < 7 ? 2.0 / 27 : (i < 19 ? 1.0 / 54 : 1.0 / 216)));
}

__host__ __device__ constexpr int find_pop(int cx, int cy, int cz) {
    for (int i = 0; i < 27; i++)
        if (kCX[i] == cx && kCY[i] == cy && kCZ[i] == cz) return i;
    return -1;
}

// ---- node-type word -------------------------------------------------------
enum : uint32_t {
    kOffInterp = 23, kOffSave = 22, kOffUsed = 21, kOffScheme = 18, kOffDir = 13,
    kOffGeom = 12, kOffUx = 9, kOffUy = 6, kOffUz = 3, kOffRho = 0
};
enum Scheme : int { kNull = 0, kVelZouHe = 1, kVelBB = 2, kPresZouHe = 3, kFreeSlip = 4,
                    kBounceBack = 5, kInterpBB = 6, kSpecial = 7 };
enum Direction : int { dN = 0, dS, dW, dE, dF, dB, dNW, dNE, dNF, dNB, dSW, dSE, dSF, dSB,
                       dWF, dWB, dEF, dEB, dNWF, dNWB, dNEF, dNEB, dSWF, dSWB, dSEF, dSEB };

__host__ __device__ constexpr uint32_t make_word(bool used, bool save, int scheme, int dir, int geom,
                                                  int uxi, int uyi, int uzi, int rhoi) {
    return ((uint32_t)used << kOffUsed) | ((uint32_t)save << kOffSave) | ((uint32_t)scheme << kOffScheme) |
           ((uint32_t)dir << kOffDir) | ((uint32_t)geom << kOffGeom) | ((uint32_t)uxi << kOffUx) |
           ((uint32_t)uyi << kOffUy) | ((uint32_t)uzi << kOffUz) | ((uint32_t)rhoi << kOffRho);
}
__host__ __device__ inline bool w_used(uint32_t w) { return (w >> kOffUsed) & 1u; }
__host__ __device__ inline int w_scheme(uint32_t w) { return (w >> kOffScheme) & 7u; }
__host__ __device__ inline int w_dir(uint32_t w) { return (w >> kOffDir) & 31u; }
__host__ __device__ inline int w_geom(uint32_t w) { return (w >> kOffGeom) & 1u; }
__host__ __device__ inline int w_uxi(uint32_t w) { return (w >> kOffUx) & 7u; }
__host__ __device__ inline int w_uyi(uint32_t w) { return (w >> kOffUy) & 7u; }
__host__ __device__ inline int w_uzi(uint32_t w) { return (w >> kOffUz) & 7u; }
__host__ __device__ inline int w_rhoi(uint32_t w) { return (w >> kOffRho) & 7u; }
__host__ __device__ inline unsigned w_interp(uint32_t w) { return (w >> kOffInterp) & 0xffu; }

// ---- wall faces of each direction code -------------------------------------
// A face is named by its outward normal: N=+y, S=-y, W=-x, E=+x, F=+z, B=-z.
enum : unsigned { fN = 1, fS = 2, fW = 4, fE = 8, fF = 16, fB = 32 };
constexpr uint8_t kDirFaces[26] = {
    fN, fS, fW, fE, fF, fB,
    fN | fW, fN | fE, fN | fF, fN | fB, fS | fW, fS | fE, fS | fF, fS | fB,
    fW | fF, fW | fB, fE | fF, fE | fB,
    fN | fW | fF, fN | fW | fB, fN | fE | fF, fN | fE | fB,
    fS | fW | fF, fS | fW | fB, fS | fE | fF, fS | fE | fB};

// Population i arrives at a node on these faces from outside the lattice.
__host__ __device__ constexpr bool incoming(int i, unsigned faces) {
    return ((faces & fN) && kCY[i] == -1) || ((faces & fS) && kCY[i] == 1) ||
           ((faces & fW) && kCX[i] == 1) || ((faces & fE) && kCX[i] == -1) ||
           ((faces & fF) && kCZ[i] == -1) || ((faces & fB) && kCZ[i] == 1);
}

// Halfway bounce-back assigns every incoming population whose opposite is
// outgoing; populations incoming from two faces at once (both i and opp(i)
// incoming) keep their streamed value.
template <int Q>
__host__ __device__ constexpr uint32_t bb_mask(int dir) {
    uint32_t m = 0;
    for (int i = 1; i < Q; i++)
        if (incoming(i, kDirFaces[dir]) && !incoming(opp(i), kDirFaces[dir])) m |= (1u << i);
    return m;
}

// ---- simulation parameters in kernel form ---------------------------------
struct FastDiv {      // n / d for n < 2^31 by multiply-shift
    uint64_t m;
    uint32_t d, shift;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    f.shift = 32 + l;
    f.m = ((1ull << f.shift) + d - 1) / d;
    return f;
}

// Elements between two planes of the state (population or coefficient
// planes): the node count rounded up to 64, so every plane starts on a
// 256-byte boundary (FP32). With planes packed back to back, node counts that
// are not a multiple of 16 cost a quarter to a third of the step rate
// (D3Q19 FP32 185x119x81: 23 GLUPS; the same lattice with NX rounded up to
// 16: 32 GLUPS).
constexpr size_t kPlaneAlign = 64;
inline int64_t plane_stride(size_t n) { return (int64_t)((n + kPlaneAlign - 1) / kPlaneAlign * kPlaneAlign); }

// Record of one node in the immersed boundary stage array: rho and the
// velocity the interpolation reads, u + (f - f0) / 2 rho (the constant-force
// velocity plus half the IBM force increments so far over rho), one vector
// load per stencil point.
template <typename T>
struct alignas(4 * sizeof(T)) StageRec {
    T rho, ux, uy, uz;
};

template <typename T>
struct Params {
    int nx, ny, nz;
    uint32_t n;                 // nodes
    FastDiv div_nx, div_nxy;
    int64_t stride;             // elements between population planes
    const void* src;            // state read this step
    void* dst;                  // state written this step
    const void* held;           // idle lattice (the step before the last; equals dst in a step)
    int plates;                 // rheometer: the planes z = 0, NZ-1 are unused
    int first_bc;               // the boundary pass of the first step after init
    const uint32_t* mask;       // node words (mask geometry only)
    uint32_t region_word[27];   // node word per box region (box geometry)
    T ux_bc[8], uy_bc[8], uz_bc[8], rho_bc[8];
    T fx, fy, fz;
    T omega;                    // Newtonian 1/tau
    T omega_p, s_y;             // Bingham
    T k_cons, n_minus_1, gamma_0;   // power law
    int clamp;
    T omega_lo, omega_hi;       // omega bounds from the tau clamp
    T* omega_state;             // power-law omega per node
    T* m_rho; T* m_ux; T* m_uy; T* m_uz; T* m_omega;
    int save;
    int zouhe_textbook;
    int bgk;                 // BGK collision instead of the regularized one
    T tau, les_c;            // LES: tau_t = (sqrt(tau^2 + les_c |S|) - tau) / 2 when les_c != 0
    T* visc_turb;            // LES: tau_t / 3 of every node, every step (null: not kept)
    int aux_mode;               // aux kernel: AuxMode
    int direct;                 // aux kernel: state is post-stream (no gather, no BC)
    uint32_t aux_base, aux_count;   // aux kernel: node range (count 0: every node, or every node of the extra plane)
    int sr;                     // 16-bit storage: stochastic rounding of the stored populations
    uint32_t sr_seed;           // its random stream of this step
    int aux_f64;                // kAuxMaterialize: doubles into dst, planes of aux_count values
    T duct_R, duct_r, w_out, u_out, w_in, u_in;
    // immersed boundary: nodes in 8^3 tiles flagged by the last IBM stage take
    // their force from ibm_f*, every other node the constant (fx, fy, fz)
    const uint32_t* tile_flag;
    const T* ibm_fx; const T* ibm_fy; const T* ibm_fz;
    StageRec<T>* ibm_stage;     // stage records of the listed tiles (written by k_ibm_macr)
    int tiles_x, tiles_y;
};

// Tiles of the IBM stage: 16 x 2 x 2 nodes. A sphere's stencil shell is a
// few nodes thick, so small tiles keep the stage's passes near the shell; 16
// nodes along x make a tile row one 64-byte segment of an FP32 population
// plane (4 x 4 x 4 tiles read 16-byte rows and fetched 3.2 times the bytes
// they used).
constexpr int kTileShiftX = 4, kTileShiftY = 1, kTileShiftZ = 1;
constexpr int kTileNodes = 1 << (kTileShiftX + kTileShiftY + kTileShiftZ);
__host__ __device__ constexpr int tile_shift(int a) { return a == 0 ? kTileShiftX : (a == 1 ? kTileShiftY : kTileShiftZ); }
// tiles along axis a of n nodes
__host__ __device__ constexpr int tiles_along(int n, int a) { return (n + (1 << tile_shift(a)) - 1) >> tile_shift(a); }
// tile holding node (x, y, z)
__host__ __device__ inline uint32_t tile_index(int x, int y, int z, int tiles_x, int tiles_y) {
    return (uint32_t)(x >> kTileShiftX) +
           (uint32_t)tiles_x * ((uint32_t)(y >> kTileShiftY) + (uint32_t)tiles_y * (uint32_t)(z >> kTileShiftZ));
}
// node l (x fastest) of tile t
__host__ __device__ inline void tile_node(uint32_t t, int l, int tiles_x, int tiles_y, int& x, int& y, int& z) {
    const int tx = (int)(t % (uint32_t)tiles_x);
    const int ty = (int)((t / (uint32_t)tiles_x) % (uint32_t)tiles_y);
    const int tz = (int)(t / ((uint32_t)tiles_x * (uint32_t)tiles_y));
    x = (tx << kTileShiftX) + (l & ((1 << kTileShiftX) - 1));
    y = (ty << kTileShiftY) + ((l >> kTileShiftX) & ((1 << kTileShiftY) - 1));
    z = (tz << kTileShiftZ) + (l >> (kTileShiftX + kTileShiftY));
}

// kAuxMomentsOutsideTiles: moments with the constant force for nodes outside
// the flagged tiles; flagged tiles keep the macroscopics of the IBM stage.
// kAuxGhost: the extra z plane of the reference's population layout (deviations
// as doubles, Q planes of NX*NY, into dst).
enum AuxMode : int { kAuxMoments = 0, kAuxMaterialize = 1, kAuxMomentsOutsideTiles = 2, kAuxGhost = 3 };

// State representations: Q planes of post-collision deviations, or the 10
// coefficient planes of the regularized post-collision state.
enum Rep : int { kRepPop = 0, kRepMom = 1 };
constexpr int kCoef = 10;

// Launchers implemented in kernels.hip. prec: 8, 4 or 2 (storage bytes).
// in_rep / out_rep: state representation of the source and destination
// lattices, 0 = populations, 1 = moment coefficients (float / double only;
// coefficients as output need in_rep 1 or a direct step)
int launch_step(int q, int prec, int rheo, bool mask_geom, bool direct, const void* params, void* stream,
                bool ibm = false, int in_rep = 0, int out_rep = 0);
int launch_aux(int q, int prec, int rheo, bool mask_geom, const void* params, void* stream, int in_rep = 0);
int launch_reduce(int prec, const void* rho, const void* ux, const void* uy, const void* uz, void* ux_old,
                  void* uy_old, void* uz_old, uint32_t n, double* partials, double* out3, void* stream);
int launch_fill_feq(int q, int prec, void* state, int64_t stride, uint32_t n, const double* rho,
                    const double* ux, const double* uy, const double* uz, void* stream, int out_rep = 0);
int launch_convert(int prec, void* buf, double* dbl, size_t count, int to_double, void* stream);
int launch_copy_probe(const void* src, void* dst, size_t bytes, void* stream);

} // namespace vlbm
