// Time step of the moment-coefficient representation (kRepMom) with LDS
// staging: one workgroup owns a TX x TY tile of the xy-plane and marches over
// a chunk of z-planes. The 10 post-collision coefficients of planes z-1, z
// and z+1 (tile plus a one-node halo in x and y) sit in a three-slot LDS ring;
// the loads of plane z+2 are issued into registers before plane z is
// computed. Every node reconstructs its Q pulled populations from LDS, applies
// its boundary condition (which reads post-collision populations of the node
// and its neighbours from the same LDS planes), collides and stores its 10
// coefficients: 10 coefficient reads (plus the halo, served mostly by L2) and
// 10 writes per node instead of Q and Q.
#pragma once
#include "kernels_impl.h"

namespace vlbm {

constexpr int kTileX = 64;
// Tile rows: 8 for float D3Q19 box-map kernels (a 79.2 KB ring, 512 threads:
// their 123-131 VGPRs give 4 waves per SIMD with two workgroups per CU, and
// the y halo costs 25 % instead of 50 %), 4 for other float kernels (47.5 KB;
// their 145-166 VGPRs at 64 x 8 would leave one 512-thread workgroup per
// CU), 2 for double (63.4 KB).
template <typename T, int Q, bool MASK> struct TileShape {
    static constexpr int TY = sizeof(T) == 4 ? ((Q == 19 && !MASK) ? 8 : 4) : 2;
    static constexpr int NT = kTileX * TY;
    // waves per SIMD the register allocation must allow (the second launch
    // bound): 4 for 64 x 8 tiles (two workgroups per CU, at most 128 VGPRs)
    static constexpr int MIN_WAVES = TY == 8 ? 4 : 1;
    static constexpr int HX = kTileX + 2, HY = TY + 2, HP = HX * HY;
    static constexpr int PER = (HP + NT - 1) / NT;       // halo positions per thread
};

// Post-collision population i of the node at offset (dx, dy, dz) from the
// tile node at LDS position base. The ring slots of planes z-1, z and z+1 are
// packed 4 bits each into one int (slot of plane z+dz at bit 4 (dz+1)): a
// select among three struct fields becomes an indexed access to the struct
// in private memory, a shift does not.
template <int Q, typename T, bool MASK>
struct TilePC {
    const T* sm;
    int slots;
    int base;
    __device__ __forceinline__ T operator()(int i, int dx, int dy, int dz) const {
        using SH = TileShape<T, Q, MASK>;
        const int slot = (slots >> (4 * (dz + 1))) & 15;
        const T* q = sm + slot * (kCoef * SH::HP) + base + dy * SH::HX + dx;
        return recon<Q, T, T>(i, q, SH::HP);
    }
    // the node's own 10 coefficients, read once for the several populations a
    // bounce-back rebuilds from them
    static constexpr bool kCoefs = true;
    __device__ __forceinline__ void own(T* c) const {
        using SH = TileShape<T, Q, MASK>;
        const T* q = sm + ((slots >> 4) & 15) * (kCoef * SH::HP) + base;
#pragma unroll
        for (int k = 0; k < kCoef; k++) c[k] = q[k * SH::HP];
    }
    // the coefficient layout is refused for the rheometer, the one reader of held values
    __device__ __forceinline__ T held(int) const { return (T)0; }
};

__device__ __forceinline__ int wrap_mod(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// One step over the lattice, coefficient planes src -> dst (p.src, p.dst).
// Grid: 8 ceil(n / 8) workgroups of TX*TY threads for the n = ceil(nx/TX) *
// ceil(ny/TY) * ceil(nz/zc) tiles.
template <int Q, typename T, int RHEO, bool MASK, bool IBM>
__global__ __launch_bounds__((TileShape<T, Q, MASK>::NT), (TileShape<T, Q, MASK>::MIN_WAVES)) void k_step_tile(const Params<T> p,
                                                                                                     int zc) {
    using SH = TileShape<T, Q, MASK>;
    constexpr int TX = kTileX, TY = SH::TY, NT = SH::NT, HX = SH::HX, HP = SH::HP, PER = SH::PER;
    __shared__ T sm[3 * kCoef * HP];
    const int ntx = (p.nx + TX - 1) / TX, nty = (p.ny + TY - 1) / TY;
    // Workgroups go to the 8 XCDs in turn (blockIdx mod 8), each XCD with its
    // own L2. Block b takes tile (b mod 8) ceil(n / 8) + b / 8 of the n tiles,
    // so x- and y-neighbouring tiles run on one XCD at about the same time
    // and share the cache lines of their halos (a tile row's x halo alone
    // fetches 2 extra lines). The grid is 8 ceil(n / 8) blocks.
    const int n_tiles = ntx * nty * ((p.nz + zc - 1) / zc);
    const int b = (int)(blockIdx.x % 8u) * ((n_tiles + 7) / 8) + (int)(blockIdx.x / 8u);
    if (b >= n_tiles) return;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * zc;
    const int z1 = min(z0 + zc, p.nz);
    const int t = threadIdx.x, lx = t % TX, ly = t / TX;
    const uint32_t plane = (uint32_t)p.nx * (uint32_t)p.ny;
    const T* src = (const T*)p.src;

    // halo positions owned by this thread: in-plane offsets of their nodes
    uint32_t poff[PER];
    bool pok[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int r = t + q * NT;
        pok[q] = r < HP;
        const int hx = r % HX, hy = r / HX;
        poff[q] = (uint32_t)wrap_mod(y0 + hy - 1, p.ny) * (uint32_t)p.nx + (uint32_t)wrap_mod(x0 + hx - 1, p.nx);
    }
    T reg[PER][kCoef];
    auto fetch = [&](int zg) {
        const uint32_t zo = (uint32_t)wrap_mod(zg, p.nz) * plane;
#pragma unroll
        for (int q = 0; q < PER; q++)
            if (pok[q]) {
#pragma unroll
                for (int k = 0; k < kCoef; k++) reg[q][k] = src[(int64_t)k * p.stride + zo + poff[q]];
            }
    };
    auto stash = [&](int slot) {
#pragma unroll
        for (int q = 0; q < PER; q++)
            if (pok[q]) {
#pragma unroll
                for (int k = 0; k < kCoef; k++) sm[(slot * kCoef + k) * HP + t + q * NT] = reg[q][k];
            }
    };

    const int x = x0 + lx, y = y0 + ly;
    const bool active = x < p.nx && y < p.ny;
    Node<T> nd;
    nd.x = x; nd.y = y;
    nd.xm = x == 0 ? p.nx - 1 : x - 1;
    nd.xp = x == p.nx - 1 ? 0 : x + 1;
    nd.ym = y == 0 ? p.ny - 1 : y - 1;
    nd.yp = y == p.ny - 1 ? 0 : y + 1;

    fetch(z0 - 1); stash(2);
    fetch(z0); stash(0);
    fetch(z0 + 1);
    int s_m = 2, s_0 = 0, s_p = 1;
    for (int z = z0; z < z1; z++) {
        stash(s_p);
        __syncthreads();
        if (z + 1 < z1) fetch(z + 2);          // in flight while plane z is computed
        if (active) {
            nd.z = z;
            nd.zm = z == 0 ? p.nz - 1 : z - 1;
            nd.zp = z == p.nz - 1 ? 0 : z + 1;
            nd.idx = (uint32_t)z * plane + (uint32_t)y * (uint32_t)p.nx + (uint32_t)x;
            const uint32_t word = node_word<MASK>(p, nd);
            if (w_used(word)) {
                const TilePC<Q, T, MASK> pc{sm, s_m | (s_0 << 4) | (s_p << 8), (ly + 1) * HX + (lx + 1)};
                T g[Q];
                gather<Q, T>(g, pc);
                if (w_scheme(word) != kNull) apply_bc<Q, T, MASK>(g, p, nd, word, pc);
                collide_store<Q, T, RHEO, IBM, kRepMom>(p, nd.idx, nd, g);
            }
        }
        __syncthreads();
        const int old = s_m;
        s_m = s_0; s_0 = s_p; s_p = old;
    }
}

// z-chunk per workgroup: 32 planes, fewer until the lattice has enough
// workgroups to fill the device in whole rounds: 2,048 for 64 x 8 tiles (two
// per CU: 4 rounds), 4,096 for 64 x 4 and 64 x 2 tiles (at 256^3 the 64 x 4
// D3Q27 kernel, three per CU, ran 29.7 GLUPS with 2,048 workgroups of 32
// planes, 31.0 with 4,096 of 16, 25.9 with 1,024 of 64).
inline int tile_zchunk(int nx, int ny, int nz, int ty) {
    const long tiles = (long)((nx + kTileX - 1) / kTileX) * ((ny + ty - 1) / ty);
    const long want = ty == 8 ? 2048 : 4096;
    int zc = 32;
    while (zc > 4 && tiles * ((nz + zc - 1) / zc) < want) zc /= 2;
    return zc < nz ? zc : nz;
}

}  // namespace vlbm
