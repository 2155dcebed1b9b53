// Immersed boundary stage for rigid spheres: device data layout shared by the
// kernels (ibm_kernels.h) and the host side (capi.cpp).
//
// Replaces src/CUDA/IBM/ibm.cu:6-1197, src/CUDA/IBM/collision/ibmCollision.cu
// and the particle structs of src/CUDA/IBM/structs/. All particle arithmetic
// is FP64; the fluid-side fields (macroscopics, force field, spread force) are
// in the lattice's compute type.
#pragma once
#include <cmath>
#include <cstdint>

#include "lattice.h"

namespace vlbm {

constexpr int kMaxColl = 13;        // history slots: 7 walls + 6 partners (MAX_ACTIVE_COLLISIONS)
constexpr int kFirstPartnerSlot = 7;
constexpr int kMaxContacts = 16;    // sphere-sphere contacts recorded per sphere and step
constexpr int kIbmBlock = 256;
constexpr int kBoxNb = 8;           // overlapping spreading boxes listed per sphere
constexpr int kMaxSortKeys = 4096;  // spheres handled by the one-block cell sort (keys held in LDS)

// One sphere. The first 48 doubles are the C-ABI particle record
// (vlbm_ibm_get_particles) in this order.
struct IbmSphere {
    double pos[3], vel[3], w[3], f[3], M[3], I[6];
    double pos_old[3], vel_old[3], w_old[3], w_avg[3], w_pos[3], f_old[3], M_old[3], dP[3], dL[3];
    double radius, volume, density;
    int movable, first_node, n_nodes, cell;
    int partner[kMaxColl];
    int last[kMaxColl];
    double tang[kMaxColl][3];
    // collision.semiAxis{,2,3} points of the reference's ParticleCenter: moved
    // with the sphere, unused by sphere contacts, kept for the checkpoint file
    double axis[3][3];
};

// Sphere-sphere contact of sphere i (the lower index) with sphere j.
struct IbmContact {
    int j, pad;
    double fd[3];   // force on j; -fd acts on i
    double mi[3];   // torque on i
    double mj[3];   // torque on j
};

// Error bits raised on the device (read back by vlbm_ibm_status).
enum IbmError : int {
    kIbmErrContacts = 1,     // more than kMaxContacts contacts of one sphere
    kIbmErrPartners = 2,     // more than kMaxContacts partners of one sphere found as the higher index
};

struct IbmParams {
    int nx, ny, nz;
    int periodic[3];
    int tiles[3];
    uint32_t n_tiles;
    double thickness, alpha, fluid_density, g[3];
    double pp_kn, pw_kn, pp_kt, pw_kt;     // Hertz / Mindlin constants
    double pp_damp, pw_damp;               // -2 ln e / sqrt(pi^2 + ln^2 e)
    double pp_mu, pw_mu;                   // friction coefficients
    int ref_float;                         // float sqrt / rsqrt / fabs helpers as the reference
    int n_spheres, n_nodes;
    int step;                              // label of the stage (collision history)
    IbmSphere* sph;
    double *px, *py, *pz, *pS;             // Lagrangian nodes: position, area
    double *pfx, *pfy, *pfz;               // accumulated force of the stage
    double *pdx, *pdy, *pdz;               // force of the last iteration
    double *pmx, *pmy, *pmz;               // its moment about the sphere centre
    int* pid;                              // sphere of each node
    uint32_t* tile_flag;                   // 1: the tile is in tile_list
    uint32_t* tile_list;                   // flagged tiles, in no particular order
    uint32_t* tile_count;
    int cells[3];
    double cell_w[3];
    uint64_t* cell_keys;                   // [n_spheres] (cell << 32 | sphere), ascending
    IbmContact* contacts;                  // [n_spheres][kMaxContacts]
    int* n_contacts;
    int* error;
    int box_edge;                          // LDS spreading box edge (0: global atomics)
    int box_plain;                         // box nodes no other box covers may take plain stores
    double box_reach;                      // ceil(r_max) + 1/2: node offsets within it stay in the box
    int* box_nb;                           // [n_spheres][kBoxNb + 1]: count (-1: too many), overlapping boxes
    int* box_stray;                        // 1: some node lies beyond box_reach of its centre (this stage)
    // interpolation work: chunks of at most kChunkNodes nodes of one sphere;
    // the chunks of sphere s are [chunk0[s], chunk0[s + 1])
    int n_chunks;
    int* chunk_sphere;
    int* chunk_first;                      // first node of the chunk
    int* chunk_count;
    int* chunk0;                           // [n_spheres + 1]
    double* chunk_sums;                    // [n_chunks][6]: force, moment of the chunk's nodes
};

constexpr int kChunkNodes = 64;

// Fluid-side fields of the stage on the listed tiles. The velocity the
// reference carries (u plus the auxiliary F / 2 rho of every iteration so
// far) is u + (f - f0) / (2 rho) here: the force field f starts the stage at
// the constant force f0 and gains exactly the spread forces, so only f is
// updated while forces spread.
template <typename T>
struct IbmFluid {
    StageRec<T>* stage;        // rho and u + (f - f0) / 2 rho of the listed nodes
    T *fx, *fy, *fz;           // force field (read by the next step in listed tiles)
    T *out_rho, *out_ux, *out_uy, *out_uz;   // saved macroscopics of the handle
};

// Launchers (kernels_ibm.hip and the per-precision kernel files).
int launch_ibm_macr(int q, int prec, bool mask_geom, const void* params, const void* fluid,
                    const uint32_t* tile_list, const uint32_t* tile_count, uint32_t max_tiles, void* stream,
                    int in_rep = 0);
int launch_ibm_stage(const IbmParams& P, int tbytes, const void* fluid, int stage_part, void* stream);

// Parts of one stage, enqueued in this order (immersedBoundaryMethod, ibm.cu:6-192).
enum IbmStagePart : int {
    kIbmBegin = 0,        // old values, cell index, node force reset, tile list
    kIbmContacts = 1,     // cell sort, contacts, force and torque sums, velocity update
    kIbmIteration = 2,    // interpolation / spreading, sphere sums, velocity update, Euler sum
    kIbmMove = 3,         // sphere and node movement
    kIbmOutput = 4,       // saved macroscopics on the listed tiles (not part of a stage)
};

// Edge of the LDS box that collects one sphere's spread force (FP32 fluids):
// the 4^3 stencils of nodes within the radius of the centre fit in
// 2 ceil(r) + 6 nodes per axis.
inline int ibm_box_edge(double r_max) { return 2 * (int)std::ceil(r_max) + 6; }

}  // namespace vlbm
