// Internal state of a lattice handle and helpers shared by the C-ABI
// translation units (capi.cpp, ibm_capi.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/vlbm.h"
#include "lattice.h"

namespace vlbm {
struct IbmState;
namespace detail {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(VLBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace detail
}  // namespace vlbm

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct vlbm_ctx {
    vlbm_desc d;
    size_t n = 0;             // nodes
    int64_t stride = 0;       // elements between population planes
    int sbytes = 0;           // storage bytes per population (the launchers take d.precision)
    int tbytes = 0;           // compute / macroscopic bytes per value
    void* lat[2] = {nullptr, nullptr};
    int cur = 0;              // lattice holding the current state
    bool direct = true;       // current state is post-stream (after init / load)
    bool mom = false;         // lattices hold moment coefficients (kRepMom)
    int planes = 0;           // planes per lattice: Q, or 10 coefficients
    void* pops = nullptr;     // Q-plane population buffer of a coefficient handle:
                              // loaded post-stream state, or scratch of vlbm_get_pops
    void* scratch_pops = nullptr;   // materialized populations of a rheometer handle (its idle lattice is read),
                                    // or the staged popAux of a loaded checkpoint
    bool aux_pending = false; // after the next (direct) step the idle lattice takes the staged plate slots
    bool direct_pops = false; // the current (direct) state is in pops
    long long steps = 0;
    long long lbm_steps = 0;  // LBM steps since init / load (the IBM stage alone does not count)
    bool resumed = false;     // loaded from a checkpoint taken after at least one step
    uint32_t* d_mask = nullptr;
    std::vector<uint32_t> words;   // host map
    uint32_t region_word[27];
    bool mask_geom = false;
    void* omega_state = nullptr;
    void* visc_turb = nullptr;    // LES eddy viscosity per node (desc.les_visc_turb)
    void* macr[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // rho, ux, uy, uz, omega
    void* uold[3] = {nullptr, nullptr, nullptr};
    double* partials = nullptr;
    double* out3 = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;           // device current at vlbm_create (the stream's device)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t device_bytes = 0;
    vlbm::IbmState* ibm = nullptr;   // immersed boundary stage (vlbm_ibm_enable)
    // what the step kernel needs of the IBM stage (null without IBM)
    const uint32_t* tile_flag = nullptr;
    void* ibm_force[3] = {nullptr, nullptr, nullptr};
    int tiles_x = 0, tiles_y = 0;
};

namespace vlbm {
namespace detail {

int validate_desc(const vlbm_desc* d);                 // capi.cpp
bool box_builder_words(int builder, uint32_t out[27]);  // capi.cpp: node word per box region

// Physics and geometry part of the kernel parameters (no buffers).
template <typename T>
inline void fill_params_desc(const vlbm_desc& d, const uint32_t region_word[27], Params<T>& p) {
    std::memset(&p, 0, sizeof(p));
    p.nx = d.nx; p.ny = d.ny; p.nz = d.nz;
    p.n = (uint32_t)((size_t)d.nx * d.ny * d.nz);
    p.div_nx = make_fastdiv((uint32_t)d.nx);
    p.div_nxy = make_fastdiv((uint32_t)(d.nx * d.ny));
    p.stride = plane_stride(p.n);
    for (int r = 0; r < 27; r++) p.region_word[r] = region_word[r];
    for (int k = 0; k < 8; k++) {
        p.ux_bc[k] = (T)d.ux_bc[k]; p.uy_bc[k] = (T)d.uy_bc[k];
        p.uz_bc[k] = (T)d.uz_bc[k]; p.rho_bc[k] = (T)d.rho_bc[k];
    }
    p.fx = (T)d.fx; p.fy = (T)d.fy; p.fz = (T)d.fz;
    p.omega = (T)(1.0 / d.tau);
    p.omega_p = (T)(1.0 / (3.0 * d.eta_p + 0.5));
    p.s_y = (T)d.s_y;
    p.k_cons = (T)d.k_consistency;
    p.n_minus_1 = (T)(d.n_index - 1.0);
    p.gamma_0 = (T)d.gamma_0;
    p.clamp = d.omega_clamp;
    if (d.omega_clamp) { p.omega_lo = (T)(1.0 / d.tau_max); p.omega_hi = (T)(1.0 / d.tau_min); }
    p.zouhe_textbook = d.zouhe_textbook;
    p.bgk = d.collision == VLBM_COLLISION_BGK ? 1 : 0;
    p.tau = (T)d.tau;
    p.les_c = (T)d.les_const;
    p.duct_R = (T)d.duct_outer_radius; p.duct_r = (T)d.duct_inner_radius;
    p.w_out = (T)d.outer_rotation; p.u_out = (T)d.outer_velocity;
    p.w_in = (T)d.inner_rotation; p.u_in = (T)d.inner_velocity;
    p.plates = d.builder == VLBM_BUILDER_RHEOMETER ? 1 : 0;
    p.sr = (d.precision == VLBM_FP16 || d.precision == VLBM_BF16) && d.rounding == VLBM_ROUND_STOCHASTIC ? 1 : 0;
}

template <typename T>
inline void fill_params(const vlbm_ctx* h, Params<T>& p) {
    fill_params_desc(h->d, h->region_word, p);
    p.stride = h->stride;
    p.held = h->lat[1 - h->cur];
    // step 0 after init reads the post-stream state in place; step 1 (or a
    // pass over the state it reads) holds the first boundary pass, unless the
    // state comes from a checkpoint of a run that has had its first one
    p.first_bc = (h->lbm_steps == 1 && !h->resumed) ? 1 : 0;
    p.mask = h->d_mask;
    p.omega_state = (T*)h->omega_state;
    p.visc_turb = (T*)h->visc_turb;
    p.m_rho = (T*)h->macr[0]; p.m_ux = (T*)h->macr[1]; p.m_uy = (T*)h->macr[2];
    p.m_uz = (T*)h->macr[3]; p.m_omega = (T*)h->macr[4];
    p.tile_flag = h->tile_flag;
    p.ibm_fx = (const T*)h->ibm_force[0]; p.ibm_fy = (const T*)h->ibm_force[1]; p.ibm_fz = (const T*)h->ibm_force[2];
    p.tiles_x = h->tiles_x; p.tiles_y = h->tiles_y;
}

// Runs one kernel with parameters of the handle's compute type.
template <typename F>
inline int with_params(vlbm_ctx* h, F&& fn) {
    if (h->tbytes == 8) {
        Params<double> p;
        fill_params(h, p);
        return fn((void*)&p, p);
    }
    Params<float> p;
    fill_params(h, p);
    return fn((void*)&p, p);
}

inline int check_launch(int rc, const char* what) {
    if (rc != 0) {
        hipError_t e = hipGetLastError();
        return fail(VLBM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
    return VLBM_OK;
}

// immersed boundary stage (ibm_capi.cpp)
void ibm_release(vlbm_ctx* h);
int ibm_stage(vlbm_ctx* h);        // one stage on the current state
int ibm_fill_macr(vlbm_ctx* h);    // macroscopics outside the flagged tiles
int ibm_reset_tiles(vlbm_ctx* h);  // no flagged tiles: every node takes the constant force

// rheometer checkpoints (capi.cpp): the reference's popAux holds each plate
// node's post-collision populations in the slots fed by an unused plane
int plate_aux(vlbm_ctx* h, double* pop_aux);              // pop_aux (pop on entry) gets those slots
int plate_aux_stage(vlbm_ctx* h, const double* pop_aux);  // after vlbm_load_pops

// Representation and buffer of the current state.
inline int state_rep(const vlbm_ctx* h) { return (h->mom && !h->direct_pops) ? 1 : 0; }
inline const void* state_src(const vlbm_ctx* h) { return h->direct_pops ? h->pops : h->lat[h->cur]; }

inline void release(vlbm_ctx* h) {
    if (!h) return;
    ibm_release(h);
    if (h->pops) (void)hipFree(h->pops);
    if (h->visc_turb) (void)hipFree(h->visc_turb);
    if (h->scratch_pops) (void)hipFree(h->scratch_pops);
    for (void* p : h->lat) if (p) (void)hipFree(p);
    for (void* p : h->macr) if (p) (void)hipFree(p);
    for (void* p : h->uold) if (p) (void)hipFree(p);
    if (h->omega_state) (void)hipFree(h->omega_state);
    if (h->d_mask) (void)hipFree(h->d_mask);
    if (h->partials) (void)hipFree(h->partials);
    if (h->out3) (void)hipFree(h->out3);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

inline int dalloc(vlbm_ctx* h, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(VLBM_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    h->device_bytes += bytes;
    return VLBM_OK;
}

// Host copy of a macroscopic field (compute type) as doubles.
inline int fetch_field(vlbm_ctx* h, const void* dev, double* out) {
    if (h->tbytes == 8) {
        HIP_TRY(hipMemcpyAsync(out, dev, h->n * 8, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return VLBM_OK;
    }
    std::vector<float> tmp(h->n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dev, h->n * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < h->n; k++) out[k] = tmp[k];
    return VLBM_OK;
}

inline int upload_field(vlbm_ctx* h, void* dev, const double* in) {
    if (h->tbytes == 8) {
        HIP_TRY(hipMemcpyAsync(dev, in, h->n * 8, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return VLBM_OK;
    }
    std::vector<float> tmp(h->n);
    for (size_t k = 0; k < h->n; k++) tmp[k] = (float)in[k];
    HIP_TRY(hipMemcpyAsync(dev, tmp.data(), h->n * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return VLBM_OK;
}

inline double pop_weight(int q, int i) { return q == 19 ? weight<19>(i) : weight<27>(i); }
}  // namespace detail
}  // namespace vlbm
