// C-ABI of the MI355X LBM library: lattice handles, boundary-map builders,
// initialisation, stepping, reductions and the reference's output layout.
// Declarations and the reference symbol each entry point replaces are in
// include/vlbm.h.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/vlbm.h"
#include "lattice.h"
#include "ctx.h"

using namespace vlbm;
using namespace vlbm::detail;

namespace vlbm {
namespace detail {
thread_local std::string g_err;
}  // namespace detail
}  // namespace vlbm

namespace {

// (error state and helpers: ctx.h)

// ---------------------------------------------------------------------------
// Boundary-map builders.
//
// Box builders are described by what sits on each wall face; the node word of
// a region (corner / edge / face / interior, classified in the reference's
// if-chain priority) follows from the faces of that region. Duct builders
// evaluate the circle geometry per node.
// ---------------------------------------------------------------------------
constexpr unsigned kRegionFaces[27] = {
    fS | fW | fB, fS | fW | fF, fS | fE | fB, fS | fE | fF, fN | fW | fB, fN | fW | fF, fN | fE | fB, fN | fE | fF,
    fS | fW, fS | fE, fN | fW, fN | fE,
    fS | fB, fS | fF, fN | fB, fN | fF, fW | fB, fW | fF, fE | fB, fE | fF,
    fS, fN, fW, fE, fB, fF, 0};

int host_region(int x, int y, int z, int nx, int ny, int nz) {
    const bool s = (y == 0), n = (y == ny - 1), w = (x == 0), e = (x == nx - 1);
    const bool b = (z == 0), f = (z == nz - 1);
    const bool c[27] = {s && w && b, s && w && f, s && e && b, s && e && f, n && w && b, n && w && f, n && e && b,
                        n && e && f, s && w, s && e, n && w, n && e, s && b, s && f, n && b, n && f, w && b, w && f,
                        e && b, e && f, s, n, w, e, b, f, true};
    for (int r = 0; r < 27; r++)
        if (c[r]) return r;
    return 26;
}

int dir_of_faces(unsigned faces) {
    for (int d = 0; d < 26; d++)
        if (kDirFaces[d] == faces) return d;
    return -1;
}

uint32_t plain_word() { return make_word(true, false, kNull, 0, 0, 0, 0, 0, 0); }
uint32_t bc_word(int scheme, int dir, int uxi = 0, int uyi = 0, int uzi = 0, int rhoi = 0) {
    return make_word(true, false, scheme, dir, 0, uxi, uyi, uzi, rhoi);
}

bool is_box_builder(int b) { return b >= VLBM_BUILDER_TAYLOR_GREEN && b <= VLBM_BUILDER_SQUARE_DUCT_ZOUHE; }
bool is_duct_builder(int b) {
    return b == VLBM_BUILDER_CIRCULAR_DUCT_IBB || b == VLBM_BUILDER_ANNULAR_DUCT_IBB || b == VLBM_BUILDER_RHEOMETER;
}

// Node word of each box region for a builder.
bool box_region_words(int builder, uint32_t out[27]) {
    for (int r = 0; r < 27; r++) {
        const unsigned rf = kRegionFaces[r];
        uint32_t w = plain_word();
        auto walls = [&](unsigned wallset) -> bool {
            const unsigned hit = rf & wallset;
            if (!hit) return false;
            w = bc_word(kBounceBack, dir_of_faces(hit));
            return true;
        };
        switch (builder) {
        case VLBM_BUILDER_TAYLOR_GREEN: break;
        case VLBM_BUILDER_PLATES_BB_Y: walls(fN | fS); break;
        case VLBM_BUILDER_PLATES_BB_X: walls(fW | fE); break;
        case VLBM_BUILDER_PLATES_BB_Z: walls(fB | fF); break;
        case VLBM_BUILDER_SQUARE_DUCT_BB: walls(fN | fS | fW | fE); break;
        case VLBM_BUILDER_CLOSED_BOX_BB: walls(fN | fS | fW | fE | fF | fB); break;
        case VLBM_BUILDER_SQUARE_DUCT_ZOUHE:
            if (!walls(fN | fS | fW | fE)) {
                if (rf & fB) w = bc_word(kVelZouHe, dB, 0, 0, 1, 0);
                else if (rf & fF) w = bc_word(kPresZouHe, dF, 0, 0, 0, 0);
            }
            break;
        case VLBM_BUILDER_PLATES_ZOUHE:
            if (rf & fS) w = bc_word(kVelZouHe, dS);
            else if (rf & fN) w = bc_word(kVelZouHe, dN);
            else if (rf & fB) w = bc_word(kPresZouHe, dB, 0, 0, 0, 1);
            else if (rf & fF) w = bc_word(kPresZouHe, dF, 0, 0, 0, 2);
            break;
        case VLBM_BUILDER_LID_DRIVEN_BB:
            // moving lid on N (ux index 1); lid edges and corners stay plain bounce-back
            if ((rf & (fN | fS | fW | fE)) == fN) w = bc_word(kVelBB, dN, 1, 0, 0, 0);
            else walls(fN | fS | fW | fE);
            break;
        case VLBM_BUILDER_COUETTE_ZOUHE:
            if (rf & fS) w = bc_word(kVelZouHe, dS, 0, 0, 2, 0);
            else if (rf & fN) w = bc_word(kVelZouHe, dN, 0, 0, 1, 0);
            break;
        case VLBM_BUILDER_INFLOW_FREE_SLIP:
            if (rf & fS) w = bc_word(kFreeSlip, dS);
            else if (rf & fN) w = bc_word(kFreeSlip, dN);
            else if (rf & fB) w = bc_word(kVelBB, dB, 0, 0, 1, 0);
            else if (rf & fF) w = bc_word(kVelBB, dF, 0, 0, 1, 0);
            break;
        case VLBM_BUILDER_PLATES_BB_FREE_SLIP:
            if (rf & fS) w = bc_word(kBounceBack, dS);
            else if (rf & fN) w = bc_word(kFreeSlip, dN);
            break;
        default: return false;
        }
        out[r] = w;
    }
    return true;
}

// distPoints2D of the reference: only the first difference is cast to float
// (src/CUDA/globalFunctions.h:158-162). R is the build's dfloat.
template <typename R>
R ref_dist2d(R x1, R y1, R x2, R y2) {
    return (R)std::sqrt((float)(x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
}

// Circular / annular duct words (circularDuctInterpBounceBack.cu:35-121,
// annularDuctInterpBounceBack.cu): nodes outside the wall are unused, nodes
// within sqrt(2) of a wall are interpolated-bounce-back nodes with the D2Q9
// set of populations whose source lies outside, and the next ring keeps its
// post-collision populations.
template <typename R>
void duct_words(const vlbm_desc& d, std::vector<uint32_t>& out) {
    const bool annular = d.builder == VLBM_BUILDER_ANNULAR_DUCT_IBB;
    const R Ro = (R)d.duct_outer_radius, Ri = (R)d.duct_inner_radius;
    const R xc = (R)(((d.nx - 1) / 2.0) + 0.5), yc = (R)(((d.ny - 1) / 2.0) + 0.5);
    const float s2 = std::sqrt((float)2);
    const int dxs[8] = {1, 0, -1, 0, 1, -1, -1, 1}, dys[8] = {0, 1, 0, -1, 1, 1, -1, -1};
    std::vector<uint32_t> plane((size_t)d.nx * d.ny);
    for (int y = 0; y < d.ny; y++)
        for (int x = 0; x < d.nx; x++) {
            const R xn = (R)(x + 0.5), yn = (R)(y + 0.5);
            const R dist = ref_dist2d<R>(xn, yn, xc, yc);
            bool used = true, save = false;
            int scheme = kNull, dir = annular ? dF : dN;
            bool mark_unknowns = annular;
            auto outside = [&](R dd) { return annular ? (dd > Ro || dd < Ri) : (dd > Ro); };
            if (outside(dist)) {
                used = false;
            } else if (dist > (Ro - s2)) {
                scheme = kInterpBB; dir = dN; mark_unknowns = true;
            } else if (dist > (Ro - 2 * s2)) {
                save = true;
                mark_unknowns = true;   // the reference's unknown-population loop runs for this ring too
            } else if (annular && dist < Ri + s2) {
                scheme = kInterpBB; dir = dS;
            } else if (annular && dist < Ri + 2 * s2) {
                save = true;
            }
            uint32_t w = make_word(used, save, scheme, dir, 0, 0, 0, 0, 0);
            if (mark_unknowns) {
                for (int i = 0; i < 8; i++) {
                    const R xa = xn - dxs[i], ya = yn - dys[i];
                    bool unknown;
                    if (xa < d.nx && xa > 0 && ya < d.ny && ya > 0) unknown = outside(ref_dist2d<R>(xa, ya, xc, yc));
                    else unknown = true;
                    if (unknown) w |= (1u << i) << kOffInterp;
                }
            }
            plane[(size_t)y * d.nx + x] = w;
        }
    out.resize((size_t)d.nx * d.ny * d.nz);
    for (int z = 0; z < d.nz; z++) std::memcpy(&out[(size_t)z * d.nx * d.ny], plane.data(), plane.size() * 4);
    if (d.builder != VLBM_BUILDER_RHEOMETER) return;
    // parallelPlatesRheometer.cu:70-89: planes z = 0 and NZ-1 unused; on z = 1
    // (BACK) and NZ-2 (FRONT) gpuSchSpecial, unless the wall classification
    // makes the node an interpolated bounce-back node
    const size_t np = (size_t)d.nx * d.ny;
    for (size_t k = 0; k < np; k++) {
        out[k] = 0;
        out[(size_t)(d.nz - 1) * np + k] = 0;
        for (int z : {1, d.nz - 2}) {
            uint32_t& w = out[(size_t)z * np + k];
            if (!w_used(w) || w_scheme(w) == kInterpBB || z <= 0 || z >= d.nz - 1) continue;
            const int dir = z == 1 ? dB : dF;
            w = (w & ~((7u << kOffScheme) | (31u << kOffDir))) | ((uint32_t)kSpecial << kOffScheme) | ((uint32_t)dir << kOffDir);
        }
    }
}

int validate(const vlbm_desc* d) {
    if (!d) return fail(VLBM_ERR_INVALID, "descriptor is NULL");
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(VLBM_ERR_INVALID, "NX, NY and NZ must be >= 1");
    const double n = (double)d->nx * d->ny * d->nz;
    if (n >= 2147483648.0) return fail(VLBM_ERR_INVALID, "lattice has 2^31 nodes or more");
    if (d->q != 19 && d->q != 27) return fail(VLBM_ERR_INVALID, "velocity set must be D3Q19 or D3Q27");
    if (d->precision != VLBM_FP64 && d->precision != VLBM_FP32 && d->precision != VLBM_FP16 &&
        d->precision != VLBM_BF16)
        return fail(VLBM_ERR_INVALID, "precision must be VLBM_FP64, VLBM_FP32, VLBM_FP16 or VLBM_BF16");
    if (d->rounding != VLBM_ROUND_STOCHASTIC && d->rounding != VLBM_ROUND_NEAREST)
        return fail(VLBM_ERR_INVALID, "rounding must be VLBM_ROUND_STOCHASTIC or VLBM_ROUND_NEAREST");
    if (d->rheology < 0 || d->rheology > 2) return fail(VLBM_ERR_INVALID, "unknown rheology");
    if (!is_box_builder(d->builder) && !is_duct_builder(d->builder))
        return fail(VLBM_ERR_INVALID, "unknown boundary-condition builder");
    if (!(d->tau > 0.5) && d->rheology == VLBM_NEWTONIAN) return fail(VLBM_ERR_INVALID, "tau must exceed 0.5");
    if (d->rheology == VLBM_BINGHAM && !(d->eta_p > 0)) return fail(VLBM_ERR_INVALID, "eta_p must be > 0");
    if (d->rheology == VLBM_POWERLAW && !(d->k_consistency > 0)) return fail(VLBM_ERR_INVALID, "K must be > 0");
    if (d->omega_clamp && !(d->tau_min > 0.5 && d->tau_max >= d->tau_min))
        return fail(VLBM_ERR_INVALID, "tau clamp needs 0.5 < tau_min <= tau_max");
    if (d->collision != VLBM_COLLISION_REGULARIZED && d->collision != VLBM_COLLISION_BGK)
        return fail(VLBM_ERR_INVALID, "collision must be VLBM_COLLISION_REGULARIZED or VLBM_COLLISION_BGK");
    if (!(d->les_const >= 0)) return fail(VLBM_ERR_INVALID, "LES constant must be >= 0");
    if (d->les_const > 0 && d->rheology != VLBM_NEWTONIAN)
        return fail(VLBM_ERR_UNSUPPORTED, "LES applies to Newtonian fluids only");
    if (d->les_const > 0 && !(d->tau > 0.5)) return fail(VLBM_ERR_INVALID, "LES needs tau > 0.5");
    if (d->les_visc_turb && !(d->les_const > 0))
        return fail(VLBM_ERR_INVALID, "the eddy viscosity export needs LES (les_const > 0)");
    if (d->layout < VLBM_LAYOUT_AUTO || d->layout > VLBM_LAYOUT_MOMENTS) return fail(VLBM_ERR_INVALID, "unknown layout");
    if (d->layout == VLBM_LAYOUT_MOMENTS && (d->collision != VLBM_COLLISION_REGULARIZED || d->precision == VLBM_FP16 ||
                                             d->precision == VLBM_BF16))
        return fail(VLBM_ERR_UNSUPPORTED, "the moment layout holds the regularized state in float or double storage");
    if (d->q == 27 && d->builder != VLBM_BUILDER_TAYLOR_GREEN) {
        // Zou-He and velocity bounce-back exist for D3Q19 only in the reference
        // (boundaryConditionsHandler.cu:274,318,361); refuse instead of silently skipping.
        uint32_t words[27];
        if (is_box_builder(d->builder) && box_region_words(d->builder, words)) {
            for (int r = 0; r < 27; r++) {
                const int s = w_scheme(words[r]);
                if (s == kVelZouHe || s == kPresZouHe || s == kVelBB)
                    return fail(VLBM_ERR_UNSUPPORTED, "Zou-He and velocity bounce-back are D3Q19-only");
            }
        }
    }
    if (d->builder == VLBM_BUILDER_RHEOMETER) {
        if (d->q != 19) return fail(VLBM_ERR_UNSUPPORTED, "the rheometer's special boundary is D3Q19-only");
        if (d->layout == VLBM_LAYOUT_MOMENTS)
            return fail(VLBM_ERR_UNSUPPORTED, "the rheometer needs the population layout");
        if (d->nz < 3) return fail(VLBM_ERR_INVALID, "the rheometer needs NZ >= 3");
    }
    if (is_duct_builder(d->builder)) {
        if (!(d->duct_outer_radius > 0)) return fail(VLBM_ERR_INVALID, "duct radius must be > 0");
        if (d->builder == VLBM_BUILDER_ANNULAR_DUCT_IBB && !(d->duct_inner_radius > 0 && d->duct_inner_radius < d->duct_outer_radius))
            return fail(VLBM_ERR_INVALID, "annular duct needs 0 < inner radius < outer radius");
    }
    return VLBM_OK;
}

int build_words(const vlbm_desc& d, std::vector<uint32_t>& out) {
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    if (is_duct_builder(d.builder)) {
        if (d.precision == VLBM_FP64) duct_words<double>(d, out);
        else duct_words<float>(d, out);
        return VLBM_OK;
    }
    uint32_t words[27];
    if (!box_region_words(d.builder, words)) return fail(VLBM_ERR_INVALID, "unknown builder");
    out.resize(n);
    size_t k = 0;
    for (int z = 0; z < d.nz; z++)
        for (int y = 0; y < d.ny; y++)
            for (int x = 0; x < d.nx; x++) out[k++] = words[host_region(x, y, z, d.nx, d.ny, d.nz)];
    return VLBM_OK;
}

}  // namespace

namespace vlbm {
namespace detail {

int validate_desc(const vlbm_desc* d) { return validate(d); }

bool box_builder_words(int builder, uint32_t out[27]) {
    return is_box_builder(builder) && box_region_words(builder, out);
}

}  // namespace detail
}  // namespace vlbm

extern "C" {

const char* vlbm_last_error(void) { return g_err.c_str(); }

void vlbm_desc_defaults(vlbm_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    // src/CUDA/var.h:97-117 and src/CUDA/NNF/nnf.h:10-24
    d->nx = 128; d->ny = 128; d->nz = 384;
    d->q = 19;
    d->precision = VLBM_FP32;
    d->rheology = VLBM_NEWTONIAN;
    d->builder = VLBM_BUILDER_CLOSED_BOX_BB;
    d->tau = 0.6;
    d->rho0 = 1.0;
    d->eta_p = 0.3; d->s_y = 1e-4;
    d->n_index = 0.5; d->k_consistency = 1.0 * (0.6 - 0.5) / 3; d->gamma_0 = 0;
    d->tau_min = 0.51; d->tau_max = 2.0;
    const double umax = 0.1;
    const double ux[8] = {0, umax, 0, 0, 0, 0, 0, 0};
    const double uy[8] = {0, umax / 2, -umax / 2, 0, 0, 0, 0, 0};
    const double uz[8] = {0, umax, -umax, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) {
        d->ux_bc[k] = ux[k]; d->uy_bc[k] = uy[k]; d->uz_bc[k] = uz[k]; d->rho_bc[k] = 1.0;
    }
    d->duct_outer_radius = d->ny / 2.0 - 1.5;
    d->duct_inner_radius = d->duct_outer_radius / 4.0;
}

int vlbm_build_map(const vlbm_desc* d, uint32_t* out) {
    int rc = validate(d);
    if (rc) return rc;
    if (!out) return fail(VLBM_ERR_INVALID, "output is NULL");
    std::vector<uint32_t> w;
    rc = build_words(*d, w);
    if (rc) return rc;
    std::memcpy(out, w.data(), w.size() * 4);
    return VLBM_OK;
}

int vlbm_create(const vlbm_desc* d, vlbm_h* out) {
    if (!out) return fail(VLBM_ERR_INVALID, "output handle pointer is NULL");
    *out = nullptr;
    int rc = validate(d);
    if (rc) return rc;
    // the rheometer's plate velocity divides by the node's distance from the
    // axis (parallelPlatesRheometer.cu:153-158), which is 0 on the axis node
    // of an odd x odd cross-section: the reference's state turns NaN there
    // (its node map is still well defined, so vlbm_build_map accepts it)
    if (d->builder == VLBM_BUILDER_RHEOMETER && d->nx % 2 == 1 && d->ny % 2 == 1)
        return fail(VLBM_ERR_INVALID, "the rheometer needs an even NX or NY (odd x odd puts a plate node on the "
                                      "axis, where the wall velocity is 0/0)");
    vlbm_ctx* h = new (std::nothrow) vlbm_ctx;
    if (!h) return fail(VLBM_ERR_ALLOC, "host allocation failed");
    h->d = *d;
    h->n = (size_t)d->nx * d->ny * d->nz;
    h->stride = plane_stride(h->n);
    h->sbytes = d->precision == VLBM_BF16 ? 2 : d->precision;
    h->tbytes = d->precision == VLBM_FP64 ? 8 : 4;
    rc = build_words(*d, h->words);
    if (rc) { release(h); return rc; }
    h->mask_geom = is_duct_builder(d->builder);
    if (!h->mask_geom) box_region_words(d->builder, h->region_word);
    else for (int r = 0; r < 27; r++) h->region_word[r] = 0;

    if (hipGetDevice(&h->device) != hipSuccess) { release(h); return fail(VLBM_ERR_DEVICE, "hipGetDevice failed"); }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { release(h); return fail(VLBM_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        release(h);
        return fail(VLBM_ERR_DEVICE, "hipEventCreate failed");
    }
    // AUTO: coefficients in float storage, where they were measured faster on
    // MI355X (GLUPS, coefficients vs populations, 256^3 unless stated): D3Q19
    // box maps (Bingham duct 42.1 vs 38.2, Bingham plates 44.4 vs 39.4,
    // power-law plates 37.7 vs 37.1), D3Q19 ducts with the node map
    // (circular 32.6 vs 30.8, power-law circular 28.5 vs 27.9, annular 32.0
    // vs 29.1; circular 128^3 25.5 vs 22.6, 64 x 64 x 512 22.5 vs 16.5) and
    // D3Q27 (power-law plates 31.0 vs 24.8, circular duct 29.2 vs 21.6). FP64
    // runs faster with populations, and the rheometer's plates hold
    // populations. The tile kernel needs about 2,048 workgroups (64 x 8 or
    // 64 x 4 tiles, z-chunks of at least 4 planes) to fill the device;
    // smaller lattices keep populations (one thread per node).
    const int tile_rows = (d->q == 19 && !h->mask_geom) ? 8 : 4;
    const long tiles = (long)((d->nx + 63) / 64) * ((d->ny + tile_rows - 1) / tile_rows) * ((d->nz + 3) / 4);
    const bool mom_auto = d->collision == VLBM_COLLISION_REGULARIZED && d->precision == VLBM_FP32 && d->nx >= 64 &&
                          d->builder != VLBM_BUILDER_RHEOMETER && tiles >= 2048;
    h->mom = d->layout == VLBM_LAYOUT_MOMENTS || (d->layout == VLBM_LAYOUT_AUTO && mom_auto);
    h->planes = h->mom ? kCoef : d->q;
    const size_t lat_bytes = (size_t)h->planes * (size_t)h->stride * h->sbytes;
    for (int k = 0; k < 2; k++) {
        if ((rc = dalloc(h, &h->lat[k], lat_bytes))) { release(h); return rc; }
        if (hipMemsetAsync(h->lat[k], 0, lat_bytes, h->stream) != hipSuccess) { release(h); return fail(VLBM_ERR_DEVICE, "hipMemset"); }
    }
    for (int k = 0; k < 5; k++) {
        if (k == 4 && d->rheology == VLBM_NEWTONIAN) continue;
        if ((rc = dalloc(h, &h->macr[k], h->n * h->tbytes))) { release(h); return rc; }
    }
    for (int k = 0; k < 3; k++)
        if ((rc = dalloc(h, &h->uold[k], h->n * h->tbytes))) { release(h); return rc; }
    if (d->rheology == VLBM_POWERLAW)
        if ((rc = dalloc(h, &h->omega_state, h->n * h->tbytes))) { release(h); return rc; }
    if (d->les_visc_turb) {
        // INIT_VISC_TURB = 0 (LES/lesVar.h:21, lbmInitialization.cu:114-115)
        if ((rc = dalloc(h, &h->visc_turb, h->n * h->tbytes))) { release(h); return rc; }
        if (hipMemsetAsync(h->visc_turb, 0, h->n * h->tbytes, h->stream) != hipSuccess) {
            release(h);
            return fail(VLBM_ERR_DEVICE, "hipMemset");
        }
    }
    if (h->mask_geom) {
        if ((rc = dalloc(h, (void**)&h->d_mask, h->n * 4))) { release(h); return rc; }
        e = hipMemcpy(h->d_mask, h->words.data(), h->n * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) { release(h); return fail(VLBM_ERR_DEVICE, "mask upload failed"); }
    }
    if ((rc = dalloc(h, (void**)&h->partials, 1024 * 3 * sizeof(double)))) { release(h); return rc; }
    if ((rc = dalloc(h, (void**)&h->out3, 3 * sizeof(double)))) { release(h); return rc; }
    if (hipStreamSynchronize(h->stream) != hipSuccess) { release(h); return fail(VLBM_ERR_DEVICE, "stream sync failed"); }
    rc = vlbm_init(h, nullptr, nullptr, nullptr, nullptr);
    if (rc) { release(h); return rc; }
    *out = h;
    return VLBM_OK;
}

void vlbm_destroy(vlbm_h h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    release(h);
}

int vlbm_set_map(vlbm_h h, const uint32_t* words) {
    if (!h || !words) return fail(VLBM_ERR_INVALID, "handle or words NULL");
    const vlbm_desc& d = h->d;
    if (d.builder == VLBM_BUILDER_RHEOMETER)
        return fail(VLBM_ERR_UNSUPPORTED, "the rheometer's map is fixed by its plates");
    if (h->ibm) return fail(VLBM_ERR_UNSUPPORTED, "set the node map before enabling the immersed boundary stage");
    for (size_t k = 0; k < h->n; k++) {
        const uint32_t w = words[k];
        const int s = w_scheme(w), dir = w_dir(w);
        if (dir > dSEB) return fail(VLBM_ERR_INVALID, "node " + std::to_string(k) + ": direction code beyond 25");
        if (s == kSpecial)
            return fail(VLBM_ERR_UNSUPPORTED, "node " + std::to_string(k) + ": the special scheme is the rheometer's");
        if ((s == kVelZouHe || s == kPresZouHe) && dir > dB)
            return fail(VLBM_ERR_INVALID, "node " + std::to_string(k) + ": Zou-He is defined on the six faces only");
        if (d.q == 27 && (s == kVelZouHe || s == kPresZouHe || s == kVelBB))
            return fail(VLBM_ERR_UNSUPPORTED, "Zou-He and velocity bounce-back are D3Q19-only");
        if (s == kInterpBB && !(d.duct_outer_radius > 0))
            return fail(VLBM_ERR_INVALID, "interpolated bounce-back needs the duct radius");
    }
    if (!h->d_mask) {
        int rc = dalloc(h, (void**)&h->d_mask, h->n * 4);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(h->d_mask, words, h->n * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->words.assign(words, words + h->n);
    h->mask_geom = true;
    return VLBM_OK;
}

int vlbm_bc_nodes(vlbm_h h, uint64_t* out, size_t cap, size_t* count) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    size_t k = 0;
    for (size_t i = 0; i < h->n; i++) {
        const uint32_t w = h->words[i];
        if (w_used(w) && w_scheme(w) != kNull) {
            if (out && k < cap) out[k] = i;
            k++;
        }
    }
    if (count) *count = k;
    if (out && k > cap) return fail(VLBM_ERR_INVALID, "output capacity too small");
    return VLBM_OK;
}

int vlbm_init(vlbm_h h, const double* rho, const double* ux, const double* uy, const double* uz) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    const size_t n = h->n;
    // macroscopic fields as doubles on the device for the equilibrium kernel
    std::vector<double> r(n), vx(n), vy(n), vz(n);
    for (size_t k = 0; k < n; k++) {
        r[k] = rho ? rho[k] : h->d.rho0;
        vx[k] = ux ? ux[k] : 0.0;
        vy[k] = uy ? uy[k] : 0.0;
        vz[k] = uz ? uz[k] : 0.0;
    }
    double* dev = nullptr;
    HIP_TRY(hipMalloc((void**)&dev, 4 * n * sizeof(double)));
    int rc = VLBM_OK;
    do {
        if (hipMemcpyAsync(dev, r.data(), n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(dev + n, vx.data(), n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(dev + 2 * n, vy.data(), n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(dev + 3 * n, vz.data(), n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            rc = fail(VLBM_ERR_DEVICE, "upload of initial fields failed");
            break;
        }
        h->cur = 0;
        // both lattices: slots that no step writes (the rheometer's unused
        // planes) hold the initial state, as both of the reference's buffers do
        bool failed = false;
        for (int k = 0; k < 2 && !failed; k++)
            failed = launch_fill_feq(h->d.q, h->d.precision, h->lat[k], h->stride, (uint32_t)n, dev, dev + n, dev + 2 * n,
                                     dev + 3 * n, h->stream, h->mom ? kRepMom : kRepPop) != 0;
        if (failed) {
            rc = check_launch(-1, "equilibrium kernel");
            break;
        }
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = fail(VLBM_ERR_DEVICE, "init sync failed"); break; }
    } while (0);
    (void)hipFree(dev);
    if (rc) return rc;
    // saved macroscopics and the residual reference start at the initial fields
    if ((rc = upload_field(h, h->macr[0], r.data()))) return rc;
    if ((rc = upload_field(h, h->macr[1], vx.data()))) return rc;
    if ((rc = upload_field(h, h->macr[2], vy.data()))) return rc;
    if ((rc = upload_field(h, h->macr[3], vz.data()))) return rc;
    if ((rc = upload_field(h, h->uold[0], vx.data()))) return rc;
    if ((rc = upload_field(h, h->uold[1], vy.data()))) return rc;
    if ((rc = upload_field(h, h->uold[2], vz.data()))) return rc;
    std::vector<double> zero(n, 0.0);
    if (h->macr[4] && (rc = upload_field(h, h->macr[4], zero.data()))) return rc;
    // power-law omega starts at 0 (src/CUDA/lbmInitialization.cu:110-112)
    if (h->omega_state && (rc = upload_field(h, h->omega_state, zero.data()))) return rc;
    h->direct = true;
    h->direct_pops = false;
    h->aux_pending = false;
    h->resumed = false;
    h->steps = 0;
    h->lbm_steps = 0;
    return ibm_reset_tiles(h);
}

// Q-plane population buffer of a coefficient handle, allocated on first use.
static int need_pops(vlbm_h h) {
    if (h->pops) return VLBM_OK;
    return dalloc(h, &h->pops, (size_t)h->d.q * (size_t)h->stride * h->sbytes);
}

int vlbm_set_omega(vlbm_h h, const double* omega) {
    if (!h || !omega) return fail(VLBM_ERR_INVALID, "handle or omega is NULL");
    if (!h->omega_state) return fail(VLBM_ERR_INVALID, "omega state exists for the power law only");
    return upload_field(h, h->omega_state, omega);
}

// Q population planes of the state (plane stride h->stride) <-> Q packed
// planes of n doubles.
static int convert_planes(vlbm_h h, void* state, double* dbl, int to_double) {
    for (int i = 0; i < h->d.q; i++)
        if (launch_convert(h->d.precision, (char*)state + (size_t)i * (size_t)h->stride * h->sbytes, dbl + (size_t)i * h->n,
                           h->n, to_double, h->stream))
            return -1;
    return 0;
}

// (pop, z) of the rheometer's plate slots whose source is an unused plane.
static std::vector<std::pair<int, int>> plate_slots(const vlbm_desc& d) {
    std::vector<std::pair<int, int>> out;
    if (d.builder != VLBM_BUILDER_RHEOMETER) return out;
    for (int i = 1; i < d.q; i++)
        for (int z = 1; z <= d.nz - 2; z += (d.nz - 3 > 0 ? d.nz - 3 : 1)) {
            const int zs = z - kCZ[i];
            if (kCZ[i] != 0 && (zs == 0 || zs == d.nz - 1)) out.push_back({i, z});
        }
    return out;
}

static int upload_ref_pops(vlbm_h h, const double* pop, void* target);

}  // extern "C"

namespace vlbm {
namespace detail {

int plate_aux(vlbm_ctx* h, double* pop_aux) {
    const vlbm_desc& d = h->d;
    const auto slots = plate_slots(d);
    // right after init both of the reference's buffers hold the initial state;
    // right after a checkpoint load the staged popAux still holds the slots
    if (slots.empty() || (h->direct && !h->aux_pending)) return VLBM_OK;
    void* source = h->aux_pending ? h->scratch_pops : h->lat[h->cur];
    const size_t plane = (size_t)d.nx * d.ny, total = (size_t)d.q * h->n;
    double* dbl = nullptr;
    HIP_TRY(hipMalloc((void**)&dbl, total * 8));
    std::vector<double> host(total);
    int rc = VLBM_OK;
    if (convert_planes(h, source, dbl, 1)) rc = check_launch(-1, "conversion kernel");
    else if (hipMemcpyAsync(host.data(), dbl, total * 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
             hipStreamSynchronize(h->stream) != hipSuccess)
        rc = fail(VLBM_ERR_DEVICE, "population download failed");
    (void)hipFree(dbl);
    if (rc) return rc;
    for (const auto& s : slots) {
        const double w = pop_weight(d.q, s.first);
        double* dst = pop_aux + (size_t)s.first * plane * (d.nz + 1) + (size_t)s.second * plane;
        const double* src = host.data() + (size_t)s.first * h->n + (size_t)s.second * plane;
        for (size_t k = 0; k < plane; k++) dst[k] = src[k] + w;
    }
    return VLBM_OK;
}

int plate_aux_stage(vlbm_ctx* h, const double* pop_aux) {
    const vlbm_desc& d = h->d;
    if (plate_slots(d).empty()) return VLBM_OK;
    if (!h->scratch_pops) HIP_TRY(hipMalloc(&h->scratch_pops, (size_t)d.q * h->stride * h->sbytes));
    int rc = upload_ref_pops(h, pop_aux, h->scratch_pops);
    if (rc) return rc;
    h->aux_pending = true;
    return VLBM_OK;
}

}  // namespace detail
}  // namespace vlbm

extern "C" {

// Reference-layout populations (Q, NZ+1, NY, NX) into Q deviation planes of target.
static int upload_ref_pops(vlbm_h h, const double* pop, void* target) {
    const vlbm_desc& d = h->d;
    const size_t plane = (size_t)d.nx * d.ny;
    std::vector<double> dev_in((size_t)d.q * h->n);
    for (int i = 0; i < d.q; i++) {
        const double w = pop_weight(d.q, i);
        const double* srcp = pop + (size_t)i * plane * (d.nz + 1);   // skip the ghost plane
        double* dstp = dev_in.data() + (size_t)i * h->n;
        for (size_t k = 0; k < h->n; k++) dstp[k] = srcp[k] - w;
    }
    double* dbl = nullptr;
    HIP_TRY(hipMalloc((void**)&dbl, dev_in.size() * 8));
    int rc = VLBM_OK;
    if (hipMemcpyAsync(dbl, dev_in.data(), dev_in.size() * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = fail(VLBM_ERR_DEVICE, "population upload failed");
    else if (convert_planes(h, target, dbl, 0))
        rc = check_launch(-1, "conversion kernel");
    else if (hipStreamSynchronize(h->stream) != hipSuccess)
        rc = fail(VLBM_ERR_DEVICE, "sync failed");
    (void)hipFree(dbl);
    return rc;
}

int vlbm_load_pops(vlbm_h h, const double* pop) {
    if (!h || !pop) return fail(VLBM_ERR_INVALID, "handle or populations NULL");
    const vlbm_desc& d = h->d;
    h->aux_pending = false;
    h->resumed = false;
    // a coefficient handle keeps the loaded populations until the first step
    int rc = h->mom ? need_pops(h) : VLBM_OK;
    if (rc) return rc;
    if ((rc = upload_ref_pops(h, pop, h->mom ? h->pops : h->lat[0]))) return rc;
    // nodes that never collide (outside duct walls, the rheometer's unused
    // planes) keep the loaded values in both lattices, as in both of the
    // reference's buffers; the rheometer also reads the idle lattice (held
    // values) in the first steps
    if (!h->mom && h->mask_geom && (rc = upload_ref_pops(h, pop, h->lat[1]))) return rc;
    h->cur = 0;
    h->direct = true;
    h->direct_pops = h->mom;
    h->steps = 0;
    h->lbm_steps = 0;
    return ibm_reset_tiles(h);
}

int vlbm_get_pops(vlbm_h h, double* pop) {
    if (!h || !pop) return fail(VLBM_ERR_INVALID, "handle or output NULL");
    const vlbm_desc& d = h->d;
    const void* source = state_src(h);
    if (!h->direct || (h->mom && !h->direct_pops)) {
        // post-stream + BC populations of the current state into the idle
        // lattice (population layout) or the population buffer (coefficients)
        int rc = h->mom ? need_pops(h) : VLBM_OK;
        if (rc) return rc;
        void* target = h->mom ? h->pops : h->lat[1 - h->cur];
        // the rheometer's plates read the idle lattice in the next step: keep it
        if (d.builder == VLBM_BUILDER_RHEOMETER) {
            if (!h->scratch_pops) HIP_TRY(hipMalloc(&h->scratch_pops, (size_t)d.q * h->stride * h->sbytes));
            target = h->scratch_pops;
        }
        rc = with_params(h, [&](void* pv, auto& p) {
            p.src = h->lat[h->cur];
            p.dst = target;
            p.aux_mode = kAuxMaterialize;
            p.direct = h->direct ? 1 : 0;
            return check_launch(launch_aux(d.q, d.precision, d.rheology, h->mask_geom, pv, h->stream, state_rep(h)),
                                "materialize kernel");
        });
        if (rc) return rc;
        source = target;
    }
    const size_t total = (size_t)d.q * h->n;
    const size_t plane = (size_t)d.nx * d.ny;
    const size_t ghost = (size_t)d.q * plane;
    double* dbl = nullptr;
    HIP_TRY(hipMalloc((void**)&dbl, (total + ghost) * 8));
    std::vector<double> host(total + ghost, 0.0);
    int rc = VLBM_OK;
    if (convert_planes(h, (void*)source, dbl, 1)) rc = check_launch(-1, "conversion kernel");
    // the extra plane z = NZ: what the last step pushed across the z ends
    // (the rest state before any step)
    if (!rc && !h->direct) {
        rc = with_params(h, [&](void* pv, auto& p) {
            p.src = h->lat[h->cur];
            p.dst = dbl + total;
            p.aux_mode = kAuxGhost;
            p.direct = 0;
            return check_launch(launch_aux(d.q, d.precision, d.rheology, h->mask_geom, pv, h->stream, state_rep(h)),
                                "extra plane kernel");
        });
    }
    if (!rc && (hipMemcpyAsync(host.data(), dbl, (h->direct ? total : total + ghost) * 8, hipMemcpyDeviceToHost,
                               h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess))
        rc = fail(VLBM_ERR_DEVICE, "population download failed");
    (void)hipFree(dbl);
    if (rc) return rc;
    for (int i = 0; i < d.q; i++) {
        const double w = pop_weight(d.q, i);
        double* dstp = pop + (size_t)i * plane * (d.nz + 1);
        const double* srcp = host.data() + (size_t)i * h->n;
        for (size_t k = 0; k < h->n; k++) dstp[k] = srcp[k] + w;
        const double* gp = host.data() + total + (size_t)i * plane;
        for (size_t k = 0; k < plane; k++) dstp[h->n + k] = gp[k] + w;
    }
    return VLBM_OK;
}

// n_steps steps; with IBM each LBM step is followed by the IBM stage, and
// ev (2 events per step plus one, or null) records the boundaries of both.
static int enqueue_steps(vlbm_h h, int n_steps, int save_last, hipEvent_t* ev = nullptr) {
    const vlbm_desc& d = h->d;
    const bool ibm = h->ibm != nullptr;
    if (ev) HIP_TRY(hipEventRecord(ev[0], h->stream));
    for (int s = 0; s < n_steps; s++) {
        const int other = 1 - h->cur;
        const bool save = save_last && (s == n_steps - 1);
        int rc = with_params(h, [&](void* pv, auto& p) {
            p.src = state_src(h);
            p.dst = h->lat[other];
            // IBM builds take the macroscopics from the stage (save_macr_to_array
            // is false with IBM, src/CUDA/main.cu:293-297)
            p.save = (save && !ibm) ? 1 : 0;
            p.sr_seed = (uint32_t)h->steps * 0x9e3779b9u + 0x7f4a7c15u;
            return check_launch(launch_step(d.q, d.precision, d.rheology, h->mask_geom, h->direct, pv, h->stream, ibm,
                                            state_rep(h), h->mom ? kRepMom : kRepPop),
                                "step kernel");
        });
        if (rc) return rc;
        h->cur = other;
        h->direct = false;
        h->direct_pops = false;
        if (h->aux_pending) {
            // rheometer resumed from a checkpoint: the direct step has read the
            // loaded state, and the idle lattice now takes popAux's plate slots
            h->aux_pending = false;
            const size_t plane = (size_t)d.nx * d.ny;
            for (const auto& sl : plate_slots(d)) {
                const size_t off = ((size_t)sl.first * (size_t)h->stride + (size_t)sl.second * plane) * h->sbytes;
                HIP_TRY(hipMemcpyAsync((char*)h->lat[1 - h->cur] + off, (const char*)h->scratch_pops + off,
                                       plane * h->sbytes, hipMemcpyDeviceToDevice, h->stream));
            }
        }
        if (ev) HIP_TRY(hipEventRecord(ev[2 * s + 1], h->stream));
        if (ibm) {
            if ((rc = ibm_stage(h))) return rc;
            if (save && (rc = ibm_fill_macr(h))) return rc;
        }
        if (ev) HIP_TRY(hipEventRecord(ev[2 * s + 2], h->stream));
        h->steps++;
        h->lbm_steps++;
    }
    return VLBM_OK;
}

int vlbm_step(vlbm_h h, int n_steps, int save_last) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    if (n_steps < 0) return fail(VLBM_ERR_INVALID, "n_steps must be >= 0");
    return enqueue_steps(h, n_steps, save_last);
}

int vlbm_time_steps(vlbm_h h, int n_steps, double* ms) {
    if (!h || !ms) return fail(VLBM_ERR_INVALID, "handle or output NULL");
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    int rc = enqueue_steps(h, n_steps, 0);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return VLBM_OK;
}

int vlbm_ibm_time_steps(vlbm_h h, int n_steps, double* ms_lbm, double* ms_ibm) {
    if (!h || !ms_lbm || !ms_ibm) return fail(VLBM_ERR_INVALID, "handle or outputs NULL");
    if (!h->ibm) return fail(VLBM_ERR_INVALID, "IBM is not enabled on this handle (vlbm_ibm_enable)");
    if (n_steps < 1) return fail(VLBM_ERR_INVALID, "n_steps must be >= 1");
    std::vector<hipEvent_t> ev((size_t)2 * n_steps + 1, nullptr);
    int rc = VLBM_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { rc = fail(VLBM_ERR_DEVICE, "hipEventCreate failed"); break; }
    if (!rc) rc = enqueue_steps(h, n_steps, 0, ev.data());
    if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(VLBM_ERR_DEVICE, "event sync failed");
    double a = 0, b = 0;
    for (int s = 0; s < n_steps && !rc; s++) {
        float t0 = 0, t1 = 0;
        if (hipEventElapsedTime(&t0, ev[2 * s], ev[2 * s + 1]) != hipSuccess ||
            hipEventElapsedTime(&t1, ev[2 * s + 1], ev[2 * s + 2]) != hipSuccess)
            rc = fail(VLBM_ERR_DEVICE, "hipEventElapsedTime failed");
        a += t0;
        b += t1;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    *ms_lbm = a;
    *ms_ibm = b;
    return rc;
}

int vlbm_compute_macroscopics(vlbm_h h) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    // with IBM the listed tiles take the macroscopics of the last stage
    if (h->ibm) return ibm_fill_macr(h);
    const vlbm_desc& d = h->d;
    return with_params(h, [&](void* pv, auto& p) {
        p.src = state_src(h);
        p.dst = h->lat[1 - h->cur];
        p.aux_mode = kAuxMoments;
        p.direct = h->direct ? 1 : 0;
        return check_launch(launch_aux(d.q, d.precision, d.rheology, h->mask_geom, pv, h->stream, state_rep(h)),
                            "moments kernel");
    });
}

int vlbm_macroscopics(vlbm_h h, double* rho, double* ux, double* uy, double* uz, double* omega) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    double* outs[5] = {rho, ux, uy, uz, omega};
    for (int k = 0; k < 5; k++) {
        if (!outs[k]) continue;
        if (!h->macr[k]) return fail(VLBM_ERR_INVALID, "omega exists for non-Newtonian fluids only");
        int rc = fetch_field(h, h->macr[k], outs[k]);
        if (rc) return rc;
    }
    return VLBM_OK;
}

int vlbm_visc_turb(vlbm_h h, double* out) {
    if (!h || !out) return fail(VLBM_ERR_INVALID, "handle or output NULL");
    if (!h->visc_turb) return fail(VLBM_ERR_INVALID, "the handle keeps no eddy viscosity (desc.les_visc_turb)");
    return fetch_field(h, h->visc_turb, out);
}

int vlbm_reduce(vlbm_h h, double* residual, double* mean_rho) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    if (launch_reduce(h->tbytes, h->macr[0], h->macr[1], h->macr[2], h->macr[3], h->uold[0], h->uold[1],
                      h->uold[2], (uint32_t)h->n, h->partials, h->out3, h->stream))
        return check_launch(-1, "reduction kernel");
    double s[3];
    HIP_TRY(hipMemcpyAsync(s, h->out3, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (residual) *residual = (s[1] != 0) ? s[0] / s[1] : 1.0;
    if (mean_rho) *mean_rho = s[2] / (double)h->n;
    return VLBM_OK;
}

// getVarFilename (src/CUDA/lbmReport.cpp:55-83): step zero-padded to at least 6 digits.
static std::string var_filename(const char* path, const char* id, const char* var, unsigned step) {
    unsigned n_zeros = 0, pot10 = 10;
    const unsigned aux1 = 1000000;
    if (step != 0)
        for (n_zeros = 0; (unsigned long long)step * pot10 < aux1; pot10 *= 10) n_zeros++;
    else
        n_zeros = 6;
    std::string s = std::string(path) + "/" + id + "/" + id + "_" + var;
    for (unsigned i = 0; i < n_zeros; i++) s += "0";
    return s + std::to_string(step) + ".bin";
}

int vlbm_save_macr_bin(vlbm_h h, const char* path, const char* id, unsigned step) {
    if (!h || !path || !id) return fail(VLBM_ERR_INVALID, "NULL argument");
    // saveAllMacrBin (lbmReport.cpp:133-172): rho, u, omega (non-Newtonian),
    // visc_turb (LES_EXPORT_VISC_TURBULENT)
    const char* names[6] = {"rho", "ux", "uy", "uz", "omega", "visc_turb"};
    const void* fields[6] = {h->macr[0], h->macr[1], h->macr[2], h->macr[3], h->macr[4], h->visc_turb};
    std::vector<double> buf(h->n);
    for (int k = 0; k < 6; k++) {
        if (!fields[k] || (k == 4 && h->d.rheology == VLBM_NEWTONIAN)) continue;
        int rc = fetch_field(h, fields[k], buf.data());
        if (rc) return rc;
        const std::string fn = var_filename(path, id, names[k], step);
        FILE* f = std::fopen(fn.c_str(), "wb");
        if (!f) return fail(VLBM_ERR_IO, "cannot open " + fn);
        size_t wrote;
        if (h->d.precision == VLBM_FP64) {
            wrote = std::fwrite(buf.data(), 8, h->n, f);
        } else {
            std::vector<float> fl(h->n);
            for (size_t i = 0; i < h->n; i++) fl[i] = (float)buf[i];
            wrote = std::fwrite(fl.data(), 4, h->n, f);
        }
        std::fclose(f);
        if (wrote != h->n) return fail(VLBM_ERR_IO, "short write to " + fn);
    }
    return VLBM_OK;
}

int vlbm_save_info(vlbm_h h, const char* path, const char* id, unsigned total_steps, double mlups, double seconds) {
    if (!h || !path || !id) return fail(VLBM_ERR_INVALID, "NULL argument");
    const vlbm_desc& d = h->d;
    const std::string fn = std::string(path) + "/" + id + "/" + id + "_info.txt";
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) return fail(VLBM_ERR_IO, "cannot open " + fn);
    const double pop_bytes = (double)d.q * d.nx * d.ny * (d.nz + 1) * (d.precision == VLBM_FP64 ? 8 : 4);
    const double bw = seconds > 0 ? pop_bytes * 2.0 / (seconds * (double)(1u << 30)) * total_steps : 0.0;
    std::fprintf(f, "---------------------------- SIMULATION INFORMATION ----------------------------\n");
    std::fprintf(f, "      Simulation ID: %s\n", id);
    std::fprintf(f, "       Velocity set: D3Q%d\n", d.q);
    std::fprintf(f, "          Precision: %s\n", d.precision == VLBM_FP64 ? "double" : "float");
    std::fprintf(f, "                 NX: %d\n", d.nx);
    std::fprintf(f, "                 NY: %d\n", d.ny);
    std::fprintf(f, "                 NZ: %d\n", d.nz);
    std::fprintf(f, "           NZ_TOTAL: %d\n", d.nz);
    std::fprintf(f, "                Tau: %.6e\n", d.tau);
    std::fprintf(f, "               Umax: %.6e\n", d.ux_bc[1]);
    std::fprintf(f, "                 FX: %.6e\n", d.fx);
    std::fprintf(f, "                 FY: %.6e\n", d.fy);
    std::fprintf(f, "                 FZ: %.6e\n", d.fz);
    std::fprintf(f, "             Nsteps: %u\n", total_steps);
    std::fprintf(f, "              MLUPS: %.1f\n", mlups);
    std::fprintf(f, "          Bandwidht: %.1f (Gb/s)\n", bw);
    std::fprintf(f, "       Time elapsed: %.3f (s)\n", seconds);
    std::fprintf(f, "--------------------------------------------------------------------------------\n");
    if (d.rheology != VLBM_NEWTONIAN) {
        std::fprintf(f, "\n------------------------------ NON NEWTONIAN FLUID -----------------------------\n");
        if (d.rheology == VLBM_POWERLAW) {
            std::fprintf(f, "              Model: Power-Law\n");
            std::fprintf(f, "        Power index: %.6e\n", d.n_index);
            std::fprintf(f, " Consistency factor: %.6e\n", d.k_consistency);
            std::fprintf(f, "            Gamma 0: %.6e\n", d.gamma_0);
        } else {
            std::fprintf(f, "              Model: Bingham\n");
            std::fprintf(f, "  Plastic viscosity: %.6e\n", d.eta_p);
            std::fprintf(f, "       Yield stress: %.6e\n", d.s_y);
            std::fprintf(f, "      Plastic omega: %.6e\n", 1.0 / (3 * d.eta_p + 0.5));
        }
        std::fprintf(f, "--------------------------------------------------------------------------------\n");
    }
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if (!ok) return fail(VLBM_ERR_IO, "write error on " + fn);
    return VLBM_OK;
}

int vlbm_steps_done(vlbm_h h, long long* steps) {
    if (!h || !steps) return fail(VLBM_ERR_INVALID, "NULL argument");
    *steps = h->steps;
    return VLBM_OK;
}

int vlbm_sync(vlbm_h h) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return VLBM_OK;
}

int vlbm_copy_probe(size_t bytes, int iters, double* ms) {
    if (!ms || iters < 1 || bytes < 16) return fail(VLBM_ERR_INVALID, "bytes >= 16, iters >= 1 and an output are required");
    bytes = bytes / 16 * 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = VLBM_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = fail(VLBM_ERR_ALLOC, "copy buffers");
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = fail(VLBM_ERR_DEVICE, "stream / events");
    if (!rc && hipMemsetAsync(a, 0, bytes, st) != hipSuccess) rc = fail(VLBM_ERR_DEVICE, "hipMemset");
    if (!rc && launch_copy_probe(a, b, bytes, st)) rc = check_launch(-1, "copy kernel");   // warm
    if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = fail(VLBM_ERR_DEVICE, "event");
    for (int k = 0; k < iters && !rc; k++)
        if (launch_copy_probe(a, b, bytes, st)) rc = check_launch(-1, "copy kernel");
    if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = fail(VLBM_ERR_DEVICE, "event");
    float t = 0;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(VLBM_ERR_DEVICE, "elapsed time");
    *ms = rc ? 0.0 : t / iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int vlbm_device_bytes(vlbm_h h, size_t* bytes) {
    if (!h || !bytes) return fail(VLBM_ERR_INVALID, "NULL argument");
    *bytes = h->device_bytes;
    return VLBM_OK;
}

int vlbm_layout(vlbm_h h, int* layout) {
    if (!h || !layout) return fail(VLBM_ERR_INVALID, "NULL argument");
    *layout = h->mom ? VLBM_LAYOUT_MOMENTS : VLBM_LAYOUT_POPULATIONS;
    return VLBM_OK;
}

}  // extern "C"
