// C-ABI of the immersed boundary stage (declarations in include/vlbm.h):
// particle set-up on the host (sphere meshes), device state, the per-step
// launch sequence and host access to particles, nodes and fields.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "ibm.h"
#include "ibm_state.h"

using namespace vlbm;
using namespace vlbm::detail;

namespace {

constexpr double kSqrt2 = 1.41421356237309504880168872420969807856967187537;

int ibm_alloc(vlbm_ctx* h, void** p, size_t bytes) {
    int rc = dalloc(h, p, bytes ? bytes : 8);
    if (rc == VLBM_OK) h->ibm->allocs.push_back(*p);
    return rc;
}

// Uniform cell list: cells at least one contact distance (2 r_max, with a
// margin for the float gap test) wide.
void update_cells(const vlbm_ctx* h, IbmState& s) {
    // LDS spreading box for FP32 fluids when it fits the workgroup's LDS
    s.P.box_edge = 0;
    if (h->tbytes == 4 && s.r_max > 0) {
        const int edge = ibm_box_edge(s.r_max);
        const size_t bytes = 3 * sizeof(int) * (size_t)edge * edge * edge;
        hipDeviceProp_t prop;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
            bytes + 4096 <= prop.sharedMemPerBlock)
            s.P.box_edge = edge;
    }
    // plain stores into box nodes need every lattice node to appear once in a
    // box: no periodic axis shorter than the box
    s.P.box_plain = 0;
    s.P.box_reach = std::ceil(s.r_max) + 0.5;
    if (s.P.box_edge > 0) {
        const int n3[3] = {h->d.nx, h->d.ny, h->d.nz};
        s.P.box_plain = 1;
        for (int a = 0; a < 3; a++)
            if (s.P.periodic[a] && n3[a] < s.P.box_edge) s.P.box_plain = 0;
    }
    const int n[3] = {h->d.nx, h->d.ny, h->d.nz};
    const double cs = 2.0 * s.r_max * 1.0001;
    for (int a = 0; a < 3; a++) {
        int c = cs > 0 ? (int)(n[a] / cs) : 1;
        c = c < 1 ? 1 : c;
        s.P.cells[a] = c;
        s.P.cell_w[a] = (double)n[a] / c;
    }
}

// Layered polar mesh of makeSpherePolar (src/CUDA/IBM/structs/particle.cpp:698-995):
// rings at equally spaced polar angles, node counts ~ cos(theta), ring areas
// from the spherical-zone formula, alternate rings phase-shifted; optional
// Coulomb relaxation with equal areas.
void polar_mesh(double diameter, const double c[3], double scale, double breugem, int coulomb,
                std::vector<double>& xyz, std::vector<double>& area) {
    double r = diameter / 2.0;
    const double s_total = 4.0 * M_PI * r * r;
    r -= breugem;
    const unsigned layers = (unsigned)(2.0 * std::sqrt(2) * r / scale + 1.0);
    std::vector<unsigned> per(layers + 1);
    std::vector<double> theta(layers + 1), height(layers + 1), zone(layers + 1);
    unsigned total = 0;
    for (unsigned i = 0; i <= layers; i++) {
        theta[i] = M_PI * ((double)i / (double)layers - 0.5);
        per[i] = (unsigned)(1.5 + std::cos(theta[i]) * layers * std::sqrt(3));
        total += per[i];
        height[i] = r * std::sin(theta[i]);
    }
    // zone areas from the south pole up to mid-layer heights, then differences
    for (unsigned i = 0; i < layers; i++) zone[i] = (height[i] + height[i + 1]) / 2.0 - height[0];
    zone[layers] = 2 * r;
    for (unsigned i = 0; i <= layers; i++) zone[i] = 2 * M_PI * r * zone[i];
    for (unsigned i = layers; i > 0; i--) zone[i] = zone[i] - zone[i - 1];
    zone[0] = zone[layers];
    std::vector<double> P(3 * (size_t)total, 0.0), A(total, 0.0);
    P[2] = r * std::sin(theta[0]);
    A[0] = zone[0];
    unsigned k = 1;
    double phase = 0.0;
    for (unsigned i = 1; i < layers; i++) {
        if (i % 2 == 1) phase = phase + M_PI / per[i];
        for (unsigned j = 0; j < per[i]; j++, k++) {
            const double ang = (double)j * 2.0 * M_PI / per[i] + phase;
            P[3 * k + 0] = r * std::cos(theta[i]) * std::cos(ang);
            P[3 * k + 1] = r * std::cos(theta[i]) * std::sin(ang);
            P[3 * k + 2] = r * std::sin(theta[i]);
            A[k] = zone[i] / per[i];
        }
    }
    P[3 * (total - 1) + 2] = r * std::sin(theta[layers]);
    A[total - 1] = zone[layers];
    if (coulomb != 0) {
        std::vector<double> push(3 * (size_t)total);
        for (int it = 0; it < coulomb; it++) {
            std::fill(push.begin(), push.end(), 0.0);
            for (unsigned i = 0; i < total; i++)
                for (unsigned j = i + 1; j < total; j++) {
                    const double dx = P[3 * j] - P[3 * i], dy = P[3 * j + 1] - P[3 * i + 1], dz = P[3 * j + 2] - P[3 * i + 2];
                    const double m = dx * dx + dy * dy + dz * dz;
                    push[3 * i] -= dx / m; push[3 * i + 1] -= dy / m; push[3 * i + 2] -= dz / m;
                    push[3 * j] -= -dx / m; push[3 * j + 1] -= -dy / m; push[3 * j + 2] -= -dz / m;
                }
            for (unsigned i = 0; i < total; i++) {
                for (int a = 0; a < 3; a++) P[3 * i + a] += push[3 * i + a] / 0.001;
                const double m = std::sqrt(P[3 * i] * P[3 * i] + P[3 * i + 1] * P[3 * i + 1] + P[3 * i + 2] * P[3 * i + 2]);
                for (int a = 0; a < 3; a++) P[3 * i + a] *= r / m;
            }
        }
        for (unsigned i = 0; i < total; i++) A[i] = s_total / total;
    }
    xyz.resize(3 * (size_t)total);
    area.assign(A.begin(), A.end());
    for (unsigned i = 0; i < total; i++)
        for (int a = 0; a < 3; a++) xyz[3 * i + a] = P[3 * i + a] + c[a];
}

// Fibonacci sphere lattice: z_k = 1 - (2k+1)/n, golden-angle longitudes,
// equal areas 4 pi r^2 / n.
void fibonacci_mesh(double r, const double c[3], int n, std::vector<double>& xyz, std::vector<double>& area) {
    const double golden = M_PI * (3.0 - std::sqrt(5.0));
    xyz.resize(3 * (size_t)n);
    area.assign((size_t)n, 4.0 * M_PI * r * r / n);
    for (int k = 0; k < n; k++) {
        const double z = 1.0 - (2.0 * k + 1.0) / n;
        const double rho = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double phi = golden * k;
        xyz[3 * k + 0] = c[0] + r * rho * std::cos(phi);
        xyz[3 * k + 1] = c[1] + r * rho * std::sin(phi);
        xyz[3 * k + 2] = c[2] + r * z;
    }
}

template <typename T>
IbmFluid<T> fluid_of(const vlbm_ctx* h) {
    const IbmState& s = *h->ibm;
    IbmFluid<T> fl;
    fl.stage = (StageRec<T>*)s.stage;
    fl.fx = (T*)s.force[0]; fl.fy = (T*)s.force[1]; fl.fz = (T*)s.force[2];
    fl.out_rho = (T*)h->macr[0]; fl.out_ux = (T*)h->macr[1]; fl.out_uy = (T*)h->macr[2]; fl.out_uz = (T*)h->macr[3];
    return fl;
}

int stage_part(vlbm_ctx* h, int part) {
    IbmState& s = *h->ibm;
    int rc;
    if (h->tbytes == 8) {
        const IbmFluid<double> fl = fluid_of<double>(h);
        rc = launch_ibm_stage(s.P, 8, &fl, part, h->stream);
    } else {
        const IbmFluid<float> fl = fluid_of<float>(h);
        rc = launch_ibm_stage(s.P, 4, &fl, part, h->stream);
    }
    return check_launch(rc, "IBM stage kernels");
}

int sync(vlbm_ctx* h) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    return VLBM_OK;
}

int need_ibm(vlbm_ctx* h) {
    if (!h) return fail(VLBM_ERR_INVALID, "handle is NULL");
    if (!h->ibm) return fail(VLBM_ERR_INVALID, "IBM is not enabled on this handle (vlbm_ibm_enable)");
    return VLBM_OK;
}

int download_spheres(vlbm_ctx* h, std::vector<IbmSphere>& out) {
    IbmState& s = *h->ibm;
    out.resize((size_t)s.P.n_spheres);
    if (out.empty()) return VLBM_OK;
    HIP_TRY(hipMemcpyAsync(out.data(), s.P.sph, out.size() * sizeof(IbmSphere), hipMemcpyDeviceToHost, h->stream));
    return sync(h);
}

int add_particle(vlbm_ctx* h, const double* state, double radius, double density, int movable, int n,
                 const double* xyz, const double* area, int* id) {
    int rc = need_ibm(h);
    if (rc) return rc;
    IbmState& s = *h->ibm;
    if (!state || radius <= 0 || density <= 0 || n < 0 || (n > 0 && (!xyz || !area)))
        return fail(VLBM_ERR_INVALID, "invalid sphere: state, radius > 0, density > 0 and the nodes are required");
    if (s.P.n_spheres + 1 > s.d.max_spheres) return fail(VLBM_ERR_INVALID, "sphere capacity (max_spheres) exceeded");
    if (s.P.n_nodes + n > s.d.max_nodes) return fail(VLBM_ERR_INVALID, "node capacity (max_nodes) exceeded");
    IbmSphere c;
    std::memset(&c, 0, sizeof(c));
    for (int a = 0; a < 3; a++) {
        c.pos[a] = c.pos_old[a] = state[a];
        c.vel[a] = c.vel_old[a] = state[3 + a];
        c.w[a] = c.w_avg[a] = c.w_old[a] = state[6 + a];
    }
    c.radius = radius;
    c.volume = radius * radius * radius * 4 * M_PI / 3;
    c.density = density;
    const double inertia = 2.0 * c.volume * density * radius * radius / 5.0;
    c.I[0] = c.I[1] = c.I[2] = inertia;
    c.movable = movable ? 1 : 0;
    c.first_node = s.P.n_nodes;
    c.n_nodes = n;
    for (int k = 0; k < kMaxColl; k++) { c.partner[k] = -1; c.last[k] = -1; }
    // collision.semiAxis = (r, r, r) as a point; semiAxis2/3 are not set for
    // spheres there (particle.cpp:777-778) and start at 0 here
    for (int a = 0; a < 3; a++) c.axis[0][a] = radius;
    const int sid = s.P.n_spheres;
    HIP_TRY(hipMemcpyAsync(s.P.sph + sid, &c, sizeof(c), hipMemcpyHostToDevice, h->stream));
    if (n > 0) {
        std::vector<double> col((size_t)n);
        double* dst[3] = {s.P.px, s.P.py, s.P.pz};
        for (int a = 0; a < 3; a++) {
            for (int k = 0; k < n; k++) col[k] = xyz[3 * k + a];
            HIP_TRY(hipMemcpyAsync(dst[a] + c.first_node, col.data(), n * 8, hipMemcpyHostToDevice, h->stream));
            rc = sync(h);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(s.P.pS + c.first_node, area, n * 8, hipMemcpyHostToDevice, h->stream));
        std::vector<int> ids((size_t)n, sid);
        HIP_TRY(hipMemcpyAsync(s.P.pid + c.first_node, ids.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        double* zero[9] = {s.P.pfx, s.P.pfy, s.P.pfz, s.P.pdx, s.P.pdy, s.P.pdz, s.P.pmx, s.P.pmy, s.P.pmz};
        for (double* z : zero) HIP_TRY(hipMemsetAsync(z + c.first_node, 0, n * 8, h->stream));
        rc = sync(h);   // host buffers of this block stay alive until the copies are done
        if (rc) return rc;
    }
    // interpolation chunks of this sphere
    {
        std::vector<int> cs, cf, cc;
        for (int k = 0; k < n; k += kChunkNodes) {
            cs.push_back(sid);
            cf.push_back(c.first_node + k);
            cc.push_back(std::min(kChunkNodes, n - k));
        }
        const int nc = (int)cs.size();
        const int end = s.P.n_chunks + nc;
        if (nc) {
            HIP_TRY(hipMemcpyAsync(s.P.chunk_sphere + s.P.n_chunks, cs.data(), nc * 4, hipMemcpyHostToDevice, h->stream));
            HIP_TRY(hipMemcpyAsync(s.P.chunk_first + s.P.n_chunks, cf.data(), nc * 4, hipMemcpyHostToDevice, h->stream));
            HIP_TRY(hipMemcpyAsync(s.P.chunk_count + s.P.n_chunks, cc.data(), nc * 4, hipMemcpyHostToDevice, h->stream));
        }
        HIP_TRY(hipMemcpyAsync(s.P.chunk0 + sid + 1, &end, 4, hipMemcpyHostToDevice, h->stream));
        rc = sync(h);
        if (rc) return rc;
        s.P.n_chunks = end;
    }
    s.P.n_spheres++;
    s.P.n_nodes += n;
    s.r_max = std::max(s.r_max, radius);
    update_cells(h, s);
    if (id) *id = sid;
    return VLBM_OK;
}

}  // namespace

namespace vlbm {
namespace detail {

void ibm_release(vlbm_ctx* h) {
    if (!h || !h->ibm) return;
    for (void* p : h->ibm->allocs) (void)hipFree(p);
    delete h->ibm;
    h->ibm = nullptr;
    h->tile_flag = nullptr;
    h->ibm_force[0] = h->ibm_force[1] = h->ibm_force[2] = nullptr;
}

// One stage of immersedBoundaryMethod (src/CUDA/IBM/ibm.cu:6-192) on the
// current state, labelled with the steps done before it.
int ibm_stage(vlbm_ctx* h) {
    IbmState& s = *h->ibm;
    s.P.step = (int)h->steps;
    int rc = stage_part(h, kIbmBegin);
    if (rc) return rc;
    const vlbm_desc& d = h->d;
    const uint32_t max_tiles = s.P.n_tiles;
    rc = with_params(h, [&](void* pv, auto& p) {
        using T = typename std::remove_reference<decltype(*p.m_rho)>::type;
        p.src = state_src(h);
        p.dst = h->lat[1 - h->cur];
        p.direct = h->direct ? 1 : 0;
        p.ibm_stage = (StageRec<T>*)s.stage;
        return check_launch(launch_ibm_macr(d.q, d.precision, h->mask_geom, pv, nullptr, s.P.tile_list, s.P.tile_count,
                                            max_tiles, h->stream, state_rep(h)),
                            "IBM macroscopics kernel");
    });
    if (rc) return rc;
    if ((rc = stage_part(h, kIbmContacts))) return rc;
    for (int it = 0; it < s.d.max_iterations; it++)
        if ((rc = stage_part(h, kIbmIteration))) return rc;
    return stage_part(h, kIbmMove);
}

// Macroscopics of the nodes outside the flagged tiles (constant force), so
// that the saved arrays hold the whole lattice after IBM steps.
int ibm_fill_macr(vlbm_ctx* h) {
    const vlbm_desc& d = h->d;
    int rc = with_params(h, [&](void* pv, auto& p) {
        p.src = state_src(h);
        p.dst = h->lat[1 - h->cur];
        p.aux_mode = kAuxMomentsOutsideTiles;
        p.direct = h->direct ? 1 : 0;
        return check_launch(launch_aux(d.q, d.precision, d.rheology, h->mask_geom, pv, h->stream, state_rep(h)),
                            "moments kernel");
    });
    if (rc) return rc;
    return stage_part(h, kIbmOutput);
}

int ibm_reset_tiles(vlbm_ctx* h) {
    if (!h->ibm) return VLBM_OK;
    HIP_TRY(hipMemsetAsync(h->ibm->P.tile_flag, 0, (size_t)h->ibm->P.n_tiles * 4, h->stream));
    HIP_TRY(hipMemsetAsync(h->ibm->P.tile_count, 0, 4, h->stream));
    return sync(h);
}

}  // namespace detail
}  // namespace vlbm

extern "C" {

void vlbm_ibm_desc_defaults(vlbm_ibm_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->max_iterations = 1;                        // IBM_MAX_ITERATION
    d->periodic_x = 1; d->periodic_y = 0; d->periodic_z = 1;   // ibmBoundaryCondition.h:20-55
    d->thickness = 1.0;                           // IBM_THICKNESS
    d->alpha = 0.5;                               // IBM_MOVEMENT_DISCRETIZATION
    d->fluid_density = 1.0;                       // FLUID_DENSITY
    d->pp_friction = 0.1; d->pw_friction = 0.1;   // ibmBoundaryCondition.h:83-86
    d->pp_rest = 0.9; d->pw_rest = 0.9;
    d->young = 0.5; d->poisson = 0.3;             // ibmBoundaryCondition.h:90-96
    d->wall_young = 100.0; d->wall_poisson = 0.3;
    d->ref_float_helpers = 1;
    d->mesh_scale = 1.0;                          // MESH_SCALE
    d->breugem = 0.0;                             // BREUGEM_PARAMETER
    d->coulomb = 0;                               // MESH_COULOMB
    d->max_spheres = 256;
    d->max_nodes = 262144;
}

int vlbm_ibm_enable(vlbm_h h, const vlbm_ibm_desc* d) {
    if (!h || !d) return fail(VLBM_ERR_INVALID, "handle or IBM descriptor is NULL");
    if (h->ibm) return fail(VLBM_ERR_INVALID, "IBM is already enabled on this handle");
    if (d->max_iterations < 0) return fail(VLBM_ERR_INVALID, "max_iterations must be >= 0");
    if (d->max_spheres < 1 || d->max_spheres > kMaxSortKeys)
        return fail(VLBM_ERR_INVALID, "max_spheres must be in [1, " + std::to_string(kMaxSortKeys) + "]");
    if (d->max_nodes < 0) return fail(VLBM_ERR_INVALID, "max_nodes must be >= 0");
    if (!(d->young > 0) || !(d->wall_young > 0) || !(d->pp_rest > 0) || !(d->pw_rest > 0))
        return fail(VLBM_ERR_INVALID, "Young's moduli and restitution coefficients must be > 0");
    if (!(d->mesh_scale > 0)) return fail(VLBM_ERR_INVALID, "mesh_scale must be > 0");
    h->ibm = new IbmState;
    IbmState& s = *h->ibm;
    s.d = *d;
    std::memset(&s.P, 0, sizeof(s.P));
    IbmParams& P = s.P;
    P.nx = h->d.nx; P.ny = h->d.ny; P.nz = h->d.nz;
    P.periodic[0] = d->periodic_x != 0; P.periodic[1] = d->periodic_y != 0; P.periodic[2] = d->periodic_z != 0;
    const int n3[3] = {P.nx, P.ny, P.nz};
    for (int a = 0; a < 3; a++) P.tiles[a] = tiles_along(n3[a], a);
    P.n_tiles = (uint32_t)P.tiles[0] * (uint32_t)P.tiles[1] * (uint32_t)P.tiles[2];
    P.thickness = d->thickness; P.alpha = d->alpha; P.fluid_density = d->fluid_density;
    P.g[0] = d->gx; P.g[1] = d->gy; P.g[2] = d->gz;
    // contact constants (src/CUDA/IBM/ibmBoundaryCondition.h:90-107, ibmCollision.cu:422)
    const double pG = d->young / (2.0 + 2.0 * d->poisson), wG = d->wall_young / (2.0 + 2.0 * d->wall_poisson);
    P.pp_kn = (4.0 / 3.0) / ((1 - d->poisson * d->poisson) / d->young + (1 - d->poisson * d->poisson) / d->young);
    P.pw_kn = (4.0 / 3.0) / ((1 - d->poisson * d->poisson) / d->young + (1 - d->wall_poisson * d->wall_poisson) / d->wall_young);
    P.pp_kt = 4.0 * kSqrt2 / ((2 - d->poisson) / pG + (2 - d->poisson) / pG);
    P.pw_kt = 4.0 * kSqrt2 / ((2 - d->poisson) / pG + (2 - d->wall_poisson) / wG);
    const double lp = std::log(d->pp_rest), lw = std::log(d->pw_rest);
    P.pp_damp = (-2.0 * lp / (std::sqrt(M_PI * M_PI + lp * lp)));
    P.pw_damp = (-2.0 * lw / (std::sqrt(M_PI * M_PI + lw * lw)));
    P.pp_mu = d->pp_friction; P.pw_mu = d->pw_friction;
    P.ref_float = d->ref_float_helpers != 0;
    update_cells(h, s);
    const size_t ms = (size_t)d->max_spheres, mn = (size_t)d->max_nodes;
    int rc = VLBM_OK;
    auto A = [&](void** p, size_t bytes) { if (!rc) rc = ibm_alloc(h, p, bytes); };
    A((void**)&P.sph, ms * sizeof(IbmSphere));
    double** dn[13] = {&P.px, &P.py, &P.pz, &P.pS, &P.pfx, &P.pfy, &P.pfz, &P.pdx, &P.pdy, &P.pdz, &P.pmx, &P.pmy, &P.pmz};
    for (double** q : dn) A((void**)q, mn * 8);
    A((void**)&P.pid, mn * 4);
    A((void**)&P.tile_flag, (size_t)P.n_tiles * 4);
    A((void**)&P.tile_list, (size_t)P.n_tiles * 4);
    A((void**)&P.tile_count, 4);
    A((void**)&P.cell_keys, ms * sizeof(uint64_t));
    A((void**)&P.contacts, ms * kMaxContacts * sizeof(IbmContact));
    A((void**)&P.n_contacts, ms * 4);
    const size_t mc = mn / kChunkNodes + ms + 1;
    A((void**)&P.chunk_sphere, mc * 4);
    A((void**)&P.chunk_first, mc * 4);
    A((void**)&P.chunk_count, mc * 4);
    A((void**)&P.chunk0, (ms + 1) * 4);
    A((void**)&P.chunk_sums, mc * 6 * sizeof(double));
    A((void**)&P.error, 4);
    A((void**)&P.box_nb, ms * (kBoxNb + 1) * 4);
    A((void**)&P.box_stray, 4);
    for (int a = 0; a < 3; a++) A(&s.force[a], h->n * h->tbytes);
    A(&s.stage, h->n * 4 * h->tbytes);
    if (rc) { ibm_release(h); return rc; }
    hipError_t e = hipSuccess;
    for (int a = 0; a < 3 && e == hipSuccess; a++) e = hipMemsetAsync(s.force[a], 0, h->n * h->tbytes, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.tile_flag, 0, (size_t)P.n_tiles * 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.tile_count, 0, 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.error, 0, 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.n_contacts, 0, ms * 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.chunk0, 0, (ms + 1) * 4, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        ibm_release(h);
        return fail(VLBM_ERR_DEVICE, std::string("IBM set-up: ") + hipGetErrorString(e));
    }
    h->tile_flag = P.tile_flag;
    for (int a = 0; a < 3; a++) h->ibm_force[a] = s.force[a];
    h->tiles_x = P.tiles[0];
    h->tiles_y = P.tiles[1];
    return VLBM_OK;
}

int vlbm_ibm_add_sphere(vlbm_h h, const double state[9], double radius, double density, int movable, int mesh,
                        int n_fibonacci, int* id) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (!state || !(radius > 0)) return fail(VLBM_ERR_INVALID, "state and radius > 0 are required");
    std::vector<double> xyz, area;
    if (mesh == VLBM_MESH_POLAR) {
        const IbmState& s = *h->ibm;
        polar_mesh(2.0 * radius, state, s.d.mesh_scale, s.d.breugem, s.d.coulomb, xyz, area);
    } else if (mesh == VLBM_MESH_FIBONACCI) {
        if (n_fibonacci < 1) return fail(VLBM_ERR_INVALID, "n_fibonacci must be >= 1");
        fibonacci_mesh(radius, state, n_fibonacci, xyz, area);
    } else {
        return fail(VLBM_ERR_INVALID, "unknown mesh kind");
    }
    return add_particle(h, state, radius, density, movable, (int)area.size(), xyz.data(), area.data(), id);
}

int vlbm_ibm_add_particle(vlbm_h h, const double state[9], double radius, double density, int movable,
                          int n_nodes, const double* xyz, const double* area, int* id) {
    return add_particle(h, state, radius, density, movable, n_nodes, xyz, area, id);
}

int vlbm_ibm_counts(vlbm_h h, int* n_spheres, int* n_nodes) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (n_spheres) *n_spheres = h->ibm->P.n_spheres;
    if (n_nodes) *n_nodes = h->ibm->P.n_nodes;
    return VLBM_OK;
}

int vlbm_ibm_get_particles(vlbm_h h, double* out) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (!out) return fail(VLBM_ERR_INVALID, "output is NULL");
    std::vector<IbmSphere> sp;
    if ((rc = download_spheres(h, sp))) return rc;
    for (size_t k = 0; k < sp.size(); k++) std::memcpy(out + VLBM_IBM_RECORD * k, &sp[k], VLBM_IBM_RECORD * sizeof(double));
    return VLBM_OK;
}

int vlbm_ibm_set_particles(vlbm_h h, const double* in) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (!in) return fail(VLBM_ERR_INVALID, "input is NULL");
    std::vector<IbmSphere> sp;
    if ((rc = download_spheres(h, sp))) return rc;
    for (size_t k = 0; k < sp.size(); k++) std::memcpy(&sp[k], in + VLBM_IBM_RECORD * k, VLBM_IBM_RECORD * sizeof(double));
    if (!sp.empty())
        HIP_TRY(hipMemcpyAsync(h->ibm->P.sph, sp.data(), sp.size() * sizeof(IbmSphere), hipMemcpyHostToDevice, h->stream));
    return sync(h);
}

int vlbm_ibm_get_collisions(vlbm_h h, int* partner, double* tang, int* last) {
    int rc = need_ibm(h);
    if (rc) return rc;
    std::vector<IbmSphere> sp;
    if ((rc = download_spheres(h, sp))) return rc;
    for (size_t k = 0; k < sp.size(); k++)
        for (int i = 0; i < kMaxColl; i++) {
            if (partner) partner[k * kMaxColl + i] = sp[k].partner[i];
            if (last) last[k * kMaxColl + i] = sp[k].last[i];
            if (tang)
                for (int a = 0; a < 3; a++) tang[(k * kMaxColl + i) * 3 + a] = sp[k].tang[i][a];
        }
    return VLBM_OK;
}

int vlbm_ibm_get_nodes(vlbm_h h, double* pos, double* force, double* dforce) {
    int rc = need_ibm(h);
    if (rc) return rc;
    const IbmParams& P = h->ibm->P;
    const size_t n = (size_t)P.n_nodes;
    if (!n) return VLBM_OK;
    double* src[3][3] = {{P.px, P.py, P.pz}, {P.pfx, P.pfy, P.pfz}, {P.pdx, P.pdy, P.pdz}};
    double* dst[3] = {pos, force, dforce};
    for (int g = 0; g < 3; g++) {
        if (!dst[g]) continue;
        for (int a = 0; a < 3; a++)
            HIP_TRY(hipMemcpyAsync(dst[g] + a * n, src[g][a], n * 8, hipMemcpyDeviceToHost, h->stream));
    }
    return sync(h);
}

int vlbm_ibm_get_force(vlbm_h h, double* f) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (!f) return fail(VLBM_ERR_INVALID, "output is NULL");
    const IbmState& s = *h->ibm;
    std::vector<uint32_t> flags(s.P.n_tiles);
    HIP_TRY(hipMemcpyAsync(flags.data(), s.P.tile_flag, flags.size() * 4, hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    const double f0[3] = {h->d.fx, h->d.fy, h->d.fz};
    for (int a = 0; a < 3; a++) {
        if ((rc = fetch_field(h, s.force[a], f + a * h->n))) return rc;
        size_t k = 0;
        for (int z = 0; z < h->d.nz; z++)
            for (int y = 0; y < h->d.ny; y++)
                for (int x = 0; x < h->d.nx; x++, k++) {
                    const uint32_t t = tile_index(x, y, z, s.P.tiles[0], s.P.tiles[1]);
                    if (!flags[t]) f[a * h->n + k] = f0[a];
                }
    }
    return VLBM_OK;
}

int vlbm_ibm_cells(vlbm_h h, int* dims, int* cell) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if (dims)
        for (int a = 0; a < 3; a++) dims[a] = h->ibm->P.cells[a];
    if (cell) {
        std::vector<IbmSphere> sp;
        if ((rc = download_spheres(h, sp))) return rc;
        for (size_t k = 0; k < sp.size(); k++) cell[k] = sp[k].cell;
    }
    return VLBM_OK;
}

int vlbm_ibm_stage(vlbm_h h) {
    int rc = need_ibm(h);
    if (rc) return rc;
    if ((rc = ibm_stage(h))) return rc;
    h->steps++;   // the next stage carries the next label
    return sync(h);
}

int vlbm_ibm_status(vlbm_h h) {
    int rc = need_ibm(h);
    if (rc) return rc;
    int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, h->ibm->P.error, 4, hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    if (err & kIbmErrContacts) return fail(VLBM_ERR_UNSUPPORTED, "a sphere had more contacts than the contact buffer holds (16)");
    if (err & kIbmErrPartners) return fail(VLBM_ERR_UNSUPPORTED, "a sphere had more lower-index partners than the buffer holds (16)");
    return VLBM_OK;
}

}  // extern "C"
