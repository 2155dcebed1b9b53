// Checkpoint folder in the reference's layout (operateSimCheckpoint,
// src/CUDA/simCheckpoint.cu:216-318): PATH/ID/checkpoint/ holding
//   0_curr_step.bin                  int32 step
//   0_pop.bin, 0_popAux.bin          Q planes of (NZ+1)*NY*NX values, x fastest
//   0_rho.bin, 0_u/{x,y,z}.bin       NX*NY*(NZ+2B) values (B z-planes of border
//                                    each side: 3 with IBM, 0 without)
//   0_omega.bin                      NX*NY*NZ values (non-Newtonian only)
//   0_f/{x,y,z}.bin                  bordered force field (IBM only)
//   0_IBM_particles_centers.bin      ParticleCenter records (IBM only)
//   0_IBM_nodes_*                    node arrays (IBM only)
// Values are raw native-endian dfloat: double for FP64 lattices, float
// otherwise (FP16-storage lattices write their float compute type).

#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "ctx.h"
#include "ibm.h"
#include "ibm_state.h"

using namespace vlbm;
using namespace vlbm::detail;

namespace {

constexpr int kBorderIbm = 3;   // MACR_BORDER_NODES with the reference's IBM defaults (ibmVar.h:148-154)

int make_dir(const std::string& p) {
    if (::mkdir(p.c_str(), 0777) == 0 || errno == EEXIST) return VLBM_OK;
    return fail(VLBM_ERR_IO, "cannot create folder " + p);
}

int write_bytes(const std::string& fn, const void* data, size_t bytes) {
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) return fail(VLBM_ERR_IO, "cannot open " + fn);
    const size_t wrote = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const bool ok = std::fclose(f) == 0 && wrote == bytes;
    return ok ? VLBM_OK : fail(VLBM_ERR_IO, "short write to " + fn);
}

int read_bytes(const std::string& fn, void* data, size_t bytes) {
    FILE* f = std::fopen(fn.c_str(), "rb");
    if (!f) return fail(VLBM_ERR_IO, "cannot open " + fn);
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size < 0 || (size_t)size != bytes) {
        std::fclose(f);
        return fail(VLBM_ERR_INVALID, fn + " holds " + std::to_string(size) + " bytes, the lattice needs " +
                                          std::to_string(bytes));
    }
    const size_t got = bytes ? std::fread(data, 1, bytes, f) : 0;
    std::fclose(f);
    return got == bytes ? VLBM_OK : fail(VLBM_ERR_IO, "short read from " + fn);
}

// doubles <-> file values of the checkpoint's dfloat
int write_values(const std::string& fn, const double* v, size_t n, int tb) {
    if (tb == 8) return write_bytes(fn, v, n * 8);
    std::vector<float> f(n);
    for (size_t k = 0; k < n; k++) f[k] = (float)v[k];
    return write_bytes(fn, f.data(), n * 4);
}

int read_values(const std::string& fn, double* v, size_t n, int tb) {
    if (tb == 8) return read_bytes(fn, v, n * 8);
    std::vector<float> f(n);
    int rc = read_bytes(fn, f.data(), n * 4);
    if (rc) return rc;
    for (size_t k = 0; k < n; k++) v[k] = f[k];
    return VLBM_OK;
}

// A field of NX*NY*NZ values inside B border planes on each z side.
std::vector<double> bordered(const double* v, size_t plane, size_t n, int border) {
    std::vector<double> out(n + 2 * (size_t)border * plane, 0.0);
    std::copy(v, v + n, out.begin() + (size_t)border * plane);
    return out;
}

int sync(vlbm_ctx* h) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    return VLBM_OK;
}

// The reference's ParticleCenter record (particleCenter.h:17-118) with
// dfloat = T, in the host compiler's natural layout: dfloat3/4/6 members as
// arrays, ParticleType as int32, bool movable, then CollisionData.
template <typename T>
struct RefCollisionData {
    int32_t shape;                  // SPHERE = 0
    T semi_axis[3][3];              // semiAxis, semiAxis2, semiAxis3
    int32_t partner[kMaxColl];      // collisionPartnerIDs
    T tang[kMaxColl][3];            // tangentialDisplacements
    int32_t last[kMaxColl];         // lastCollisionStep
};

template <typename T>
struct RefParticleCenter {
    T pos[3], pos_old[3], vel[3], vel_old[3], w[3], w_avg[3], w_old[3], w_pos[3];
    T q_pos[4], q_pos_old[4];       // x, y, z, w
    T f[3], f_old[3], M[3], M_old[3];
    T I[6];                         // xx yy zz xy xz yz
    T dP[3], dL[3];
    T S, radius, volume, density;
    bool movable;
    RefCollisionData<T> collision;
};
static_assert(sizeof(RefParticleCenter<double>) == 992, "ParticleCenter<double> layout");
static_assert(sizeof(RefParticleCenter<float>) == 544, "ParticleCenter<float> layout");

template <typename T>
void to_ref(const IbmSphere& c, RefParticleCenter<T>& r) {
    std::memset(&r, 0, sizeof(r));
    for (int a = 0; a < 3; a++) {
        r.pos[a] = (T)c.pos[a]; r.pos_old[a] = (T)c.pos_old[a];
        r.vel[a] = (T)c.vel[a]; r.vel_old[a] = (T)c.vel_old[a];
        r.w[a] = (T)c.w[a]; r.w_avg[a] = (T)c.w_avg[a]; r.w_old[a] = (T)c.w_old[a]; r.w_pos[a] = (T)c.w_pos[a];
        r.f[a] = (T)c.f[a]; r.f_old[a] = (T)c.f_old[a]; r.M[a] = (T)c.M[a]; r.M_old[a] = (T)c.M_old[a];
        r.dP[a] = (T)c.dP[a]; r.dL[a] = (T)c.dL[a];
    }
    // spheres keep the orientation quaternion set at creation (particle.cpp:738-746)
    r.q_pos[0] = 1; r.q_pos_old[0] = 1;
    for (int a = 0; a < 6; a++) r.I[a] = (T)c.I[a];
    r.S = (T)(4.0 * M_PI * c.radius * c.radius);
    r.radius = (T)c.radius;
    r.volume = (T)c.volume;
    r.density = (T)c.density;
    r.movable = c.movable != 0;
    r.collision.shape = 0;
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) r.collision.semi_axis[k][a] = (T)c.axis[k][a];
    for (int k = 0; k < kMaxColl; k++) {
        r.collision.partner[k] = c.partner[k];
        r.collision.last[k] = c.last[k];
        for (int a = 0; a < 3; a++) r.collision.tang[k][a] = (T)c.tang[k][a];
    }
}

template <typename T>
void from_ref(const RefParticleCenter<T>& r, IbmSphere& c) {
    for (int a = 0; a < 3; a++) {
        c.pos[a] = r.pos[a]; c.pos_old[a] = r.pos_old[a];
        c.vel[a] = r.vel[a]; c.vel_old[a] = r.vel_old[a];
        c.w[a] = r.w[a]; c.w_avg[a] = r.w_avg[a]; c.w_old[a] = r.w_old[a]; c.w_pos[a] = r.w_pos[a];
        c.f[a] = r.f[a]; c.f_old[a] = r.f_old[a]; c.M[a] = r.M[a]; c.M_old[a] = r.M_old[a];
        c.dP[a] = r.dP[a]; c.dL[a] = r.dL[a];
    }
    for (int a = 0; a < 6; a++) c.I[a] = r.I[a];
    c.radius = r.radius;
    c.volume = r.volume;
    c.density = r.density;
    c.movable = r.movable ? 1 : 0;
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) c.axis[k][a] = r.collision.semi_axis[k][a];
    for (int k = 0; k < kMaxColl; k++) {
        c.partner[k] = r.collision.partner[k];
        c.last[k] = r.collision.last[k];
        for (int a = 0; a < 3; a++) c.tang[k][a] = r.collision.tang[k][a];
    }
}

int download(vlbm_ctx* h, const double* dev, double* host, size_t n) {
    if (n) HIP_TRY(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, h->stream));
    return sync(h);
}

int upload(vlbm_ctx* h, double* dev, const double* host, size_t n) {
    if (n) HIP_TRY(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, h->stream));
    return sync(h);
}

// Particle centres, node arrays and the force field (simCheckpoint.cu:252-290).
int ibm_write(vlbm_ctx* h, const std::string& pre, int tb, int border) {
    IbmState& s = *h->ibm;
    const IbmParams& P = s.P;
    const size_t ns = (size_t)P.n_spheres, nn = (size_t)P.n_nodes;
    const size_t plane = (size_t)h->d.nx * h->d.ny;
    int rc;
    std::vector<IbmSphere> sp(ns);
    if (ns) HIP_TRY(hipMemcpyAsync(sp.data(), P.sph, ns * sizeof(IbmSphere), hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    if (tb == 8) {
        std::vector<RefParticleCenter<double>> r(ns);
        for (size_t k = 0; k < ns; k++) to_ref(sp[k], r[k]);
        rc = write_bytes(pre + "IBM_particles_centers.bin", r.data(), ns * sizeof(r[0]));
    } else {
        std::vector<RefParticleCenter<float>> r(ns);
        for (size_t k = 0; k < ns; k++) to_ref(sp[k], r[k]);
        rc = write_bytes(pre + "IBM_particles_centers.bin", r.data(), ns * sizeof(r[0]));
    }
    if (rc) return rc;
    std::vector<int> pid(nn);
    if (nn) HIP_TRY(hipMemcpyAsync(pid.data(), P.pid, nn * 4, hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    std::vector<uint32_t> idx(pid.begin(), pid.end());
    if ((rc = write_bytes(pre + "IBM_nodes_centers_idx.bin", idx.data(), nn * 4))) return rc;
    // node velocities are never written by the sphere path: zero
    const std::vector<double> zero(nn, 0.0);
    double* const groups[3][3] = {{P.px, P.py, P.pz}, {P.pfx, P.pfy, P.pfz}, {P.pdx, P.pdy, P.pdz}};
    const char* gnames[5] = {"IBM_nodes_pos", "IBM_nodes_vel", "IBM_nodes_vel_old", "IBM_nodes_f", "IBM_nodes_deltaF"};
    const char* axes[3] = {"/x.bin", "/y.bin", "/z.bin"};
    std::vector<double> v(nn);
    for (int g = 0; g < 5; g++) {
        if ((rc = make_dir(pre + gnames[g]))) return rc;
        for (int a = 0; a < 3; a++) {
            const double* src = zero.data();
            if (g == 0 || g >= 3) {
                if ((rc = download(h, groups[g == 0 ? 0 : g - 2][a], v.data(), nn))) return rc;
                src = v.data();
            }
            if ((rc = write_values(pre + gnames[g] + axes[a], src, nn, tb))) return rc;
        }
    }
    if ((rc = download(h, P.pS, v.data(), nn))) return rc;
    if ((rc = write_values(pre + "IBM_nodes_S.bin", v.data(), nn, tb))) return rc;
    // force field of the next step, bordered like the macroscopics
    std::vector<double> f(3 * h->n);
    if ((rc = vlbm_ibm_get_force(h, f.data()))) return rc;
    if ((rc = make_dir(pre + "f"))) return rc;
    for (int a = 0; a < 3; a++) {
        const std::vector<double> b = bordered(f.data() + a * h->n, plane, h->n, border);
        if ((rc = write_values(pre + "f" + axes[a], b.data(), b.size(), tb))) return rc;
    }
    return VLBM_OK;
}

int ibm_read(vlbm_ctx* h, const std::string& pre, int tb, int border) {
    IbmState& s = *h->ibm;
    IbmParams& P = s.P;
    const size_t ns = (size_t)P.n_spheres, nn = (size_t)P.n_nodes;
    const size_t plane = (size_t)h->d.nx * h->d.ny;
    int rc;
    std::vector<IbmSphere> sp(ns);
    if (ns) HIP_TRY(hipMemcpyAsync(sp.data(), P.sph, ns * sizeof(IbmSphere), hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    if (tb == 8) {
        std::vector<RefParticleCenter<double>> r(ns);
        if ((rc = read_bytes(pre + "IBM_particles_centers.bin", r.data(), ns * sizeof(r[0])))) return rc;
        for (size_t k = 0; k < ns; k++) from_ref(r[k], sp[k]);
    } else {
        std::vector<RefParticleCenter<float>> r(ns);
        if ((rc = read_bytes(pre + "IBM_particles_centers.bin", r.data(), ns * sizeof(r[0])))) return rc;
        for (size_t k = 0; k < ns; k++) from_ref(r[k], sp[k]);
    }
    std::vector<uint32_t> idx(nn);
    if ((rc = read_bytes(pre + "IBM_nodes_centers_idx.bin", idx.data(), nn * 4))) return rc;
    std::vector<int> pid(nn);
    if (nn) HIP_TRY(hipMemcpyAsync(pid.data(), P.pid, nn * 4, hipMemcpyDeviceToHost, h->stream));
    if ((rc = sync(h))) return rc;
    for (size_t k = 0; k < nn; k++)
        if (idx[k] != (uint32_t)pid[k])
            return fail(VLBM_ERR_INVALID, "checkpoint node " + std::to_string(k) + " belongs to sphere " +
                                              std::to_string(idx[k]) + ", the handle's to sphere " + std::to_string(pid[k]));
    if (ns) HIP_TRY(hipMemcpyAsync(P.sph, sp.data(), ns * sizeof(IbmSphere), hipMemcpyHostToDevice, h->stream));
    if ((rc = sync(h))) return rc;
    double* const groups[3][3] = {{P.px, P.py, P.pz}, {P.pfx, P.pfy, P.pfz}, {P.pdx, P.pdy, P.pdz}};
    const char* gnames[3] = {"IBM_nodes_pos", "IBM_nodes_f", "IBM_nodes_deltaF"};
    const char* axes[3] = {"/x.bin", "/y.bin", "/z.bin"};
    std::vector<double> v(nn);
    for (int g = 0; g < 3; g++)
        for (int a = 0; a < 3; a++) {
            if ((rc = read_values(pre + gnames[g] + axes[a], v.data(), nn, tb))) return rc;
            if ((rc = upload(h, groups[g][a], v.data(), nn))) return rc;
        }
    if ((rc = read_values(pre + "IBM_nodes_S.bin", v.data(), nn, tb))) return rc;
    if ((rc = upload(h, P.pS, v.data(), nn))) return rc;
    // force field: the tiles holding a value other than the constant force
    // are flagged, so that the next step reads them; the tile list stays
    // empty until the next stage builds it
    std::vector<uint32_t> flags(P.n_tiles, 0u);
    const size_t nb = h->n + 2 * (size_t)border * plane;
    std::vector<double> b(nb);
    const double f0[3] = {h->d.fx, h->d.fy, h->d.fz};
    for (int a = 0; a < 3; a++) {
        if ((rc = read_values(pre + "f" + axes[a], b.data(), nb, tb))) return rc;
        const double* f = b.data() + (size_t)border * plane;
        if ((rc = upload_field(h, s.force[a], f))) return rc;
        const double f0t = h->tbytes == 8 ? f0[a] : (double)(float)f0[a];
        size_t k = 0;
        for (int z = 0; z < h->d.nz; z++)
            for (int y = 0; y < h->d.ny; y++)
                for (int x = 0; x < h->d.nx; x++, k++) {
                    const double fv = h->tbytes == 8 ? f[k] : (double)(float)f[k];
                    if (fv != f0t) {
                        const uint32_t t = tile_index(x, y, z, P.tiles[0], P.tiles[1]);
                        flags[t] = 1;
                    }
                }
    }
    if (!flags.empty())
        HIP_TRY(hipMemcpyAsync(P.tile_flag, flags.data(), flags.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemsetAsync(P.tile_count, 0, 4, h->stream));
    return sync(h);
}

}  // namespace

extern "C" {

int vlbm_save_checkpoint(vlbm_h h, const char* path, const char* id, int step) {
    if (!h || !path || !id) return fail(VLBM_ERR_INVALID, "NULL argument");
    const vlbm_desc& d = h->d;
    const int tb = d.precision == VLBM_FP64 ? 8 : 4;
    const int border = h->ibm ? kBorderIbm : 0;
    const size_t plane = (size_t)d.nx * d.ny;
    const std::string dir = std::string(path) + "/" + id + "/checkpoint";
    int rc;
    if ((rc = make_dir(std::string(path) + "/" + id)) || (rc = make_dir(dir))) return rc;
    const std::string pre = dir + "/0_";
    const int32_t s32 = step;
    if ((rc = write_bytes(pre + "curr_step.bin", &s32, 4))) return rc;
    // populations: the post-stream state the next step collides; the
    // reference's popAux is overwritten by that step before it is read, and
    // gets the same values here, except the rheometer's plate slots fed by an
    // unused plane (no node streams into them; they hold the plate node's
    // post-collision populations)
    {
        std::vector<double> pop((size_t)d.q * plane * (d.nz + 1));
        if ((rc = vlbm_get_pops(h, pop.data()))) return rc;
        if ((rc = write_values(pre + "pop.bin", pop.data(), pop.size(), tb))) return rc;
        if ((rc = plate_aux(h, pop.data()))) return rc;
        if ((rc = write_values(pre + "popAux.bin", pop.data(), pop.size(), tb))) return rc;
    }
    // saved macroscopics (with IBM: those of the last stage on every node)
    if (h->ibm && (rc = vlbm_compute_macroscopics(h))) return rc;
    std::vector<double> buf(h->n);
    const char* names[4] = {"rho.bin", "u/x.bin", "u/y.bin", "u/z.bin"};
    if ((rc = make_dir(pre + "u"))) return rc;
    for (int k = 0; k < 4; k++) {
        if ((rc = fetch_field(h, h->macr[k], buf.data()))) return rc;
        const std::vector<double> b = bordered(buf.data(), plane, h->n, border);
        if ((rc = write_values(pre + names[k], b.data(), b.size(), tb))) return rc;
    }
    if (d.rheology != VLBM_NEWTONIAN) {
        // the power law's omega is its state; Bingham's the saved output
        const void* om = d.rheology == VLBM_POWERLAW ? h->omega_state : h->macr[4];
        if ((rc = fetch_field(h, om, buf.data()))) return rc;
        if ((rc = write_values(pre + "omega.bin", buf.data(), h->n, tb))) return rc;
    }
    if (h->ibm && (rc = ibm_write(h, pre, tb, border))) return rc;
    return VLBM_OK;
}

int vlbm_load_checkpoint(vlbm_h h, const char* path, const char* id, int* step) {
    if (!h || !path || !id) return fail(VLBM_ERR_INVALID, "NULL argument");
    const vlbm_desc& d = h->d;
    const int tb = d.precision == VLBM_FP64 ? 8 : 4;
    const int border = h->ibm ? kBorderIbm : 0;
    const size_t plane = (size_t)d.nx * d.ny;
    const std::string pre = std::string(path) + "/" + id + "/checkpoint/0_";
    int32_t s32 = 0;
    int rc;
    if ((rc = read_bytes(pre + "curr_step.bin", &s32, 4))) return rc;
    {
        std::vector<double> pop((size_t)d.q * plane * (d.nz + 1));
        if ((rc = read_values(pre + "pop.bin", pop.data(), pop.size(), tb))) return rc;
        if ((rc = vlbm_load_pops(h, pop.data()))) return rc;
        if (d.builder == VLBM_BUILDER_RHEOMETER) {
            if ((rc = read_values(pre + "popAux.bin", pop.data(), pop.size(), tb))) return rc;
            if ((rc = plate_aux_stage(h, pop.data()))) return rc;
            h->resumed = s32 > 0;
        }
    }
    const size_t nb = h->n + 2 * (size_t)border * plane;
    std::vector<double> b(nb);
    const char* names[4] = {"rho.bin", "u/x.bin", "u/y.bin", "u/z.bin"};
    for (int k = 0; k < 4; k++) {
        if ((rc = read_values(pre + names[k], b.data(), nb, tb))) return rc;
        const double* inner = b.data() + (size_t)border * plane;
        if ((rc = upload_field(h, h->macr[k], inner))) return rc;
        // the residual's reference velocity starts at the loaded one
        if (k > 0 && (rc = upload_field(h, h->uold[k - 1], inner))) return rc;
    }
    if (d.rheology != VLBM_NEWTONIAN) {
        std::vector<double> om(h->n);
        if ((rc = read_values(pre + "omega.bin", om.data(), h->n, tb))) return rc;
        if ((rc = upload_field(h, h->macr[4], om.data()))) return rc;
        if (d.rheology == VLBM_POWERLAW && (rc = upload_field(h, h->omega_state, om.data()))) return rc;
    }
    if (h->ibm && (rc = ibm_read(h, pre, tb, border))) return rc;
    h->steps = s32;
    if (step) *step = s32;
    return VLBM_OK;
}

}  // extern "C"
