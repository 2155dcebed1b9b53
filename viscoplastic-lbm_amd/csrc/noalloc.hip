// vlbm_get_pops_noalloc: the population download of vlbm_get_pops without
// any device allocation, for callers that cannot allocate device memory
// (serving loops that snapshot states between steps, harnesses that forbid
// hipMalloc). vlbm_get_pops stages the populations in a buffer it allocates
// per call (and, for the rheometer, a materialized copy of the state); this
// entry point runs the same kernels over chunks of nodes whose values go
// through one fixed per-device array instead.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/vlbm.h"
#include "ctx.h"
#include "lattice.h"

using namespace vlbm;
using namespace vlbm::detail;

namespace {

// 8 MiB of doubles per device, held until the process ends. Every call on a
// device shares it: calls that overlap on one device race.
constexpr size_t kScratchDoubles = size_t(1) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` per device. The array's name in host code is a
// host-side stand-in, not a device address: kernels and copies take these.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int device, double** out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(VLBM_ERR_UNSUPPORTED, "device index " + std::to_string(device) + " is beyond the " +
                                              std::to_string(kMaxDevices) + " devices the scratch table holds");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(VLBM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (double*)ptr;
    }
    *out = g_scratch_ptr[device];
    return VLBM_OK;
}

// Makes the handle's device current for the call and restores the caller's.
struct DeviceScope {
    int prev = -1;
    bool ok = false;
    explicit DeviceScope(int device) {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = prev == device || hipSetDevice(device) == hipSuccess;
    }
    ~DeviceScope() {
        if (ok && prev >= 0) (void)hipSetDevice(prev);
    }
};

// One chunk of the aux kernel (kAuxMaterialize to doubles, or kAuxGhost)
// into the scratch: Q planes of `count` values.
int aux_chunk(vlbm_h h, double* buf, int mode, uint32_t base, uint32_t count) {
    const vlbm_desc& d = h->d;
    return with_params(h, [&](void* pv, auto& p) {
        p.src = state_src(h);
        p.dst = buf;
        p.aux_mode = mode;
        p.aux_f64 = 1;
        p.aux_base = base;
        p.aux_count = count;
        p.direct = h->direct ? 1 : 0;
        return check_launch(launch_aux(d.q, d.precision, d.rheology, h->mask_geom, pv, h->stream, state_rep(h)),
                            "population chunk kernel");
    });
}

}  // namespace

extern "C" int vlbm_get_pops_noalloc(vlbm_h h, double* pop) {
    if (!h || !pop) return fail(VLBM_ERR_INVALID, "handle or output NULL");
    const vlbm_desc& d = h->d;
    // nodes per chunk: Q planes of them fill at most the scratch (55,040
    // nodes for D3Q19, 38,656 for D3Q27), so lattices of any size pass
    const uint32_t chunk = (uint32_t)(kScratchDoubles / (size_t)d.q / 256 * 256);
    DeviceScope scope(h->device);
    if (!scope.ok) return fail(VLBM_ERR_DEVICE, "cannot make the handle's device current");
    double* buf = nullptr;
    int rc = scratch_on(h->device, &buf);
    if (rc) return rc;
    const size_t plane = (size_t)d.nx * d.ny;
    const size_t ref_plane_stride = plane * (d.nz + 1);     // [Q][NZ+1][NY][NX]
    auto copy_out = [&](size_t dst_off, uint32_t count) -> int {
        for (int i = 0; i < d.q; i++)
            HIP_TRY(hipMemcpyAsync(pop + (size_t)i * ref_plane_stride + dst_off, buf + (size_t)i * count,
                                   (size_t)count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        return VLBM_OK;
    };
    // the lattice's nodes: post-stream, post-BC populations (as vlbm_get_pops)
    for (size_t base = 0; base < h->n; base += chunk) {
        const uint32_t count = (uint32_t)std::min<size_t>(chunk, h->n - base);
        if ((rc = aux_chunk(h, buf, kAuxMaterialize, (uint32_t)base, count))) return rc;
        if ((rc = copy_out(base, count))) return rc;
    }
    // the extra plane z = NZ: what the last step pushed across the z ends
    if (!h->direct) {
        for (size_t base = 0; base < plane; base += chunk) {
            const uint32_t count = (uint32_t)std::min<size_t>(chunk, plane - base);
            if ((rc = aux_chunk(h, buf, kAuxGhost, (uint32_t)base, count))) return rc;
            if ((rc = copy_out(h->n + base, count))) return rc;
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int i = 0; i < d.q; i++) {
        const double w = pop_weight(d.q, i);
        double* dstp = pop + (size_t)i * ref_plane_stride;
        if (h->direct) {
            for (size_t k = 0; k < h->n; k++) dstp[k] += w;
            for (size_t k = 0; k < plane; k++) dstp[h->n + k] = w;   // the rest state
        } else {
            for (size_t k = 0; k < h->n + plane; k++) dstp[k] += w;
        }
    }
    return VLBM_OK;
}
