// Serving engine: many independent box-builder lattices resident on one GPU,
// admitted by an HBM budget and stepped together by the batched kernel
// (declarations in include/vlbm.h, kernels in kernels_batch.hip).
//
// Life of a job: submitted (pending) -> admitted when a slot and a piece of
// the arena are free (rest-state fill + the direct first step of the
// single-lattice path) -> advanced in chunks of K steps, K even, by one
// replay of a captured K-launch sequence -> residual check every
// check_every of its own steps -> retired (macroscopics copied to host, arena
// piece returned) on residual < resid_max, max_steps, or a non-finite /
// negative mean density. A job's arithmetic does not depend on what else is
// resident, so its result is the same alone and in any mix.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "batch.h"
#include "ctx.h"

using namespace vlbm;
using namespace vlbm::detail;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Job {
    vlbm_desc d;
    double resid_max = 0;
    long long max_steps = 0;
    int state = VLBM_JOB_PENDING;
    long long steps = 0;
    double residual = 1.0, mean_rho = 0;
    double t_submit = 0, t_start = -1, t_end = -1;
    size_t off = 0, bytes = 0;
    uint32_t words[27];
    int phase = 0;
    long long next_check = 0;
    bool has_result = false;
    std::vector<double> res[4];
};

}  // namespace

struct vlbm_batch {
    vlbm_batch_desc d;
    std::string cache_path;
    int tbytes = 4;
    hipStream_t stream = nullptr;
    char* arena = nullptr;
    size_t arena_bytes = 0;
    std::map<size_t, size_t> free_list;       // offset -> size, coalesced
    std::vector<Job> jobs;
    std::deque<int> pending;
    std::vector<int> running;                  // admission order = table order
    void* d_jobs = nullptr;                    // BatchJob<T>[max_jobs]
    uint32_t* d_prefix = nullptr;              // [max_jobs + 1]
    uint16_t* d_unit_job = nullptr;            // job slot of every work unit
    size_t unit_cap = 0;                       // capacity of d_unit_job (bounded by the arena)
    bool units_dirty = true;                   // running set changed since the last upload
    int units_block = 0;                       // block size the unit -> job table was built for
    size_t resident_nodes = 0;
    uint32_t* d_njobs = nullptr;
    double* d_partials = nullptr;
    size_t partials_cap = 0;                   // work units
    double* d_sums = nullptr;                  // [3 * max_jobs]
    bool table_dirty = true;
    std::vector<uint32_t> prefix;              // host copy for the current block
    int K = 0, block = 0, grid = 0;
    bool tuned = false;
    std::string tuned_key;                     // resident-node bucket the launch shape was chosen for
    bool fixed_shape = false;                  // K, block and grid given by the caller
    std::map<std::string, std::array<int, 3>> shapes;   // launch shapes chosen in this batch, by key
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_K = 0, graph_block = 0, graph_grid = 0, graph_rheo = -1;
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int sync(vlbm_batch* b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    return VLBM_OK;
}

// ---- arena ------------------------------------------------------------------
bool arena_alloc(vlbm_batch* b, size_t bytes, size_t* off) {
    auto best = b->free_list.end();
    for (auto it = b->free_list.begin(); it != b->free_list.end(); ++it)
        if (it->second >= bytes && (best == b->free_list.end() || it->second < best->second)) best = it;
    if (best == b->free_list.end()) return false;
    *off = best->first;
    const size_t rest = best->second - bytes;
    const size_t at = best->first + bytes;
    b->free_list.erase(best);
    if (rest) b->free_list[at] = rest;
    return true;
}

void arena_free(vlbm_batch* b, size_t off, size_t bytes) {
    auto it = b->free_list.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != b->free_list.end() && it->first + it->second == next->first) {
        it->second += next->second;
        b->free_list.erase(next);
    }
    if (it != b->free_list.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            b->free_list.erase(it);
        }
    }
}

// ---- job layout: lattice A, lattice B, u_old x/y/z, power-law omega ---------
struct Layout {
    size_t lat, uold, omega, total;
};

Layout layout_of(const vlbm_desc& d, int tbytes) {
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    Layout L;
    L.lat = align_up((size_t)d.q * (size_t)plane_stride(n) * d.precision);
    L.uold = align_up(n * tbytes);
    L.omega = d.rheology == VLBM_POWERLAW ? align_up(n * tbytes) : 0;
    L.total = 2 * L.lat + 3 * L.uold + L.omega;
    return L;
}

template <typename T>
void job_params(vlbm_batch* b, const Job& j, int src, Params<T>& p) {
    fill_params_desc(j.d, j.words, p);
    p.sr = 0;   // the captured step sequences replay fixed parameters: FP16 jobs round to nearest
    const Layout L = layout_of(j.d, b->tbytes);
    char* base = b->arena + j.off;
    char* lat[2] = {base, base + L.lat};
    p.src = lat[src];
    p.dst = lat[1 - src];
    if (L.omega) p.omega_state = (T*)(base + 2 * L.lat + 3 * L.uold);
}

template <typename T>
void job_entry(vlbm_batch* b, const Job& j, BatchJob<T>& e) {
    std::memset(&e, 0, sizeof(e));
    job_params<T>(b, j, 0, e.p[0]);
    job_params<T>(b, j, 1, e.p[1]);
    const Layout L = layout_of(j.d, b->tbytes);
    char* u = b->arena + j.off + 2 * L.lat;
    for (int a = 0; a < 3; a++) e.uold[a] = (T*)(u + a * L.uold);
    e.rheo = j.d.rheology;
    e.phase = j.phase;
    e.due = 0;
}

template <typename T>
int upload_table(vlbm_batch* b, const std::vector<int>& due) {
    const size_t nj = b->running.size();
    std::vector<BatchJob<T>> tab(nj);
    for (size_t k = 0; k < nj; k++) {
        job_entry<T>(b, b->jobs[b->running[k]], tab[k]);
        tab[k].due = due.empty() ? 0 : due[k];
    }
    b->prefix.assign(nj + 1, 0);
    for (size_t k = 0; k < nj; k++) {
        const Job& j = b->jobs[b->running[k]];
        const uint32_t n = (uint32_t)((size_t)j.d.nx * j.d.ny * j.d.nz);
        b->prefix[k + 1] = b->prefix[k] + (n + b->block - 1) / b->block;
    }
    const uint32_t units = b->prefix[nj];
    if (units > b->unit_cap) return fail(VLBM_ERR_INVALID, "work units exceed the table capacity");
    // the table depends on the running set and on the block size (units of
    // block nodes); the autotuner changes the block between chunks
    if ((b->units_dirty || b->units_block != b->block) && units) {
        std::vector<uint16_t> uj(units);
        for (size_t k = 0; k < nj; k++)
            for (uint32_t w = b->prefix[k]; w < b->prefix[k + 1]; w++) uj[w] = (uint16_t)k;
        HIP_TRY(hipMemcpyAsync(b->d_unit_job, uj.data(), (size_t)units * 2, hipMemcpyHostToDevice, b->stream));
        if (int rc = sync(b)) return rc;
        b->units_block = b->block;
    }
    b->units_dirty = false;
    if (units > b->partials_cap) {
        if (b->d_partials) (void)hipFree(b->d_partials);
        b->d_partials = nullptr;
        const size_t cap = (size_t)units * 2 + 1024;
        HIP_TRY(hipMalloc((void**)&b->d_partials, cap * 3 * sizeof(double)));
        b->partials_cap = cap;
    }
    const uint32_t nj32 = (uint32_t)nj;
    if (nj) HIP_TRY(hipMemcpyAsync(b->d_jobs, tab.data(), nj * sizeof(BatchJob<T>), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_prefix, b->prefix.data(), (nj + 1) * 4, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_njobs, &nj32, 4, hipMemcpyHostToDevice, b->stream));
    return sync(b);   // host staging buffers go out of scope
}

int upload(vlbm_batch* b, const std::vector<int>& due = {}) {
    const int rc = b->tbytes == 8 ? upload_table<double>(b, due) : upload_table<float>(b, due);
    if (!rc) b->table_dirty = false;
    return rc;
}

// rest state + direct first step of the single-lattice path (vlbm_init + the
// first vlbm_step), so the job continues exactly as a plain solve would
int admit(vlbm_batch* b, int id) {
    Job& j = b->jobs[id];
    const Layout L = layout_of(j.d, b->tbytes);
    const uint32_t n = (uint32_t)((size_t)j.d.nx * j.d.ny * j.d.nz);
    char* base = b->arena + j.off;
    if (launch_batch_fill_rest(j.d.q, j.d.precision, base, plane_stride(n), n, j.d.rho0, b->stream))
        return check_launch(-1, "rest-state kernel");
    HIP_TRY(hipMemsetAsync(base + 2 * L.lat, 0, 3 * L.uold + L.omega, b->stream));
    int rc;
    if (b->tbytes == 8) {
        Params<double> p;
        job_params<double>(b, j, 0, p);
        rc = launch_step(j.d.q, j.d.precision, j.d.rheology, false, true, &p, b->stream);
    } else {
        Params<float> p;
        job_params<float>(b, j, 0, p);
        rc = launch_step(j.d.q, j.d.precision, j.d.rheology, false, true, &p, b->stream);
    }
    if (rc) return check_launch(rc, "first step");
    j.phase = 1;
    j.steps = 1;
    j.next_check = 1 + b->d.check_every;
    j.state = VLBM_JOB_RUNNING;
    j.t_start = b->now();
    b->running.push_back(id);
    b->resident_nodes += n;
    b->table_dirty = true;
    b->units_dirty = true;
    return VLBM_OK;
}

// final macroscopics into the idle lattice of the job, then to the host
int retire(vlbm_batch* b, int id, int state) {
    Job& j = b->jobs[id];
    const Layout L = layout_of(j.d, b->tbytes);
    const size_t n = (size_t)j.d.nx * j.d.ny * j.d.nz;
    char* base = b->arena + j.off;
    char* idle = base + (j.phase ? 0 : L.lat);
    char* uold = base + 2 * L.lat;
    const size_t fb = n * b->tbytes;
    int rc;
    if (b->tbytes == 8) {
        Params<double> p;
        job_params<double>(b, j, j.phase, p);
        p.aux_mode = kAuxMoments;
        p.m_rho = (double*)idle; p.m_ux = (double*)(idle + fb); p.m_uy = (double*)(idle + 2 * fb);
        p.m_uz = (double*)(idle + 3 * fb); p.m_omega = (double*)uold;
        rc = launch_aux(j.d.q, j.d.precision, j.d.rheology, false, &p, b->stream);
    } else {
        Params<float> p;
        job_params<float>(b, j, j.phase, p);
        p.aux_mode = kAuxMoments;
        p.m_rho = (float*)idle; p.m_ux = (float*)(idle + fb); p.m_uy = (float*)(idle + 2 * fb);
        p.m_uz = (float*)(idle + 3 * fb); p.m_omega = (float*)uold;
        rc = launch_aux(j.d.q, j.d.precision, j.d.rheology, false, &p, b->stream);
    }
    if (rc) return check_launch(rc, "final moments");
    std::vector<char> host(4 * fb);
    HIP_TRY(hipMemcpyAsync(host.data(), idle, 4 * fb, hipMemcpyDeviceToHost, b->stream));
    if ((rc = sync(b))) return rc;
    for (int f = 0; f < 4; f++) {
        j.res[f].resize(n);
        if (b->tbytes == 8) std::memcpy(j.res[f].data(), host.data() + f * fb, fb);
        else {
            const float* src = (const float*)(host.data() + f * fb);
            for (size_t k = 0; k < n; k++) j.res[f][k] = src[k];
        }
    }
    j.has_result = true;
    j.state = state;
    j.t_end = b->now();
    arena_free(b, j.off, j.bytes);
    b->running.erase(std::find(b->running.begin(), b->running.end(), id));
    b->resident_nodes -= n;
    b->table_dirty = true;
    b->units_dirty = true;
    return VLBM_OK;
}

int admit_pending(vlbm_batch* b) {
    // arrival order, a later job may pass one that does not fit yet
    for (auto it = b->pending.begin(); it != b->pending.end();) {
        if ((int)b->running.size() >= b->d.max_jobs) break;
        Job& j = b->jobs[*it];
        const size_t nodes = (size_t)j.d.nx * j.d.ny * j.d.nz;
        // node budget: enough resident work to fill the GPU, not more, so that
        // jobs finish in arrival order instead of all together
        if (b->d.max_resident_nodes && !b->running.empty() && b->resident_nodes + nodes > b->d.max_resident_nodes) {
            ++it;
            continue;
        }
        size_t off;
        if (!arena_alloc(b, j.bytes, &off)) { ++it; continue; }
        j.off = off;
        const int id = *it;
        it = b->pending.erase(it);
        const int rc = admit(b, id);
        if (rc) return rc;
    }
    return VLBM_OK;
}

// rheology shared by every running job (the step kernel is specialised to
// it), or 3 for a mix
int common_rheo(const vlbm_batch* b) {
    int r = -1;
    for (int id : b->running) {
        const int jr = b->jobs[id].d.rheology;
        if (r < 0) r = jr;
        else if (r != jr) return 3;
    }
    return r < 0 ? 0 : r;
}

int build_graph(vlbm_batch* b) {
    if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = VLBM_OK;
    const int rheo = common_rheo(b);
    for (int k = 0; k < b->K && !rc; k++)
        if (launch_batch_step(b->d.q, b->d.precision, b->block, rheo, (uint32_t)b->grid, b->d_jobs, b->d_prefix,
                              b->d_unit_job, b->d_njobs, k, b->stream))
            rc = fail(VLBM_ERR_DEVICE, "batched step launch during capture failed");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(VLBM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    b->graph_K = b->K; b->graph_block = b->block; b->graph_grid = b->grid; b->graph_rheo = rheo;
    return VLBM_OK;
}

// K steps of every running job
int run_chunk(vlbm_batch* b) {
    if (b->d.use_graph) {
        if (!b->graph_exec || b->graph_K != b->K || b->graph_block != b->block || b->graph_grid != b->grid ||
            b->graph_rheo != common_rheo(b)) {
            const int rc = build_graph(b);
            if (rc) return rc;
        }
        HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        const int rheo = common_rheo(b);
        for (int k = 0; k < b->K; k++)
            if (launch_batch_step(b->d.q, b->d.precision, b->block, rheo, (uint32_t)b->grid, b->d_jobs, b->d_prefix,
                                  b->d_unit_job, b->d_njobs, k, b->stream))
                return check_launch(-1, "batched step");
    }
    for (int id : b->running) b->jobs[id].steps += b->K;
    return VLBM_OK;
}

// residual checks of the jobs at a check point; retirements
int check_jobs(vlbm_batch* b) {
    std::vector<int> due(b->running.size(), 0);
    bool any = false;
    for (size_t k = 0; k < b->running.size(); k++) {
        const Job& j = b->jobs[b->running[k]];
        if (j.steps >= j.next_check || j.steps >= j.max_steps) { due[k] = 1; any = true; }
    }
    if (!any) return VLBM_OK;
    int rc = upload(b, due);
    if (rc) return rc;
    const uint32_t nj = (uint32_t)b->running.size();
    if (launch_batch_check(b->d.q, b->d.precision, b->block, (uint32_t)b->grid, b->d_jobs, b->d_prefix,
                           b->d_unit_job, b->d_njobs, nj, b->d_partials, b->d_sums, b->stream))
        return check_launch(-1, "residual check");
    std::vector<double> sums(3 * (size_t)nj);
    HIP_TRY(hipMemcpyAsync(sums.data(), b->d_sums, sums.size() * 8, hipMemcpyDeviceToHost, b->stream));
    if ((rc = sync(b))) return rc;
    std::vector<std::pair<int, int>> done;
    for (size_t k = 0; k < nj; k++) {
        if (!due[k]) continue;
        Job& j = b->jobs[b->running[k]];
        const double n = (double)j.d.nx * j.d.ny * j.d.nz;
        // treatData (src/CUDA/treatData.cpp:65-68): residual 1 when sum |u| is 0
        j.residual = sums[3 * k + 1] != 0 ? sums[3 * k] / sums[3 * k + 1] : 1.0;
        j.mean_rho = sums[3 * k + 2] / n;
        j.next_check += b->d.check_every;
        if (!std::isfinite(j.mean_rho) || !std::isfinite(j.residual) || j.mean_rho < 0)
            done.emplace_back(b->running[k], VLBM_JOB_FAILED);
        else if (j.residual < j.resid_max)
            done.emplace_back(b->running[k], VLBM_JOB_DONE);
        else if (j.steps >= j.max_steps)
            done.emplace_back(b->running[k], VLBM_JOB_MAX_STEPS);
    }
    b->table_dirty = true;   // due flags
    for (auto& pr : done)
        if ((rc = retire(b, pr.first, pr.second))) return rc;
    return VLBM_OK;
}

// ---- autotuner ----------------------------------------------------------------
std::string tune_key(const vlbm_batch* b, size_t nodes) {
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    int bucket = 0;
    while ((size_t(1) << (bucket + 1)) <= nodes) bucket++;
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%s q%d s%d check%d nodes2^%d", arch.c_str(), b->d.q, b->d.precision,
                  b->d.check_every, bucket);
    return buf;
}

bool cache_lookup(const vlbm_batch* b, const std::string& key, int* K, int* block, int* grid) {
    if (b->cache_path.empty()) return false;
    FILE* f = std::fopen(b->cache_path.c_str(), "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        const char* bar = std::strchr(line, '|');
        if (!bar || std::string(line, bar - line) != key) continue;
        if (std::sscanf(bar + 1, "%d %d %d", K, block, grid) == 3) found = true;
    }
    std::fclose(f);
    return found;
}

void cache_store(const vlbm_batch* b, const std::string& key, int K, int block, int grid) {
    if (b->cache_path.empty()) return;
    FILE* f = std::fopen(b->cache_path.c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%s|%d %d %d\n", key.c_str(), K, block, grid);
    std::fclose(f);
}

// even step counts per launch that divide check_every (so every job's checks
// fall on chunk boundaries), largest first, at most 64
std::vector<int> k_candidates(int check_every) {
    std::vector<int> out;
    for (int k = 64; k >= 2; k -= 2)
        if (check_every % k == 0) out.push_back(k);
    return out;
}

// Picks launch shape and steps per launch by timing real chunks of the
// running jobs (they advance; no work is thrown away), then persists the choice.
int autotune(vlbm_batch* b) {
    size_t nodes = 0;
    for (int id : b->running) nodes += (size_t)b->jobs[id].d.nx * b->jobs[id].d.ny * b->jobs[id].d.nz;
    const std::string key = tune_key(b, nodes);
    b->tuned_key = key;
    int K = b->d.steps_per_launch, block = b->d.block, grid = b->d.grid;
    int cK, cb, cg;
    const auto seen = b->shapes.find(key);
    const bool known = seen != b->shapes.end();
    if (known) { cK = seen->second[0]; cb = seen->second[1]; cg = seen->second[2]; }
    if (known || cache_lookup(b, key, &cK, &cb, &cg)) {
        b->shapes[key] = {cK, cb, cg};
        b->K = K ? K : cK; b->block = block ? block : cb; b->grid = grid ? grid : cg;
        b->tuned = true;
        b->table_dirty = true;
        return VLBM_OK;
    }
    const std::vector<int> Ks = K ? std::vector<int>{K} : std::vector<int>{k_candidates(b->d.check_every).front()};
    const std::vector<int> blocks = block ? std::vector<int>{block} : std::vector<int>{128, 256};
    // up to one work unit per block for 2^24 resident nodes (surplus blocks exit at once)
    const std::vector<int> grids = grid ? std::vector<int>{grid}
                                        : std::vector<int>{1024, 2048, 4096, 8192, 16384, 32768, 65536};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 0;
    int bK = Ks[0], bb = blocks[0], bg = grids[0];
    int rc = VLBM_OK;
    // Each candidate runs a warm chunk (graph build) and a timed chunk of the
    // resident jobs, with the jobs' residual checks after each chunk as in
    // vlbm_batch_advance (a job's steps and results do not depend on the
    // tuning); candidates compare by node updates per second.
    auto resident = [&]() {
        double n = 0;
        for (int id : b->running) n += (double)b->jobs[id].d.nx * b->jobs[id].d.ny * b->jobs[id].d.nz;
        return n;
    };
    for (int bl : blocks) {
        for (int g : grids) {
            if (b->running.empty()) break;
            b->K = Ks[0]; b->block = bl; b->grid = g;
            if ((rc = upload(b))) break;
            if ((rc = run_chunk(b))) break;
            if ((rc = check_jobs(b))) break;
            if (b->running.empty()) break;
            if (b->table_dirty && (rc = upload(b))) break;
            const double nodes = resident();
            (void)hipEventRecord(e0, b->stream);
            if ((rc = run_chunk(b))) break;
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if ((rc = check_jobs(b))) break;
            const double rate = ms > 0 ? nodes * b->K / ms : 0;
            if (rate > best) { best = rate; bb = bl; bg = g; }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->K = bK; b->block = bb; b->grid = bg;
    b->tuned = true;
    b->table_dirty = true;
    cache_store(b, key, bK, bb, bg);
    b->shapes[key] = {bK, bb, bg};
    return VLBM_OK;
}

// The resident nodes are in another power-of-two bucket than the one the
// launch shape was chosen for.
bool retune_due(const vlbm_batch* b) {
    size_t nodes = 0;
    for (int id : b->running) nodes += (size_t)b->jobs[id].d.nx * b->jobs[id].d.ny * b->jobs[id].d.nz;
    return tune_key(b, nodes) != b->tuned_key;
}

int need(vlbm_batch* b) {
    if (!b) return fail(VLBM_ERR_INVALID, "batch handle is NULL");
    return VLBM_OK;
}

}  // namespace

extern "C" {

void vlbm_batch_desc_defaults(vlbm_batch_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->hbm_budget = (size_t)32 << 30;
    d->max_jobs = 256;
    d->q = 19;
    d->precision = VLBM_FP32;
    d->check_every = 1000;
    d->steps_per_launch = 0;
    d->block = 0;
    d->grid = 0;
    d->use_graph = 1;
    d->tune_cache = nullptr;
    d->max_resident_nodes = 0;
}

int vlbm_batch_create(const vlbm_batch_desc* d, vlbm_batch_h* out) {
    if (!d || !out) return fail(VLBM_ERR_INVALID, "descriptor or output NULL");
    *out = nullptr;
    if (d->q != 19 && d->q != 27) return fail(VLBM_ERR_INVALID, "velocity set must be D3Q19 or D3Q27");
    if (d->precision != VLBM_FP64 && d->precision != VLBM_FP32 && d->precision != VLBM_FP16)
        return fail(VLBM_ERR_INVALID, "precision must be 8, 4 or 2 bytes");
    if (d->max_jobs < 1) return fail(VLBM_ERR_INVALID, "max_jobs must be >= 1");
    if (d->check_every < 2 || d->check_every % 2) return fail(VLBM_ERR_INVALID, "check_every must be even and >= 2");
    if (d->steps_per_launch && (d->steps_per_launch % 2 || d->check_every % d->steps_per_launch))
        return fail(VLBM_ERR_INVALID, "steps_per_launch must be even and divide check_every");
    if (d->block && d->block != 128 && d->block != 256) return fail(VLBM_ERR_INVALID, "block must be 128 or 256");
    if (d->grid < 0) return fail(VLBM_ERR_INVALID, "grid must be >= 0");
    if (d->hbm_budget < kAlign) return fail(VLBM_ERR_INVALID, "hbm_budget too small");
    vlbm_batch* b = new vlbm_batch;
    b->d = *d;
    if (d->tune_cache) b->cache_path = d->tune_cache;
    b->d.tune_cache = nullptr;
    b->tbytes = d->precision == VLBM_FP64 ? 8 : 4;
    b->t0 = std::chrono::steady_clock::now();
    auto bail = [&](int code, const std::string& m) {
        vlbm_batch_destroy(b);
        return fail(code, m);
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(VLBM_ERR_DEVICE, "hipStreamCreate failed");
    b->arena_bytes = d->hbm_budget / kAlign * kAlign;
    if (hipMalloc((void**)&b->arena, b->arena_bytes) != hipSuccess) {
        b->arena = nullptr;
        return bail(VLBM_ERR_ALLOC, "arena of " + std::to_string(b->arena_bytes) + " bytes could not be allocated");
    }
    b->free_list[0] = b->arena_bytes;
    const size_t entry = b->tbytes == 8 ? sizeof(BatchJob<double>) : sizeof(BatchJob<float>);
    if (hipMalloc(&b->d_jobs, entry * d->max_jobs) != hipSuccess ||
        hipMalloc((void**)&b->d_prefix, 4 * ((size_t)d->max_jobs + 1)) != hipSuccess ||
        hipMalloc((void**)&b->d_njobs, 4) != hipSuccess ||
        hipMalloc((void**)&b->d_sums, 3 * sizeof(double) * d->max_jobs) != hipSuccess)
        return bail(VLBM_ERR_ALLOC, "job table allocation failed");
    if (d->max_jobs > 65535) return bail(VLBM_ERR_INVALID, "max_jobs must be <= 65535");
    {
        // every job holds two lattices and the previous velocity, so the arena
        // bounds the resident nodes and the 128-node work units
        const size_t per_node = 2 * (size_t)d->q * d->precision + 3 * (size_t)b->tbytes;
        b->unit_cap = b->arena_bytes / per_node / 128 + (size_t)d->max_jobs + 1;
        if (hipMalloc((void**)&b->d_unit_job, 2 * b->unit_cap) != hipSuccess)
            return bail(VLBM_ERR_ALLOC, "work-unit table allocation failed");
    }
    b->K = d->steps_per_launch ? d->steps_per_launch : k_candidates(d->check_every).front();
    b->block = d->block ? d->block : 256;
    b->grid = d->grid ? d->grid : 2048;
    b->tuned = d->steps_per_launch && d->block && d->grid;
    b->fixed_shape = b->tuned;
    const int rc = upload(b);
    if (rc) { vlbm_batch_destroy(b); return rc; }
    *out = b;
    return VLBM_OK;
}

void vlbm_batch_destroy(vlbm_batch_h b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->arena) (void)hipFree(b->arena);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    if (b->d_unit_job) (void)hipFree(b->d_unit_job);
    if (b->d_njobs) (void)hipFree(b->d_njobs);
    if (b->d_sums) (void)hipFree(b->d_sums);
    if (b->d_partials) (void)hipFree(b->d_partials);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int vlbm_batch_job_bytes(vlbm_batch_h b, const vlbm_desc* d, size_t* bytes) {
    int rc = need(b);
    if (rc) return rc;
    if (!d || !bytes) return fail(VLBM_ERR_INVALID, "descriptor or output NULL");
    *bytes = layout_of(*d, b->tbytes).total;
    return VLBM_OK;
}

int vlbm_batch_submit(vlbm_batch_h b, const vlbm_desc* d, double resid_max, long long max_steps, int* job) {
    int rc = need(b);
    if (rc) return rc;
    if (!d) return fail(VLBM_ERR_INVALID, "descriptor NULL");
    if ((rc = validate_desc(d))) return rc;
    if (d->q != b->d.q || d->precision != b->d.precision)
        return fail(VLBM_ERR_INVALID, "job velocity set and precision must match the batch");
    Job j;
    j.d = *d;
    if (!box_builder_words(d->builder, j.words))
        return fail(VLBM_ERR_UNSUPPORTED, "batched jobs use the box builders (duct builders need a node mask)");
    if (max_steps < 1) return fail(VLBM_ERR_INVALID, "max_steps must be >= 1");
    j.resid_max = resid_max;
    j.max_steps = max_steps;
    j.bytes = layout_of(*d, b->tbytes).total;
    if (j.bytes > b->arena_bytes) return fail(VLBM_ERR_ALLOC, "job needs more than the whole HBM budget");
    j.t_submit = b->now();
    b->jobs.push_back(std::move(j));
    const int id = (int)b->jobs.size() - 1;
    b->pending.push_back(id);
    if (job) *job = id;
    return VLBM_OK;
}

int vlbm_batch_advance(vlbm_batch_h b, int* n_running, int* n_pending, int* n_finished) {
    int rc = need(b);
    if (rc) return rc;
    int finished_before = 0;
    for (const Job& j : b->jobs) finished_before += j.state >= VLBM_JOB_DONE;
    if ((rc = admit_pending(b))) return rc;
    if (!b->running.empty()) {
        // (re)tune when the resident nodes enter another power-of-two bucket
        if (!b->tuned || (!b->fixed_shape && retune_due(b))) {
            if ((rc = autotune(b))) return rc;
        }
        if (!b->running.empty()) {               // tuning chunks may have finished every job
            if (b->table_dirty && (rc = upload(b))) return rc;
            if ((rc = run_chunk(b))) return rc;
            if ((rc = check_jobs(b))) return rc;
        }
        if ((rc = sync(b))) return rc;
    }
    int finished = 0;
    for (const Job& j : b->jobs) finished += j.state >= VLBM_JOB_DONE;
    if (n_running) *n_running = (int)b->running.size();
    if (n_pending) *n_pending = (int)b->pending.size();
    if (n_finished) *n_finished = finished - finished_before;
    return VLBM_OK;
}

int vlbm_batch_job_info(vlbm_batch_h b, int job, vlbm_job_info* info) {
    int rc = need(b);
    if (rc) return rc;
    if (job < 0 || job >= (int)b->jobs.size() || !info) return fail(VLBM_ERR_INVALID, "unknown job or NULL output");
    const Job& j = b->jobs[job];
    info->state = j.state;
    info->steps = j.steps;
    info->residual = j.residual;
    info->mean_rho = j.mean_rho;
    info->t_submit = j.t_submit;
    info->t_start = j.t_start;
    info->t_end = j.t_end;
    info->bytes = j.bytes;
    return VLBM_OK;
}

int vlbm_batch_job_result(vlbm_batch_h b, int job, double* rho, double* ux, double* uy, double* uz) {
    int rc = need(b);
    if (rc) return rc;
    if (job < 0 || job >= (int)b->jobs.size()) return fail(VLBM_ERR_INVALID, "unknown job");
    const Job& j = b->jobs[job];
    if (!j.has_result) return fail(VLBM_ERR_INVALID, "job has no result (still running, or released)");
    double* outs[4] = {rho, ux, uy, uz};
    for (int f = 0; f < 4; f++)
        if (outs[f]) std::memcpy(outs[f], j.res[f].data(), j.res[f].size() * sizeof(double));
    return VLBM_OK;
}

int vlbm_batch_release(vlbm_batch_h b, int job) {
    int rc = need(b);
    if (rc) return rc;
    if (job < 0 || job >= (int)b->jobs.size()) return fail(VLBM_ERR_INVALID, "unknown job");
    Job& j = b->jobs[job];
    for (auto& v : j.res) std::vector<double>().swap(v);
    j.has_result = false;
    return VLBM_OK;
}

int vlbm_batch_config(vlbm_batch_h b, int* steps_per_launch, int* block, int* grid) {
    int rc = need(b);
    if (rc) return rc;
    if (steps_per_launch) *steps_per_launch = b->K;
    if (block) *block = b->block;
    if (grid) *grid = b->grid;
    return VLBM_OK;
}

double vlbm_batch_clock(vlbm_batch_h b) { return b ? b->now() : 0.0; }

}  // extern "C"
