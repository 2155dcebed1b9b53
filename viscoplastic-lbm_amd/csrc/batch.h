// Batched serving of many independent lattices on one GPU: device-side job
// table and launchers (batch_kernels.h, kernels_batch.hip; host engine in
// batch_capi.cpp).
//
// The admitted lattices form one flat work space: job j owns work units
// [prefix[j], prefix[j+1]) of BLOCK nodes each. One persistent-grid kernel
// walks the work units, finds the job of a unit by binary search in prefix,
// and runs the same per-node step as the single-lattice kernel on that job's
// parameters. Each job keeps its own two lattices; p[k] of a job reads
// lattice k and writes lattice 1-k, and step k of a chunk uses
// p[(phase + k) & 1], so a chunk of an even number of steps leaves every job
// on the lattice it started from and the job table does not change between
// chunks.
#pragma once
#include <cstdint>
#include "lattice.h"

namespace vlbm {

template <typename T>
struct BatchJob {
    Params<T> p[2];
    T* uold[3];        // velocity at the previous residual check
    int rheo;
    int phase;         // lattice holding the state at the start of a chunk
    int due;           // residual check requested for this job
    int pad;
};

// Launchers (kernels_batch.hip). prec 8 / 4 / 2; block 128 or 256. Work
// unit w (BLOCK nodes) belongs to job unit_job[w] and starts at node
// (w - prefix[job]) * BLOCK of it. rheo: the rheology every resident job
// shares, or 3 for a mix.
int launch_batch_step(int q, int prec, int block, int rheo, uint32_t grid, const void* jobs, const uint32_t* prefix,
                      const uint16_t* unit_job, const uint32_t* n_jobs, int k, void* stream);
// per-job sums {sum |u - u_old|, sum |u|, sum rho} of the jobs flagged due;
// u_old <- u for them
int launch_batch_check(int q, int prec, int block, uint32_t grid, const void* jobs, const uint32_t* prefix,
                       const uint16_t* unit_job, const uint32_t* n_jobs, uint32_t n_jobs_host, double* partials,
                       double* sums, void* stream);
// rest state g_i = w_i (rho0 - 1) in deviation storage
int launch_batch_fill_rest(int q, int prec, void* lat, int64_t stride, uint32_t n, double rho0, void* stream);

}  // namespace vlbm
