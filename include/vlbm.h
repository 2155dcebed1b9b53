/*
 * vlbm.h — C-ABI of the MI355X-native VISCOPLASTIC-LBM time-step library.
 *
 * One handle = one lattice resident in HBM. Plain C types only: pointers,
 * sizes and a POD descriptor. Every call returns VLBM_OK (0) or an error code;
 * vlbm_last_error() gives the message for the calling thread. This replaces
 * the reference's checkCudaErrors -> exit(-1) convention
 * (src/CUDA/errorDef.h:12-47).
 *
 * Each entry point names the reference symbol it replaces.
 */
#ifndef VLBM_H
#define VLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define VLBM_OK 0
#define VLBM_ERR_INVALID 1      /* bad argument or descriptor */
#define VLBM_ERR_DEVICE 2       /* HIP runtime error */
#define VLBM_ERR_ALLOC 3        /* device or host allocation failed */
#define VLBM_ERR_IO 4           /* file could not be written */
#define VLBM_ERR_UNSUPPORTED 5  /* combination the reference does not define */

/* storage precision of populations (var.h SINGLE_PRECISION / DOUBLE_PRECISION) */
#define VLBM_FP64 8
#define VLBM_FP32 4
#define VLBM_FP16 2             /* FP16 deviations f_i - w_i, FP32 arithmetic */
#define VLBM_BF16 3             /* bfloat16 deviations f_i - w_i, FP32 arithmetic (2 bytes) */

/* rounding of 16-bit storage: stochastic (unbiased: the stored value is the
 * computed one in expectation, so increments below half a spacing survive on
 * average) or to nearest */
#define VLBM_ROUND_STOCHASTIC 0
#define VLBM_ROUND_NEAREST 1   /* batched FP16 jobs (vlbm_batch_*) always round to nearest */

/* rheology (src/CUDA/NNF/nnf.h:10-30) */
#define VLBM_NEWTONIAN 0
#define VLBM_BINGHAM 1
#define VLBM_POWERLAW 2

/* Collision schemes: regularized (src/CUDA/lbm.cu:263-461, the compiled
 * default) or BGK (src/CUDA/collisionSchemes/recursiveCollisionForceNoReg.txt). */
#define VLBM_COLLISION_REGULARIZED 0
#define VLBM_COLLISION_BGK 1

/* State layout in HBM. POPULATIONS: Q planes of post-collision deviations.
 * MOMENTS: 10 planes of coefficients (A, B, G) with g*_i = w_i (A + B.c_i +
 * c_i^T G c_i), which hold the regularized post-collision state exactly (float
 * and double storage, regularized collision only): 10 instead of Q values
 * read and written per node. AUTO: MOMENTS in float storage with the
 * regularized collision, for every builder but the rheometer, when NX >= 64
 * and the lattice holds at least 2,048 tiles of 64 x 8 nodes (64 x 4 for
 * D3Q27 and for the duct builders' node maps) by 4 planes; else POPULATIONS.
 * vlbm_layout reports the choice. */
#define VLBM_LAYOUT_AUTO 0
#define VLBM_LAYOUT_POPULATIONS 1
#define VLBM_LAYOUT_MOMENTS 2

/* boundary-condition builders (src/CUDA/boundaryConditionsBuilders/) */
#define VLBM_BUILDER_TAYLOR_GREEN 0          /* taylorGreenVortex.cu: fully periodic */
#define VLBM_BUILDER_PLATES_BB_Y 1           /* parallelPlatesBounceBack_Y.cu */
#define VLBM_BUILDER_PLATES_BB_X 2           /* parallelPlatesBounceBack_X.cu */
#define VLBM_BUILDER_PLATES_BB_Z 3           /* parallelPlatesBounceBack_Z.cu */
#define VLBM_BUILDER_SQUARE_DUCT_BB 4        /* squareDuctBounceBack.cu */
#define VLBM_BUILDER_CLOSED_BOX_BB 5         /* closedBoxBounceBack.cu */
#define VLBM_BUILDER_PLATES_ZOUHE 6          /* parallelPlatesZouHe.cu */
#define VLBM_BUILDER_LID_DRIVEN_BB 7         /* lidDrivenBounceBack.cu */
#define VLBM_BUILDER_COUETTE_ZOUHE 8         /* couetteZouHe.cu */
#define VLBM_BUILDER_INFLOW_FREE_SLIP 9      /* inflowFreeSlip.cu */
#define VLBM_BUILDER_PLATES_BB_FREE_SLIP 10  /* parallelPlatesBounceBackFreeSlip.cu */
#define VLBM_BUILDER_SQUARE_DUCT_ZOUHE 11    /* no reference file: squareDuctBounceBack walls,
                                                velocity Zou-He inlet at z=0 (uz index 1),
                                                pressure Zou-He outlet at z=NZ-1 (rho index 0) */
#define VLBM_BUILDER_CIRCULAR_DUCT_IBB 12    /* circularDuctInterpBounceBack.cu */
#define VLBM_BUILDER_ANNULAR_DUCT_IBB 13     /* annularDuctInterpBounceBack.cu */
#define VLBM_BUILDER_RHEOMETER 14            /* parallelPlatesRheometer.cu: the circular duct, planes
                                                z = 0 and NZ-1 unused, gpuSchSpecial on z = 1 (back,
                                                at rest) and NZ-2 (front, outer_rotation); D3Q19,
                                                population layout */

/*
 * Simulation descriptor: the runtime form of the compile-time constants of
 * src/CUDA/var.h:18-160, src/CUDA/NNF/nnf.h:10-24 and the builder choice.
 */
typedef struct vlbm_desc {
    int nx, ny, nz;          /* NX, NY, NZ (var.h:97-101); one GPU holds the whole lattice */
    int q;                   /* 19 or 27 (var.h D3Q19 / D3Q27) */
    int precision;           /* VLBM_FP64, VLBM_FP32, VLBM_FP16 or VLBM_BF16 */
    int rheology;            /* VLBM_NEWTONIAN / VLBM_BINGHAM / VLBM_POWERLAW */
    int builder;             /* VLBM_BUILDER_* */
    double tau;              /* TAU: Newtonian relaxation time (OMEGA = 1/TAU) */
    double rho0;             /* RHO_0: initial density */
    double fx, fy, fz;       /* FX, FY, FZ: constant body force */
    double eta_p, s_y;       /* Bingham: plastic viscosity ETA_P, yield stress S_Y */
    double n_index, k_consistency, gamma_0;   /* power law: N_INDEX, K_CONSISTENCY, GAMMA_0 */
    int omega_clamp;         /* 1: clamp the apparent tau into [tau_min, tau_max] (not in the reference) */
    double tau_min, tau_max;
    double ux_bc[8], uy_bc[8], uz_bc[8], rho_bc[8];   /* UX_BC, UY_BC, UZ_BC, RHO_BC (var.h:114-117) */
    double duct_outer_radius, duct_inner_radius;      /* OUTER_RADIUS, INNER_RADIUS (var.h:141-160) */
    double outer_rotation, outer_velocity, inner_rotation, inner_velocity;
    int zouhe_textbook;      /* 0: reproduce the reference's Zou-He lines; 1: textbook Zou-He */
    int collision;           /* VLBM_COLLISION_REGULARIZED or VLBM_COLLISION_BGK */
    double les_const;        /* LES_MODEL's Implicit_const (lbm.cu:226-242); 0 = no LES, Newtonian only */
    int layout;              /* VLBM_LAYOUT_* (not in the reference: its state is always populations) */
    int rounding;            /* VLBM_ROUND_*: rounding of 16-bit storage (not in the reference) */
    int les_visc_turb;       /* LES_EXPORT_VISC_TURBULENT (LES/lesVar.h:16): keep the eddy viscosity
                                tau_t/3 of every node (lbm.cu:229-239), read with vlbm_visc_turb and
                                saved as ID_visc_turb<step>.bin; needs les_const > 0 */
} vlbm_desc;

typedef struct vlbm_ctx* vlbm_h;

/* Fills a descriptor with the reference defaults of var.h / nnf.h. */
void vlbm_desc_defaults(vlbm_desc* d);

/* Allocates the lattice on the current device and builds its boundary map.
 * Replaces Populations::popAllocation + Macroscopics::macrAllocation
 * (src/CUDA/structs/populations.h:47-52, src/CUDA/structs/macroscopics.h:83-124)
 * and the gpuBuildBoundaryConditions launch (src/CUDA/main.cu:177-185). */
int vlbm_create(const vlbm_desc* d, vlbm_h* out);
void vlbm_destroy(vlbm_h h);
const char* vlbm_last_error(void);

/* Node-type map words, x fastest ([NZ][NY][NX]), bit-identical to the
 * reference's NodeTypeMap (src/CUDA/structs/nodeTypeMap.h:15-322) as written
 * by gpuBuildBoundaryConditions (boundaryConditionsBuilder.h:18-19). Host only. */
int vlbm_build_map(const vlbm_desc* d, uint32_t* out);

/* Node map written by a builder outside the library: NX*NY*NZ words in the
 * layout of vlbm_build_map, replacing the handle's map (the drop-in for a
 * user-written gpuBuildBoundaryConditions, boundaryConditionsBuilder.h:18-19,
 * whose words the step's dispatch reads, boundaryConditionsHandler.cu:25-79).
 * Refused: the rheometer, a handle with the immersed boundary stage, the
 * special scheme, Zou-He off the six faces, D3Q27 Zou-He / velocity
 * bounce-back, interpolated bounce-back without a duct radius. */
int vlbm_set_map(vlbm_h h, const uint32_t* words);

/* Linear indices of the used BC nodes in z,y,x order
 * (BoundaryConditionsInfo::setupBoundaryConditionsInfo,
 * src/CUDA/structs/boundaryConditionsInfo.h:76-116). out may be NULL to query the count. */
int vlbm_bc_nodes(vlbm_h h, uint64_t* out, size_t cap, size_t* count);

/* Equilibrium initialisation from (rho, u) fields of NX*NY*NZ doubles; NULL
 * pointers mean rho0 / zero velocity (gpuInitialization + gpuMacrInitValue,
 * src/CUDA/lbmInitialization.cu:48-150). Resets the step counter. */
int vlbm_init(vlbm_h h, const double* rho, const double* ux, const double* uy, const double* uz);

/* Loads / reads populations in the reference layout: Q planes of (NZ+1)*NY*NX
 * doubles, x fastest (idxPop, src/CUDA/globalFunctions.h:143-147), holding the
 * post-streaming, post-BC state that the next step collides
 * (the pop array of src/CUDA/simCheckpoint.cu:216-318). vlbm_get_pops returns
 * the reference's whole buffer: nodes that never collide hold what its push
 * streaming leaves there, and the extra plane z=NZ what the last step pushed
 * across the z ends (the rest state w_i before any step). The extra plane is
 * ignored on load. */
int vlbm_load_pops(vlbm_h h, const double* pop);
int vlbm_get_pops(vlbm_h h, double* pop);

/* vlbm_get_pops without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops snapshotting states between steps, harnesses that forbid hipMalloc).
 * Same output as vlbm_get_pops. The values pass, chunk by chunk, through one
 * 8 MiB device array per GPU, held from the first call on that GPU until the
 * process ends; lattices of any size are served in chunks of 55,040 (D3Q19)
 * or 38,656 (D3Q27) nodes. Calls that overlap on one GPU share that array and
 * race; the calls of one thread are ordered. */
int vlbm_get_pops_noalloc(vlbm_h h, double* pop);

/* Power-law apparent omega state, NX*NY*NZ doubles (macr.omega). */
int vlbm_set_omega(vlbm_h h, const double* omega);

/* Runs n_steps time steps: fused gather + BC + moments + collision.
 * Replaces gpuMacrCollisionStream + gpuPopulationsTransfer + gpuApplyBC +
 * swapPop (src/CUDA/main.cu:300-367; src/CUDA/lbm.cu:3-504,579-657).
 * save_last != 0 stores the macroscopics computed during the last step
 * (the state entering it), as save_macr_to_array does (src/CUDA/main.cu:296). */
int vlbm_step(vlbm_h h, int n_steps, int save_last);

/* Computes the macroscopics of the current state into the saved arrays
 * without advancing (the moments the next step would compute). */
int vlbm_compute_macroscopics(vlbm_h h);

/* Copies the saved macroscopics to host doubles, [NZ][NY][NX]; any pointer
 * may be NULL. omega is the apparent omega (non-Newtonian only). */
int vlbm_macroscopics(vlbm_h h, double* rho, double* ux, double* uy, double* uz, double* omega);

/* LES eddy viscosity tau_t/3 of the last step, NX*NY*NZ doubles (macr.visc_turb,
 * written every step when desc.les_visc_turb is set; 0 before the first). */
int vlbm_visc_turb(vlbm_h h, double* out);

/* Residual and mean density of the saved macroscopics with FP64
 * accumulation: residual = sum|u - u_old| / sum|u| (1 when the denominator is
 * 0), mean_rho = sum(rho)/N; u_old becomes u (treatData,
 * src/CUDA/treatData.cpp:25-78). Either output may be NULL. */
int vlbm_reduce(vlbm_h h, double* residual, double* mean_rho);

/* Writes PATH/ID/ID_{rho,ux,uy,uz[,omega]}<step>.bin in the storage
 * precision, [NZ][NY][NX], no header (saveAllMacrBin / getVarFilename,
 * src/CUDA/lbmReport.cpp:55-172). The folder must exist. */
int vlbm_save_macr_bin(vlbm_h h, const char* path, const char* id, unsigned step);

/* Writes PATH/ID/ID_info.txt with the keys read by
 * src/Post Processing/fileTreat.py:56-143 (getSimInfoString,
 * src/CUDA/lbmReport.cpp:174-386). */
int vlbm_save_info(vlbm_h h, const char* path, const char* id, unsigned total_steps,
                   double mlups, double seconds);

/* Writes the checkpoint folder PATH/ID/checkpoint/ in the layout of
 * saveSimCheckpoint / operateSimCheckpoint (src/CUDA/simCheckpoint.cu:216-318):
 * 0_curr_step.bin (int32 step), 0_pop.bin and 0_popAux.bin (Q planes of
 * (NZ+1)*NY*NX, the post-stream state the next step collides), 0_rho.bin and
 * 0_u/{x,y,z}.bin (NX*NY*(NZ+2B), B = 3 border planes with IBM, else 0),
 * 0_omega.bin (non-Newtonian), and with IBM 0_f/{x,y,z}.bin, the
 * ParticleCenter records 0_IBM_particles_centers.bin and the
 * 0_IBM_nodes_* arrays. Values are double for FP64 lattices, float otherwise.
 * PATH must exist; ID and checkpoint are created. Passing vlbm_steps_done
 * makes a loaded run continue exactly; the reference's in-loop checkpoints
 * (src/CUDA/main.cu:452-455) store the label of the step just run, one less. */
int vlbm_save_checkpoint(vlbm_h h, const char* path, const char* id, int step);

/* Loads a checkpoint folder written by vlbm_save_checkpoint or by the
 * reference (loadSimCheckpoint, src/CUDA/simCheckpoint.cu:297-305) into a
 * handle of the same descriptor (and, with IBM, the same spheres and node
 * counts); file sizes are checked against the handle. The step counter
 * becomes the stored step, which also labels the next IBM stage. */
int vlbm_load_checkpoint(vlbm_h h, const char* path, const char* id, int* step);

/* Steps executed since vlbm_init / vlbm_load_pops, or the step stored by the
 * checkpoint that vlbm_load_checkpoint read. */
int vlbm_steps_done(vlbm_h h, long long* steps);

/* Runs n_steps and reports the device time between two HIP events recorded
 * on the library's stream around them (cudaEventRecord timing of main.cu:259-267). */
int vlbm_time_steps(vlbm_h h, int n_steps, double* ms);

/* Waits for all work queued on the handle's stream. */
int vlbm_sync(vlbm_h h);

/* Device bytes held by the handle. */
int vlbm_device_bytes(vlbm_h h, size_t* bytes);

/* The state representation in use, VLBM_LAYOUT_POPULATIONS or
 * VLBM_LAYOUT_MOMENTS (what VLBM_LAYOUT_AUTO chose). No counterpart in the
 * reference, whose state is always populations. */
int vlbm_layout(vlbm_h h, int* layout);

/* Average time of a plain device-to-device copy of `bytes` (16-byte loads and
 * stores), the calibration of the HBM traffic counters. No reference
 * counterpart (measurement helper). */
int vlbm_copy_probe(size_t bytes, int iters, double* ms);

/* ------------------------------------------------------------------------
 * Immersed boundary stage for rigid spheres (src/CUDA/IBM/). When enabled,
 * every vlbm_step runs the LBM step and then the stage of
 * immersedBoundaryMethod (src/CUDA/IBM/ibm.cu:6-192), labelled with the
 * number of steps done before it, as main.cu:369-389 does with its loop
 * counter. Particle arithmetic is FP64 whatever the lattice precision.
 * ------------------------------------------------------------------------ */

/* sphere meshes for vlbm_ibm_add_sphere */
#define VLBM_MESH_POLAR 0        /* Particle::makeSpherePolar (src/CUDA/IBM/structs/particle.cpp:698-995) */
#define VLBM_MESH_FIBONACCI 1    /* Fibonacci sphere lattice of n equal-area nodes (no reference counterpart) */

/* Particle record of vlbm_ibm_get_particles / vlbm_ibm_set_particles: 48
 * doubles per sphere, in this order (ParticleCenter,
 * src/CUDA/IBM/structs/particleCenter.h:65-118): pos[3] vel[3] w[3] f[3]
 * M[3] I[6] (xx yy zz xy xz yz) pos_old[3] vel_old[3] w_old[3] w_avg[3]
 * w_pos[3] f_old[3] M_old[3] dP_internal[3] dL_internal[3]. */
#define VLBM_IBM_RECORD 48
#define VLBM_IBM_HISTORY 13      /* MAX_ACTIVE_COLLISIONS: 7 wall slots + 6 partner slots */

/* Runtime form of src/CUDA/IBM/ibmVar.h and src/CUDA/IBM/ibmBoundaryCondition.h. */
typedef struct vlbm_ibm_desc {
    int max_iterations;                       /* IBM_MAX_ITERATION */
    int periodic_x, periodic_y, periodic_z;   /* IBM_BC_{X,Y,Z}_PERIODIC; 0 = IBM_BC_*_WALL (walls at 0 and N-1) */
    double thickness;                         /* IBM_THICKNESS */
    double alpha;                             /* IBM_MOVEMENT_DISCRETIZATION */
    double fluid_density;                     /* FLUID_DENSITY */
    double gx, gy, gz;                        /* GX, GY, GZ */
    double pp_friction, pw_friction;          /* PP_FRICTION_COEF, PW_FRICTION_COEF */
    double pp_rest, pw_rest;                  /* PP_REST_COEF, PW_REST_COEF */
    double young, poisson;                    /* PARTICLE_YOUNG_MODULUS, PARTICLE_POISSON_RATIO */
    double wall_young, wall_poisson;          /* WALL_YOUNG_MODULUS, WALL_POISSON_RATIO */
    int ref_float_helpers;                    /* 1: sqrtf / rsqrtf / fabsf where the FP64 reference calls
                                                 them (vector_length, vector_normalize, sphereSphereGap,
                                                 sphereWallCollision); 0: FP64 throughout */
    double mesh_scale;                        /* MESH_SCALE (polar mesh) */
    double breugem;                           /* BREUGEM_PARAMETER (polar mesh) */
    int coulomb;                              /* MESH_COULOMB (polar mesh) */
    int max_spheres, max_nodes;               /* capacities allocated by vlbm_ibm_enable */
} vlbm_ibm_desc;

/* Reference defaults (NUM_PARTICLES-independent constants of ibmVar.h /
 * ibmBoundaryCondition.h); capacities 256 spheres, 262144 nodes. */
void vlbm_ibm_desc_defaults(vlbm_ibm_desc* d);

/* Allocates the particle arrays (ParticlesSoA, src/CUDA/IBM/structs/particle.cpp:5-63),
 * the per-node force field (macr.f) and the spread-force buffer (IBMMacrsAux,
 * src/CUDA/IBM/structs/ibmMacrsAux.h) and switches the handle to IBM steps. */
int vlbm_ibm_enable(vlbm_h h, const vlbm_ibm_desc* d);

/* Adds a sphere (makeSpherePolar / a Fibonacci lattice) with state
 * {pos[3], vel[3], w[3]}; n_fibonacci is the node count of VLBM_MESH_FIBONACCI.
 * I = 2/5 m r^2, history slots cleared. *id receives the sphere index. */
int vlbm_ibm_add_sphere(vlbm_h h, const double state[9], double radius, double density, int movable,
                        int mesh, int n_fibonacci, int* id);

/* Adds a sphere whose Lagrangian nodes are given: positions xyz[3*n]
 * (x0 y0 z0 x1 ...), areas area[n] (ParticleNode pos / S). */
int vlbm_ibm_add_particle(vlbm_h h, const double state[9], double radius, double density, int movable,
                          int n_nodes, const double* xyz, const double* area, int* id);

int vlbm_ibm_counts(vlbm_h h, int* n_spheres, int* n_nodes);

/* Particle records (VLBM_IBM_RECORD doubles each), host <-> device. */
int vlbm_ibm_get_particles(vlbm_h h, double* out);
int vlbm_ibm_set_particles(vlbm_h h, const double* in);

/* Contact history (CollisionData, src/CUDA/IBM/structs/particleCenter.h:36-63):
 * partner[n][13], tang[n][13][3], last[n][13]; any pointer may be NULL. */
int vlbm_ibm_get_collisions(vlbm_h h, int* partner, double* tang, int* last);

/* Lagrangian nodes: positions, accumulated forces of the last stage and the
 * force of its last iteration (ParticleNodeSoA pos / f / deltaF), [3][n] each. */
int vlbm_ibm_get_nodes(vlbm_h h, double* pos, double* force, double* dforce);

/* Force field the next collision uses, [3][NZ][NY][NX] doubles (macr.f). */
int vlbm_ibm_get_force(vlbm_h h, double* f);

/* Uniform cell list of the last stage: dims[3] cells per axis and the cell
 * index of each sphere (x fastest). No reference counterpart: the reference
 * tests all pairs (gpuParticlesCollisionHandler, ibmCollision.cu:2381-2448). */
int vlbm_ibm_cells(vlbm_h h, int* dims, int* cell);

/* Runs one IBM stage on the current state without an LBM step, labelled like
 * the next step would be; advances the step counter (for tests and for
 * particles without fluid coupling). */
int vlbm_ibm_stage(vlbm_h h);

/* VLBM_OK, or VLBM_ERR_UNSUPPORTED when a sphere had more contacts than the
 * contact buffer holds (16 per sphere and step). */
int vlbm_ibm_status(vlbm_h h);

/* Runs n_steps and reports the device time of the LBM step kernels and of the
 * IBM stages separately (HIP events around each part). */
int vlbm_ibm_time_steps(vlbm_h h, int n_steps, double* ms_lbm, double* ms_ibm);

/* ------------------------------------------------------------------------
 * Serving: many independent lattices resident on one GPU, admitted by an
 * HBM budget and stepped together by one persistent-grid kernel over all of
 * them, with captured step sequences (hipGraph) and an autotuned launch
 * shape. No reference counterpart: the reference runs one compile-time-sized
 * simulation per process (src/CUDA/main.cu). Jobs use the box builders and
 * start from rest (rho0, u = 0); the stop test is the reference's
 * (treatData / stopSim, src/CUDA/treatData.cpp:25-78,451-459).
 * ------------------------------------------------------------------------ */
typedef struct vlbm_batch_desc {
    size_t hbm_budget;       /* bytes of the arena that holds the jobs' lattices */
    int max_jobs;            /* resident jobs at most */
    int q, precision;        /* shared by every job of the batch */
    int check_every;         /* steps between residual checks of a job (even) */
    int steps_per_launch;    /* steps per captured launch sequence; 0 = autotune (even, divides check_every) */
    int block;               /* threads per block, 128 or 256; 0 = autotune */
    int grid;                /* blocks of the persistent grid; 0 = autotune */
    int use_graph;           /* 1: replay captured sequences (hipGraph); 0: plain launches */
    const char* tune_cache;  /* file the autotuner reads and appends (NULL: none) */
    size_t max_resident_nodes;   /* admit a pending job only while the resident lattices hold at most
                                    this many nodes (a lone job is always admitted); 0 = no limit */
} vlbm_batch_desc;

#define VLBM_JOB_PENDING 0
#define VLBM_JOB_RUNNING 1
#define VLBM_JOB_DONE 2          /* residual < resid_max */
#define VLBM_JOB_MAX_STEPS 3     /* reached max_steps (checked at its check points) */
#define VLBM_JOB_FAILED 4        /* non-finite or negative mean density */

typedef struct vlbm_job_info {
    int state;               /* VLBM_JOB_* */
    long long steps;
    double residual, mean_rho;
    double t_submit, t_start, t_end;   /* seconds on the batch clock; -1 = not yet */
    size_t bytes;            /* arena bytes while resident */
} vlbm_job_info;

typedef struct vlbm_batch* vlbm_batch_h;

void vlbm_batch_desc_defaults(vlbm_batch_desc* d);
int vlbm_batch_create(const vlbm_batch_desc* d, vlbm_batch_h* out);
void vlbm_batch_destroy(vlbm_batch_h b);
/* Arena bytes a job would hold (two lattices, previous velocity, power-law omega). */
int vlbm_batch_job_bytes(vlbm_batch_h b, const vlbm_desc* d, size_t* bytes);
/* Queues a job; it is admitted by vlbm_batch_advance when a slot and memory are free. */
int vlbm_batch_submit(vlbm_batch_h b, const vlbm_desc* d, double resid_max, long long max_steps, int* job);
/* Admits what fits, runs one chunk of steps of every resident job, checks the
 * jobs at their check points and retires the finished ones. */
int vlbm_batch_advance(vlbm_batch_h b, int* n_running, int* n_pending, int* n_finished);
int vlbm_batch_job_info(vlbm_batch_h b, int job, vlbm_job_info* info);
/* Final macroscopics of a finished job, [NZ][NY][NX] doubles each (any may be NULL). */
int vlbm_batch_job_result(vlbm_batch_h b, int job, double* rho, double* ux, double* uy, double* uz);
/* Drops the host copy of a finished job's result. */
int vlbm_batch_release(vlbm_batch_h b, int job);
/* Launch shape in use (after autotuning). */
int vlbm_batch_config(vlbm_batch_h b, int* steps_per_launch, int* block, int* grid);
/* Seconds since vlbm_batch_create (the clock of vlbm_job_info). */
double vlbm_batch_clock(vlbm_batch_h b);

#ifdef __cplusplus
}
#endif
#endif /* VLBM_H */
