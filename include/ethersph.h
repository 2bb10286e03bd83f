/*
 * ethersph.h — C ABI of the MI355X (gfx950) per-step SPH particle hot path.
 *
 * The library replaces the per-step work of EtherSPHCPU.jl's `solve!` /
 * `eachStep!` (neighbour search, pair interactions, self updates, Shepard
 * re-initialisation, particle deletion) for the four fixed-count solvers.
 * Host code (the reference's Julia types, or the Python mirror in this
 * repository) packs its particle vectors into contiguous arrays, hands them
 * over once, steps on the device and reads fields back.
 *
 * Conventions
 *  - Every entry point returns an int status (ESPH_OK == 0).  No C++
 *    exception crosses this boundary.  esph_last_error() explains the last
 *    failure of a context (thread-local message per context).
 *  - Host buffers passed in are borrowed for the duration of the call and
 *    copied; the library owns all device memory.
 *  - Vector fields are particle-major: x[i*dim + k], which is the memory
 *    layout of a Julia Matrix{Float64}(dim, n) (column-major) and of a C
 *    row-major (n, dim) array.
 *  - Particle indices are 0-based.  Particle order of downloads is the
 *    reference's order: upload order, with deleted particles removed
 *    (stable, as Base.deleteat!).
 *  - One host thread per context at a time; contexts are independent.
 *
 * Neighbour rule (replaces CellListMap.neighborlist; version unpinned, see
 * Solve/FluidWallSolve.jl:22,26-31): particles i and j are neighbours iff
 *     d2 <= R*R,  d2 = ((dx0*dx0) + dx1*dx1) [+ dx2*dx2],  dxk = xi_k - xj_k
 * evaluated in IEEE binary64 with every operation rounded separately (no
 * fused multiply-add), and i != j for pairs within one set.  r = sqrt(d2).
 * The rule holds in every precision mode.
 */
#ifndef ETHERSPH_H
#define ETHERSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESPH_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
enum {
    ESPH_OK = 0,
    ESPH_ERR_ARG = 1,      /* invalid argument                           */
    ESPH_ERR_HIP = 2,      /* HIP runtime error                          */
    ESPH_ERR_BUDGET = 3,   /* HBM budget of the context exceeded         */
    ESPH_ERR_NAN = 4,      /* non-finite particle position met in a step */
    ESPH_ERR_STATE = 5,    /* call out of order (e.g. step before upload)*/
    ESPH_ERR_CAPACITY = 6, /* output buffer too small                    */
    ESPH_ERR_NO_DEVICE = 7 /* no usable GPU                              */
};

/* ---- smoothing kernels: SPHKernel/SmoothKernel/SmoothKernel.jl:15-20 ---- */
enum {
    ESPH_CUBIC_SPLINE = 0, /* CubicSpline.jl:39-73  */
    ESPH_GAUSSIAN = 1,     /* Gaussian.jl:31-56     */
    ESPH_WENDLAND_C2 = 2,  /* WendlandC2.jl:38-63   */
    ESPH_WENDLAND_C4 = 3   /* WendlandC4.jl:38-54   */
};

/* ---- solvers: one per fixed-count eachStep! method under src/Solve/ ------ */
enum {
    ESPH_SOLVER_FLUID_WALL = 0,                     /* FluidWallSolve.jl:8-83 */
    ESPH_SOLVER_FLUID_VELOCITY = 1,                 /* FluidVelocitySolve.jl:8-102 */
    ESPH_SOLVER_FLUID_THERMOSTATIC_WALL = 2,        /* FluidThermostaticWallSolve.jl:8-125 */
    ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL = 3, /* FluidCompulsiveThermostaticWallSolve.jl:8-152 */
    ESPH_SOLVER_FLUID_INLET_OUTLET_WALL = 4             /* FluidInletOutletWallSolve.jl:8-264 */
};

/* ---- buffer velocity profiles of the inlet/outlet solver (the reference's
   user_defined_function, e.g. PoiseuilleFlow2D.jl:28-35,152-168) ------- */
enum {
    ESPH_PROFILE_NONE = 0,       /* buffer velocities are left as they are */
    ESPH_PROFILE_CHANNEL_2D = 1, /* transient plane Poiseuille: y' = y - c[1];
                                    u = f/(2 nu) y'(l - y') - sum_{n=0..terms} 4 f l^2 / (nu pi^3 (2n+1)^3)
                                        sin(pi y' (2n+1) / l) exp(-(2n+1)^2 pi^2 nu t / l^2) */
    ESPH_PROFILE_PIPE_3D = 2     /* steady pipe Poiseuille: u = f/(4 nu) (l^2 - (y-c[1])^2 - (z-c[2])^2) */
};

/* ---- particle groups (downloads) ------------------------------------------ */
enum {
    ESPH_GROUP_FLUID = 0,
    ESPH_GROUP_INLET = 1, /* inlet buffer particles (inlet/outlet solver) */
    ESPH_GROUP_OUTLET = 2 /* outlet buffer particles (inlet/outlet solver) */
};

/* ---- arithmetic forms -------------------------------------------------- */
enum {
    ESPH_FP64 = 0,     /* everything in binary64 (the reference's form)                 */
    ESPH_FP32_BULK = 1 /* pair arithmetic and pair sums in binary32; position, velocity,
                          density and temperature integration and the Shepard sums in
                          binary64                                                       */
};

/* ---- wall particle kinds: Particle/FixedParticle/WallParticle.jl ------- */
enum {
    ESPH_WALL_COMPULSIVE = 0,      /* CompulsiveWallParticle      :8-19  */
    ESPH_WALL_STATIC_VELOCITY = 1, /* StaticVelocityWallParticle  :21-34 */
    ESPH_WALL_THERMOSTATIC = 2     /* ThermostaticWallParticle    :36-51 */
};

/* ---- downloadable fluid fields (LiquidParticle.jl:8-101) --------------- */
enum {
    ESPH_FIELD_X = 0,     /* x_vec_    (dim per particle) */
    ESPH_FIELD_V = 1,     /* v_vec_    (dim per particle) */
    ESPH_FIELD_RHO = 2,   /* rho_   */
    ESPH_FIELD_P = 3,     /* p_     */
    ESPH_FIELD_C = 4,     /* c_     */
    ESPH_FIELD_MASS = 5,  /* mass_  */
    ESPH_FIELD_GAP = 6,   /* gap_   */
    ESPH_FIELD_T = 7,     /* t_     (thermal) */
    ESPH_FIELD_KAPPA = 8, /* kappa_ (thermal) */
    ESPH_FIELD_CP = 9,    /* cp_    (thermal) */
    ESPH_FIELD_ID = 10    /* upload index of each live particle, as double */
};

/* ---- neighbour pair sets ---------------------------------------------- */
enum {
    ESPH_PAIRS_FLUID_FLUID = 0, /* unordered (min, max) fluid index pairs         */
    ESPH_PAIRS_FLUID_WALL = 1   /* (fluid index, wall index), walls concatenated
                                   in the order of their groups                    */
};

/*
 * Scene descriptor.  Named fields only: the reference's
 * DensityReinitializedFixedTimeIntervalForwardEuler positional order
 * (dt, total_step, output_interval, check_interval,
 *  density_reinitialized_interval), FixedTimeIntervalDiscretization.jl:10-16,
 * is easy to swap; here each interval has its own name.
 */
typedef struct esph_scene_desc {
    int32_t dim;              /* 2 or 3 (SmoothKernel.jl:42 also allows 1; not supported) */
    int32_t kernel;           /* ESPH_CUBIC_SPLINE ...                                  */
    double influence_radius;  /* R; h = R / radius_ratio (SmoothKernel.jl:43-44)          */
    int32_t solver;           /* ESPH_SOLVER_*                                          */
    int32_t precision;        /* ESPH_FP64 / ESPH_FP32_BULK                              */
    /* CommonWeaklyCompressibleLiquidModel / ThermalWeaklyCompressibleLiquidModel
       (EquationModel/LiquidModel.jl:10-66); b and nu_0 are derived as there. */
    double rho_0;
    double c_0;
    double p_0;
    int32_t gamma;
    double mu_0;
    double body_force[3];
    double t_0;               /* thermal only */
    double kappa_0;           /* thermal only (kept for completeness; unused by the step) */
    double cp;                /* thermal only */
    double beta;              /* thermal only */
    /* time discretisation */
    double dt;
    int64_t density_reinitialized_interval; /* <= 0: never */
    int64_t check_interval;                 /* <= 0: never */
    /* `check` predicate as data: a fluid particle is deleted at a check step
       iff for some axis k: x_k > check_max[k] or x_k < check_min[k]
       (comparisons false for NaN, as the reference's `if x > bound` tests) */
    int32_t check_enabled;
    double check_min[3];
    double check_max[3];
    /* sizing */
    int64_t fluid_capacity;   /* 0: number of fluid particles at upload */
    int32_t domain_set;       /* 1: use domain_min/max for the cell grid, else the
                                 bounding box of the uploaded particles and the finite
                                 check bounds, padded by one cell                       */
    double domain_min[3];
    double domain_max[3];
    /* ESPH_SOLVER_FLUID_INLET_OUTLET_WALL: the reference's check / modify
       closures as planes along io_axis (PoiseuilleFlow2D.jl:146-150):
       after each step an inlet particle with x[axis] > inlet_enter is copied
       into the fluid and moved back by inlet_shift; a fluid particle with
       x[axis] > outlet_enter moves into the outlet; an outlet particle with
       x[axis] > outlet_exit is removed. */
    int32_t io_axis;
    double inlet_enter;
    double inlet_shift;
    double outlet_enter;
    double outlet_exit;
    /* buffer velocities set after each step at t = step * dt: inlet
       v[axis] = profile; outlet v[axis] = profile when outlet_profile is 1;
       outlet cross components are zeroed (PoiseuilleFlow2D.jl:152-168,
       PoiseuilleFlow3D.jl:143-158) */
    int32_t profile;          /* ESPH_PROFILE_* */
    int32_t profile_terms;
    int32_t outlet_profile;
    double profile_f;
    double profile_nu;
    double profile_l;         /* channel width (2D) or pipe radius (3D) */
    double profile_center[3];
    /* candidate-list skin as a fraction of R: the lists hold every particle
       within R (1 + skin) and are reused, filtered by the exact rule, until
       some particle has moved skin R / 2 (or a check or migration step
       needs a sort).  0: the library default (0.06 in 3D, 0.045 in 2D);
       negative: the lists are
       rebuilt every step.  The inlet/outlet solver rebuilds every step.
       Results follow the same neighbour rule either way. */
    double verlet_skin;
    /* further `check` predicates (used with check_enabled): a fluid particle
       is also deleted at a check step iff for some q < check_nplanes (at
       most 8): n_q . x > d_q, with check_planes[q] = (n_x, n_y, n_z, d_q) and
       n_q . x evaluated as (n_x x + n_y y) + n_z z.  Together with the box
       this expresses any convex keep-region bounded by planes. */
    int32_t check_nplanes;
    double check_planes[8][4];
} esph_scene_desc;

typedef struct esph_fluid_arrays {
    int64_t n;
    const double* x;      /* n*dim */
    const double* v;      /* n*dim */
    const double* rho;    /* n */
    const double* p;      /* n */
    const double* c;      /* n */
    const double* mass;   /* n */
    const double* gap;    /* n */
    const double* t;      /* n, thermal solvers (NULL otherwise) */
    const double* kappa;  /* n, thermal solvers */
    const double* cp;     /* n, thermal solvers */
} esph_fluid_arrays;

typedef struct esph_wall_arrays {
    int32_t kind;         /* ESPH_WALL_* */
    int64_t n;
    const double* x;      /* n*dim */
    const double* normal; /* n*dim */
    const double* gap;    /* n */
    const double* v;      /* n*dim, ESPH_WALL_STATIC_VELOCITY (NULL otherwise) */
    const double* t;      /* n, ESPH_WALL_THERMOSTATIC */
    const double* kappa;  /* n, ESPH_WALL_THERMOSTATIC */
} esph_wall_arrays;

typedef struct esph_ctx esph_ctx;
typedef struct esph_scene esph_scene;

/* ---- library / context -------------------------------------------------- */
int esph_abi_version(void);
/* Creates a context on HIP device `device`; hbm_budget_bytes == 0 means no
   limit.  The context owns one HIP stream (default priority). */
int esph_ctx_create(int32_t device, uint64_t hbm_budget_bytes, esph_ctx** out);
int esph_ctx_destroy(esph_ctx* ctx);
const char* esph_last_error(const esph_ctx* ctx);
/* Pair-pass tuner (no counterpart in the reference: tuning knobs of this
   library).  The pair passes keep 1, 2 or 4 neighbour gathers in flight per
   thread and run in workgroups of 64, 128 or 256 threads; results are
   identical for every choice.  mode 1 (default): the first step of each class
   of scenes (dim, kernel, precision, thermal mix) times the force pass at the
   three depths (256 threads), then at the fastest depth with 64 and 128
   threads, on that step's neighbour lists, and keeps the fastest; then the
   continuity and Shepard passes at all nine pairs, each keeping its own
   (esph_ctx_tuned_pass); mode 0: the given depth (and the workgroup size of
   esph_ctx_set_workgroup, 256 by default) everywhere.  Clears earlier
   choices; applies to steps launched afterwards. */
int esph_ctx_set_tuning(esph_ctx* ctx, int32_t mode, int32_t depth);
/* The workgroup size (64, 128 or 256 threads) used in tuning mode 0. */
int esph_ctx_set_workgroup(esph_ctx* ctx, int32_t threads);
/* Brick walks (no counterpart in the reference: a choice between two
   pair-walk paths of this library).  enable 1: the continuity and force
   passes of 3D FP32_BULK scenes without heat, buffer particles or varying
   gaps take the particles brick by brick (blocks of cells), with the
   quantized positions and velocity records of each brick's neighbour region
   staged in LDS; 0 (default): those scenes gather every candidate from
   device memory like the others (faster on the 1M-particle 3D dam break:
   0.27 against 0.38 ms per force pass).  Results are identical either way.
   Applies to scenes whose fluid is first uploaded afterwards. */
int esph_ctx_set_brick_walk(esph_ctx* ctx, int32_t enable);
/* Wall walks of the force pass in 3D (no counterpart in the reference: a
   choice between two forms of this library's wall loop).  enable 0
   (default): the wall candidates of a wavefront's 64 particles are shared
   out over all its lanes and each particle's wall terms summed by a
   segmented sum over the lanes (on the 1M-particle 3D dam break a tenth of
   the lanes of the per-particle form have wall candidates to walk); 1: each
   particle walks its own wall candidates, adding the terms in list order,
   as brick walks and 2D scenes always do.  The two differ in the rounding
   of the wall sums only.  Applies to scenes whose fluid is first uploaded
   afterwards. */
int esph_ctx_set_wall_lanes(esph_ctx* ctx, int32_t enable);
/* The depth chosen for a class (thermal_mix: 1 plain scenes, 2 thermal, 3
   both in one batch bucket) and, if ms3 is not NULL, the force-pass times
   (ms) measured at depths 1, 2, 4.  ESPH_ERR_STATE if not tuned yet. */
int esph_ctx_tuned_depth(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision, int32_t thermal_mix,
                         int32_t* depth, float* ms3);
/* The workgroup size chosen for a class and, if ms3 is not NULL, the
   force-pass times (ms) at 64, 128 and 256 threads (at the chosen depth).
   ESPH_ERR_STATE if not tuned yet. */
int esph_ctx_tuned_workgroup(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision,
                             int32_t thermal_mix, int32_t* threads, float* ms3);
/* The gather depth and workgroup size chosen for one pass of a class (pass
   0: forces, 1: continuity, 2: Shepard) and, if ms9 is not NULL, the times
   (ms) it was chosen from: for forces, depths 1, 2, 4 at 256 threads then
   64, 128, 256 threads at the chosen depth (ms9[6..8] zero); for continuity
   (timed dry, the state untouched) and Shepard, depth {1, 2, 4} x {64, 128,
   256} threads, depth-major.  ESPH_ERR_STATE if not tuned yet. */
int esph_ctx_tuned_pass(const esph_ctx* ctx, int32_t dim, int32_t kernel, int32_t precision, int32_t thermal_mix,
                        int32_t pass, int32_t* depth, int32_t* threads, float* ms9);
/* Bytes of device memory currently held by scenes of this context. */
int esph_ctx_memory_used(const esph_ctx* ctx, uint64_t* bytes);

/* ---- kernel evaluation (SmoothKernel.jl:37-48 + kernelValue/kernelGradient)
   evaluated on the device, for checking the device math ------------------ */
int esph_kernel_eval(esph_ctx* ctx, int32_t kernel, int32_t dim, double influence_radius,
                     int64_t n, const double* r, double* w, double* dw);

/* ---- standalone neighbour list: the CellListMap.neighborlist replacement
   used by findNeighbours! (Neighbour/Neighbour.jl:88-128).
   ny == 0: pairs within x (unordered, i < j); ny > 0: pairs (i in x, j in y).
   Writes at most `capacity` pairs as (i, j) int32 to out_pairs (2 per pair)
   and their count to *m; returns ESPH_ERR_CAPACITY (with *m set) if it does
   not fit.  Pairs are sorted by (i, j). ------------------------------------ */
int esph_neighbor_list(esph_ctx* ctx, int32_t dim, double cutoff,
                       int64_t nx, const double* x, int64_t ny, const double* y,
                       int32_t* out_pairs, int64_t capacity, int64_t* m);

/* ---- scenes --------------------------------------------------------------- */
int esph_scene_create(esph_ctx* ctx, const esph_scene_desc* desc, esph_scene** out);
int esph_scene_destroy(esph_scene* scene);
/* Wall groups are static; set them (all groups at once) before the fluid. */
int esph_scene_set_walls(esph_scene* scene, int32_t ngroups, const esph_wall_arrays* groups);
/* Uploads (or re-uploads) the fluid state and resets the step counter.  For
   the inlet/outlet solver, set the buffers first. */
int esph_scene_set_fluid(esph_scene* scene, const esph_fluid_arrays* fluid);
/* Inlet and outlet buffer particles of the inlet/outlet solver (fields x, v,
   rho, p, c, mass, gap); call before esph_scene_set_fluid. */
int esph_scene_set_buffers(esph_scene* scene, const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet);
/* Advances nsteps steps of eachStep! (stream-ordered, asynchronous). */
int esph_scene_step(esph_scene* scene, int64_t nsteps);
/* Waits for the scene's work; reports ESPH_ERR_NAN if a step met a
   non-finite position. */
int esph_scene_sync(esph_scene* scene);
int esph_scene_step_count(const esph_scene* scene, int64_t* step);
int esph_scene_fluid_count(esph_scene* scene, int64_t* n);
/* Downloads a fluid field in reference order; capacity in doubles. */
int esph_scene_get_fluid(esph_scene* scene, int32_t field, double* out, int64_t capacity,
                         int64_t* written);
/* Particle count and fields of a group (ESPH_GROUP_*), in the reference's
   order for that group's vector. */
int esph_scene_group_count(esph_scene* scene, int32_t group, int64_t* n);
int esph_scene_get_group(esph_scene* scene, int32_t group, int32_t field, double* out, int64_t capacity,
                         int64_t* written);
/* Exports the pair set the next step will use (positions as they are now),
   sorted; same output contract as esph_neighbor_list. */
int esph_scene_neighbours(esph_scene* scene, int32_t pairset, int32_t* out_pairs,
                          int64_t capacity, int64_t* m);
/* The pair decisions the next step's pair passes will take themselves
   (findNeighbours!, Neighbour/Neighbour.jl:88-128, as this library makes
   it): every candidate of each live fluid particle's list that the passes
   weight with 1 -- in ESPH_FP32_BULK the quantized-geometry decision of the
   continuity and (3D) force walks, in ESPH_FP64 the binary64 rule over the
   candidate list; wall candidates always by the binary64 rule; particles
   whose lists overflowed by the cell scan.  Directed pairs: ESPH_PAIRS_FLUID_FLUID
   lists (i, j) and (j, i) when both walks take the pair, so a pair set
   following the neighbour rule gives each unordered pair of
   esph_scene_neighbours twice.  Same output contract as esph_neighbor_list;
   ESPH_ERR_ARG for the inlet/outlet solver. */
int esph_scene_walk_pairs(esph_scene* scene, int32_t pairset, int32_t* out_pairs,
                          int64_t capacity, int64_t* m);
/* kernelValueInterpolation (DataIO/VTPToDataFrame.jl:77-142) on the resident
   fluid, FP64: at each of npoints probe points (npoints*dim doubles),
   sum_j W(r) m_j / rho_j f_j / sum_j W(r) m_j / rho_j over the fluid particles
   within R (pair rule above); none_value where there is none.  field is
   ESPH_FIELD_V (dim values per point), _RHO, _P, _C or, for thermal solvers,
   _T.  Synchronous. */
int esph_scene_probe(esph_scene* scene, int32_t field, int64_t npoints, const double* points, double none_value,
                     double* out);
/* esph_scene_probe without any device allocation.  No counterpart in the
   reference: for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
   points and values are staged in a static device array of the library
   (3 MiB per GPU, held until the process ends), so npoints is at most 65,536
   (ESPH_ERR_CAPACITY beyond).  Results equal esph_scene_probe's.  Every call
   on one GPU shares that array: calls that overlap on one GPU (two host
   threads, or contexts on two streams) race.  Synchronous. */
int esph_scene_probe_noalloc(esph_scene* scene, int32_t field, int64_t npoints, const double* points,
                             double none_value, double* out);
/* Average device time per step of each phase over the steps since the last
   call (milliseconds): [0] cell sort, [1] continuity+EOS, [2] forces+integrate,
   [3] Shepard (per step it ran), [4] 0 (the check runs inside the sort);
   and the number of steps timed.
   Timing is off unless enabled (it adds events to the stream). */
int esph_scene_enable_timing(esph_scene* scene, int32_t enable);
int esph_scene_phase_times(esph_scene* scene, double* ms5, int64_t* steps);
/* out8: [0] particle-steps whose candidate count exceeded the per-particle
   list width (those particles scanned their cells instead; results are the
   same), [1] fluid-fluid list width, [2] fluid-wall list width, [3] cells,
   [4] particles more than one cell outside the grid at the last sort (they
   share the border cells; many of them mean a blown-up simulation and slow
   steps), [5] error bits (1: non-finite position, 2: inlet copies dropped at
   capacity), [6] sorts so far (each rebuilds the candidate lists; the
   upload's included), [7] exact-list width when candidate lists are reused
   across steps (0 when they are rebuilt every step). */
int esph_scene_stats(esph_scene* scene, int64_t* out8);
/* Brick walks of the scene (see esph_ctx_set_brick_walk).  out6: [0] 1 if
   the scene walks bricks, [1..3] cells per brick along x, y, z, [4] bricks
   holding particles at the last sort, [5] brick walks so far whose
   neighbour region exceeded the LDS staging (those walked their candidates
   from device memory; results are the same). */
int esph_scene_brick_stats(esph_scene* scene, int64_t* out6);
/* Graph mode: steps are captured as HIP graphs on first use and replayed --
   one step with the Shepard pass, one without, and runs of 2, 4, 8 or 16
   steps without it (the host picks from the step count; results are the same
   as with plain launches).  Phase timing is not recorded in graph mode. */
int esph_scene_enable_graph(esph_scene* scene, int32_t enable);

/* ---- batches (ensembles) ---------------------------------------------------
   A batch advances several scenes of one context together: one launch per
   (dim, kernel, precision) class and phase for all members, driven by a
   device table of scenes and tile tables (scene, first particle).  Members
   keep their state and may still be stepped alone; the result of a scene is
   the same alone or in a batch.  Members must stay alive while the batch
   exists. */
typedef struct esph_batch esph_batch;
int esph_batch_create(esph_ctx* ctx, int32_t nscenes, esph_scene* const* scenes, esph_batch** out);
int esph_batch_destroy(esph_batch* batch);
int esph_batch_enable_graph(esph_batch* batch, int32_t enable);
int esph_batch_step(esph_batch* batch, int64_t nsteps);
int esph_batch_sync(esph_batch* batch);

/* ---- serving: continuous batching of simulation requests -------------------
   No counterpart in the reference (which runs one solve! at a time).
   esph_server_submit copies a request (descriptor, walls, fluid, number of
   steps) into a FIFO queue.  esph_server_advance admits queued requests in
   arrival order while the context's HBM budget (esph_ctx_create) and
   max_active allow, advances every running request together as one batch
   (graph-replayed) by min(max_steps, fewest steps left among them), then reads
   back and frees finished requests (their x, v, rho, p, c, id [, t] stay
   available through esph_server_result until released).  A request that
   can never fit the budget fails, and so does a running request whose
   simulation blows up (a non-finite position, or more than max(64, n/64)
   particles far outside its grid).  Not thread-safe: one host thread drives
   a server. */
enum { ESPH_REQ_QUEUED = 0, ESPH_REQ_RUNNING = 1, ESPH_REQ_DONE = 2, ESPH_REQ_FAILED = 3 };
typedef struct esph_server esph_server;
int esph_server_create(esph_ctx* ctx, int32_t max_active, esph_server** out); /* max_active <= 0: no limit */
int esph_server_destroy(esph_server* server);
/* inlet and outlet: the buffers of ESPH_SOLVER_FLUID_INLET_OUTLET_WALL
   requests (both NULL for the other solvers); results are the fluid group. */
int esph_server_submit(esph_server* server, const esph_scene_desc* desc, int32_t nwall_groups,
                       const esph_wall_arrays* walls, const esph_fluid_arrays* fluid,
                       const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet, int64_t nsteps,
                       int64_t* request_id);
int esph_server_advance(esph_server* server, int64_t max_steps, int64_t* stepped);
int esph_server_status(esph_server* server, int64_t request_id, int32_t* state, int64_t* steps_done);
/* Fields of a finished request in reference order (ESPH_FIELD_X, _V, _RHO, _P,
   _C, _ID, and _T for thermal solvers); out may be NULL to query *written. */
int esph_server_result(esph_server* server, int64_t request_id, int32_t field, double* out, int64_t capacity,
                       int64_t* written);
int esph_server_release(esph_server* server, int64_t request_id);
/* out8: [0] p50 and [1] p99 step latency (ms, over every scene-step),
   [2] completed requests, [3] completed per second since the first submit,
   [4] running, [5] queued, [6] device bytes held, [7] batch steps run. */
int esph_server_stats(esph_server* server, double* out8);

#ifdef __cplusplus
}
#endif

#endif /* ETHERSPH_H */
