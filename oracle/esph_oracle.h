/*
 * esph_oracle.h — FP64 CPU restatement of EtherSPHCPU.jl's per-step particle
 * hot path.  TEST INFRASTRUCTURE ONLY: the oracle is the checker that the
 * tests, the smoke check and the CPU-baseline leg of the benchmark compare
 * the HIP library against.  Nothing in the product links or calls it.
 *
 * It follows the reference's structure (a pair list per step, push-style
 * updates of both ends of each unordered pair, threaded with atomics like
 * Threads.atomic_add!) so that it is also the reported CPU baseline.  With
 * nthreads == 1 it is deterministic (pair order (i, j) ascending).
 */
#ifndef ESPH_ORACLE_H
#define ESPH_ORACLE_H

#include <stdint.h>
#include "../include/ethersph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SmoothKernel (SPHKernel/SmoothKernel/SmoothKernel.jl:37-48) */
typedef struct eo_kernel {
    int32_t kind;
    int32_t dim;
    double h;
    double radius_ratio;
    double influence_radius;
    double sigma;
    double kernel_0;
} eo_kernel;

int eo_kernel_init(eo_kernel* k, int32_t kind, double influence_radius, int32_t dim);
double eo_kernel_value(const eo_kernel* k, double r);
double eo_kernel_gradient(const eo_kernel* k, double r);
/* Vector form for tests; also writes h, sigma, kernel_0 into params3. */
int eo_kernel_eval(int32_t kind, int32_t dim, double influence_radius, int64_t n,
                   const double* r, double* w, double* dw, double* params3);

/* Neighbour list under the rule stated in include/ethersph.h.
   brute != 0: O(n^2) double loop; else a cell list.  Returns the number of
   pairs (writes at most cap pairs, sorted by (i, j)); -1 on bad input. */
int64_t eo_neighbor_list(int32_t dim, double cutoff, int64_t nx, const double* x,
                         int64_t ny, const double* y, int32_t brute, int32_t nthreads,
                         int32_t* out_pairs, int64_t cap);

typedef struct eo_scene eo_scene;

eo_scene* eo_scene_new(const esph_scene_desc* desc);
void eo_scene_free(eo_scene* s);
int eo_scene_set_walls(eo_scene* s, int32_t ngroups, const esph_wall_arrays* groups);
int eo_scene_set_fluid(eo_scene* s, const esph_fluid_arrays* f);
/* Inlet/outlet buffers of ESPH_SOLVER_FLUID_INLET_OUTLET_WALL. */
int eo_scene_set_buffers(eo_scene* s, const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet);
int64_t eo_scene_group_count(const eo_scene* s, int32_t group);
int eo_scene_get_group(const eo_scene* s, int32_t group, int32_t field, double* out);
/* Runs nsteps of eachStep!; returns 0. */
int eo_scene_step(eo_scene* s, int64_t nsteps, int32_t nthreads);
int64_t eo_scene_fluid_count(const eo_scene* s);
int64_t eo_scene_step_count(const eo_scene* s);
/* Field in reference order; out holds n (or n*dim) doubles. */
int eo_scene_get_fluid(const eo_scene* s, int32_t field, double* out);
/* SPH interpolation of a fluid field (V: dim values per point) at np probe
   points (VTPToDataFrame.jl:77-142); none_value where no particle is within R. */
int eo_scene_probe(const eo_scene* s, int32_t field, int64_t np, const double* pts, double none_value,
                   double* out);
/* Pair set the next step would use; sorted; returns count (see above). */
int64_t eo_scene_neighbours(eo_scene* s, int32_t pairset, int32_t* out_pairs, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif
