/*
 * esph_oracle.c — FP64 CPU restatement of EtherSPHCPU.jl's per-step hot path.
 * TEST INFRASTRUCTURE ONLY (the checker); see esph_oracle.h.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * written in the reference's left-to-right evaluation order; Julia's
 * `x^2`/`x^3` literal powers are x*x and x*x*x, its `x^n` for a variable
 * integer n is restated as binary powering (ulp-level difference only).
 * Build with -ffp-contract=off so no multiply-add is fused.
 */
#include "esph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Smoothing kernels                                                         */
/* ------------------------------------------------------------------------ */

/* kSmoothKernelParametersDict, SmoothKernel.jl:15-20 */
static int kernel_table(int32_t kind, int32_t dim, double* ratio, double* cdim) {
    static const double PI = 3.14159265358979323846;
    double c[3];
    switch (kind) {
        case ESPH_CUBIC_SPLINE:
            *ratio = 2.0; c[0] = 2.0 / 3.0; c[1] = 10.0 / 7.0 / PI; c[2] = 1.0 / PI; break;
        case ESPH_GAUSSIAN:
            *ratio = 3.0; c[0] = 1.0 / sqrt(PI); c[1] = 1.0 / PI; c[2] = 1.0 / sqrt(PI * PI * PI); break;
        case ESPH_WENDLAND_C2:
            *ratio = 2.0; c[0] = 0.0; c[1] = 7.0 / 4.0 / PI; c[2] = 21.0 / 16.0 / PI; break;
        case ESPH_WENDLAND_C4:
            *ratio = 2.0; c[0] = 5.0 / 8.0; c[1] = 9.0 / 4.0 / PI; c[2] = 495.0 / 256.0 / PI; break;
        default: return -1;
    }
    if (dim < 1 || dim > 3) return -1;
    *cdim = c[dim - 1];
    return 0;
}

/* SmoothKernel(influence_radius, dim, K), SmoothKernel.jl:37-48 */
int eo_kernel_init(eo_kernel* k, int32_t kind, double influence_radius, int32_t dim) {
    double ratio, cdim;
    if (kernel_table(kind, dim, &ratio, &cdim)) return -1;
    k->kind = kind;
    k->dim = dim;
    k->radius_ratio = ratio;
    k->influence_radius = influence_radius;
    k->h = influence_radius / ratio;
    double hd = k->h;
    if (dim == 2) hd = k->h * k->h;
    if (dim == 3) hd = k->h * k->h * k->h;
    k->sigma = cdim / hd;
    k->kernel_0 = k->sigma;
    return 0;
}

/* kernelValue: CubicSpline.jl:39-48, Gaussian.jl:31-38, WendlandC2.jl:38-45,
   WendlandC4.jl:38-45 */
double eo_kernel_value(const eo_kernel* k, double r) {
    double q = r / k->h;
    switch (k->kind) {
        case ESPH_CUBIC_SPLINE:
            if (q < 1.0) return k->sigma / 4.0 * (3.0 * (q * q) * (q - 2.0) + 4.0);
            if (q < 2.0) { double t = 2.0 - q; return k->sigma / 4.0 * (t * t * t); }
            return 0.0;
        case ESPH_GAUSSIAN:
            if (q < 3.0) return k->sigma * exp(-(q * q));
            return 0.0;
        case ESPH_WENDLAND_C2:
            if (q < 2.0) {
                double t = 2.0 - q, t2 = t * t;
                return k->sigma * (t2 * t2) * (1.0 + 2.0 * q) / 16.0;
            }
            return 0.0;
        case ESPH_WENDLAND_C4:
            if (q < 2.0) {
                double t = 2.0 - q, t2 = t * t, t6 = t2 * t2 * t2;
                return k->sigma * t6 * (35.0 * (q * q) + 36.0 * q + 12.0) / 768.0;
            }
            return 0.0;
    }
    return 0.0;
}

/* kernelGradient: CubicSpline.jl:64-73, Gaussian.jl:49-56, WendlandC2.jl:56-63,
   WendlandC4.jl:47-54 */
double eo_kernel_gradient(const eo_kernel* k, double r) {
    double q = r / k->h;
    switch (k->kind) {
        case ESPH_CUBIC_SPLINE:
            if (q < 1.0) return k->sigma / k->h * 3.0 / 4.0 * q * (3.0 * q - 4.0);
            if (q < 2.0) { double t = 2.0 - q; return -k->sigma / k->h * 3.0 / 4.0 * (t * t); }
            return 0.0;
        case ESPH_GAUSSIAN:
            if (q < 3.0) return -2.0 * k->sigma / k->h * q * exp(-(q * q));
            return 0.0;
        case ESPH_WENDLAND_C2:
            if (q < 2.0) { double t = 2.0 - q; return -k->sigma / k->h * 5.0 / 8.0 * q * (t * t * t); }
            return 0.0;
        case ESPH_WENDLAND_C4:
            if (q < 2.0) {
                double t = 2.0 - q, t2 = t * t, t5 = t2 * t2 * t;
                return -k->sigma / k->h * (35.0 / 96.0 * (q * q) + 7.0 / 48.0 * q) * t5;
            }
            return 0.0;
    }
    return 0.0;
}

int eo_kernel_eval(int32_t kind, int32_t dim, double influence_radius, int64_t n,
                   const double* r, double* w, double* dw, double* params3) {
    eo_kernel k;
    if (eo_kernel_init(&k, kind, influence_radius, dim)) return -1;
    for (int64_t i = 0; i < n; ++i) {
        if (w) w[i] = eo_kernel_value(&k, r[i]);
        if (dw) dw[i] = eo_kernel_gradient(&k, r[i]);
    }
    if (params3) { params3[0] = k.h; params3[1] = k.sigma; params3[2] = k.kernel_0; }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Neighbour search (replaces CellListMap.neighborlist; rule in ethersph.h)  */
/* ------------------------------------------------------------------------ */

static inline double dist2(int dim, const double* a, const double* b) {
    double d0 = a[0] - b[0];
    double s = d0 * d0;
    double d1 = a[1] - b[1];
    s = s + d1 * d1;
    if (dim == 3) { double d2 = a[2] - b[2]; s = s + d2 * d2; }
    return s;
}

typedef struct { int64_t cnt; int64_t* i; int64_t* j; } pairvec;

typedef struct {
    int dim;
    double lo[3], inv;
    int64_t nc[3];
    int64_t* start; /* ncell+1 */
    int64_t* idx;   /* particle indices sorted by cell */
} cellgrid;

static int64_t cell_coord(const cellgrid* g, int k, double x) {
    double f = floor((x - g->lo[k]) * g->inv);
    if (!(f >= 0.0)) return 0; /* also NaN */
    if (f >= (double)(g->nc[k] - 1)) return g->nc[k] - 1;
    return (int64_t)f;
}

static int64_t cell_of(const cellgrid* g, const double* x) {
    int64_t c = cell_coord(g, 0, x[0]) + g->nc[0] * cell_coord(g, 1, x[1]);
    if (g->dim == 3) c += g->nc[0] * g->nc[1] * cell_coord(g, 2, x[2]);
    return c;
}

/* Grid over the bounding box of both sets; cells a little wider than the
   cutoff so that neighbours are never more than one cell apart. */
static void grid_build(cellgrid* g, int dim, double cutoff, int64_t n, const double* p,
                       int64_t n2, const double* p2) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int pass = 0; pass < 2; ++pass) {
        const double* q = pass ? p2 : p;
        int64_t m = pass ? n2 : n;
        for (int64_t i = 0; i < m; ++i)
            for (int k = 0; k < dim; ++k) {
                double v = q[i * dim + k];
                if (isfinite(v)) { if (v < lo[k]) lo[k] = v; if (v > hi[k]) hi[k] = v; }
            }
    }
    g->dim = dim;
    double cs = cutoff * 1.0001;
    g->inv = 1.0 / cs;
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
        if (k >= dim || !(lo[k] <= hi[k])) { g->lo[k] = 0; g->nc[k] = 1; continue; }
        g->lo[k] = lo[k];
        int64_t c = (int64_t)floor((hi[k] - lo[k]) / cs) + 1;
        if (c > 4096) c = 4096; /* clamping keeps the search exact, only slower */
        g->nc[k] = c;
        ncell *= c;
    }
    if (dim == 2) g->nc[2] = 1;
    g->start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    g->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { key[i] = cell_of(g, p + i * dim); g->start[key[i] + 1]++; }
    for (int64_t c = 0; c < ncell; ++c) g->start[c + 1] += g->start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)(ncell + 1));
    for (int64_t i = 0; i < n; ++i) g->idx[fill[key[i]]++] = i;
    free(fill);
    free(key);
}

static void grid_free(cellgrid* g) { free(g->start); free(g->idx); }

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Builds the CSR list nbr[start[i]..start[i+1]) of neighbours of x_i among
   `y` (self == 1: y is x and only j > i is kept), sorted ascending. */
static void neighbour_csr(int dim, double cutoff, int64_t nx, const double* x, int64_t ny,
                          const double* y, int self, int brute, int nthreads,
                          int64_t** out_start, int64_t** out_nbr) {
    const double R2 = cutoff * cutoff;
    int64_t* start = (int64_t*)calloc((size_t)nx + 1, sizeof(int64_t));
    cellgrid g;
    memset(&g, 0, sizeof(g));
    if (!brute) grid_build(&g, dim, cutoff, ny, y, nx, x);
    /* two passes: count, then fill */
    int64_t* nbr = NULL;
    for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) if (nthreads > 1)
        for (int64_t i = 0; i < nx; ++i) {
            const double* xi = x + i * dim;
            int64_t cnt = 0;
            int64_t base = pass ? start[i] : 0;
            if (brute) {
                for (int64_t j = self ? i + 1 : 0; j < ny; ++j)
                    if (dist2(dim, xi, y + j * dim) <= R2) {
                        if (pass) nbr[base + cnt] = j;
                        cnt++;
                    }
            } else {
                int64_t c[3] = {0, 0, 0};
                for (int k = 0; k < dim; ++k) c[k] = cell_coord(&g, k, xi[k]);
                int64_t zlo = dim == 3 ? -1 : 0, zhi = dim == 3 ? 1 : 0;
                for (int64_t dz = zlo; dz <= zhi; ++dz)
                    for (int64_t dy = -1; dy <= 1; ++dy)
                        for (int64_t dx = -1; dx <= 1; ++dx) {
                            int64_t a = c[0] + dx, b = c[1] + dy, e = c[2] + dz;
                            if (a < 0 || b < 0 || e < 0 || a >= g.nc[0] || b >= g.nc[1] || e >= g.nc[2]) continue;
                            int64_t cell = a + g.nc[0] * (b + g.nc[1] * e);
                            for (int64_t s = g.start[cell]; s < g.start[cell + 1]; ++s) {
                                int64_t j = g.idx[s];
                                if (self && j <= i) continue;
                                if (dist2(dim, xi, y + j * dim) <= R2) {
                                    if (pass) nbr[base + cnt] = j;
                                    cnt++;
                                }
                            }
                        }
                if (pass) qsort(nbr + base, (size_t)cnt, sizeof(int64_t), cmp_i64);
            }
            if (!pass) start[i + 1] = cnt;
        }
        if (!pass) {
            for (int64_t i = 0; i < nx; ++i) start[i + 1] += start[i];
            nbr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(start[nx] > 0 ? start[nx] : 1));
        }
    }
    if (!brute) grid_free(&g);
    *out_start = start;
    *out_nbr = nbr;
}

int64_t eo_neighbor_list(int32_t dim, double cutoff, int64_t nx, const double* x, int64_t ny,
                         const double* y, int32_t brute, int32_t nthreads, int32_t* out_pairs,
                         int64_t cap) {
    if ((dim != 2 && dim != 3) || nx < 0 || ny < 0 || !(cutoff > 0)) return -1;
    int self = ny == 0;
    int64_t *start, *nbr;
    neighbour_csr(dim, cutoff, nx, x, self ? nx : ny, self ? x : y, self, brute,
                  nthreads < 1 ? 1 : nthreads, &start, &nbr);
    int64_t m = start[nx];
    if (out_pairs) {
        int64_t w = 0;
        for (int64_t i = 0; i < nx && w < cap; ++i)
            for (int64_t s = start[i]; s < start[i + 1] && w < cap; ++s, ++w) {
                out_pairs[2 * w] = (int32_t)i;
                out_pairs[2 * w + 1] = (int32_t)nbr[s];
            }
    }
    free(start);
    free(nbr);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Scene state (reference particle vectors as SoA, reference order)          */
/* ------------------------------------------------------------------------ */

/* A buffer particle group of the inlet/outlet solver (a Vector of
   CommonLiquidParticle in the reference), reference order. */
typedef struct eo_buf {
    int64_t n, cap, next_id;
    double *x, *v, *rho, *p, *c, *m, *gap;
    int64_t* id;
} eo_buf;

/* Pair records fluid i -- buffer particle j (both movable: v_ij = v_i - v_j). */
typedef struct eo_bpairs {
    int64_t n;
    int64_t *i, *j;
    double *r, *W, *dW, *g, *vv; /* g, vv: dim per pair */
} eo_bpairs;

static void buf_free(eo_buf* b);
static void bpairs_free(eo_bpairs* q);

struct eo_scene {
    esph_scene_desc d;
    eo_kernel k;
    double b, nu_0; /* LiquidModel.jl:30-31 */
    int thermal;    /* solver carries ThermalWeaklyCompressibleLiquidModel */
    int64_t step;
    /* fluid */
    int64_t n;
    double *x, *v, *dv, *rho, *drho, *p, *c, *m, *gap, *t, *dtt, *kappa, *cp, *sw, *ss;
    int64_t* id;
    /* walls, groups concatenated */
    int64_t nw;
    double *wx, *wn, *wgap, *wv, *wt, *wkappa;
    int32_t* wkind;
    /* pair records of the current step (Neighbour.jl:31-40) */
    int64_t nff, nfw;
    int64_t *ffi, *ffj, *fwi, *fwj;
    double *ffr, *ffW, *ffdW, *ffg, *ffvv, *ffrv; /* g, vv, rv: dim per pair */
    double *fwr, *fwW, *fwdW, *fwg, *fwvv, *fwrv;
    /* inlet/outlet solver (FluidInletOutletWallSolve.jl) */
    int64_t fcap;        /* allocated fluid capacity (the fluid grows by inlet copies) */
    int64_t next_id;     /* id of the next fluid particle pushed */
    eo_buf inlet, outlet;
    eo_bpairs fi, fo;    /* fluid-inlet and fluid-outlet pair records */
};

static double* dup_or_zero(const double* src, int64_t n) {
    double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (src && n > 0) memcpy(p, src, sizeof(double) * (size_t)n);
    return p;
}

eo_scene* eo_scene_new(const esph_scene_desc* desc) {
    eo_scene* s = (eo_scene*)calloc(1, sizeof(eo_scene));
    s->d = *desc;
    if (eo_kernel_init(&s->k, desc->kernel, desc->influence_radius, desc->dim)) { free(s); return NULL; }
    /* CommonWeaklyCompressibleLiquidModel, LiquidModel.jl:30-31 */
    s->b = desc->c_0 * desc->c_0 * desc->rho_0 / (double)desc->gamma;
    s->nu_0 = desc->mu_0 / desc->rho_0;
    s->thermal = desc->solver == ESPH_SOLVER_FLUID_THERMOSTATIC_WALL ||
                 desc->solver == ESPH_SOLVER_FLUID_COMPULSIVE_THERMOSTATIC_WALL;
    return s;
}

static void free_pairs(eo_scene* s) {
    free(s->ffi); free(s->ffj); free(s->ffr); free(s->ffW); free(s->ffdW); free(s->ffg); free(s->ffvv); free(s->ffrv);
    free(s->fwi); free(s->fwj); free(s->fwr); free(s->fwW); free(s->fwdW); free(s->fwg); free(s->fwvv); free(s->fwrv);
    s->ffi = s->ffj = s->fwi = s->fwj = NULL;
    s->ffr = s->ffW = s->ffdW = s->ffg = s->ffvv = s->ffrv = NULL;
    s->fwr = s->fwW = s->fwdW = s->fwg = s->fwvv = s->fwrv = NULL;
    s->nff = s->nfw = 0;
}

static void free_fluid(eo_scene* s) {
    free(s->x); free(s->v); free(s->dv); free(s->rho); free(s->drho); free(s->p); free(s->c); free(s->m);
    free(s->gap); free(s->t); free(s->dtt); free(s->kappa); free(s->cp); free(s->sw); free(s->ss); free(s->id);
    s->x = NULL;
}

void eo_scene_free(eo_scene* s) {
    if (!s) return;
    free_fluid(s);
    buf_free(&s->inlet);
    buf_free(&s->outlet);
    bpairs_free(&s->fi);
    bpairs_free(&s->fo);
    free(s->wx); free(s->wn); free(s->wgap); free(s->wv); free(s->wt); free(s->wkappa); free(s->wkind);
    free_pairs(s);
    free(s);
}

int eo_scene_set_walls(eo_scene* s, int32_t ngroups, const esph_wall_arrays* g) {
    const int D = s->d.dim;
    int64_t nw = 0;
    for (int q = 0; q < ngroups; ++q) nw += g[q].n;
    free(s->wx); free(s->wn); free(s->wgap); free(s->wv); free(s->wt); free(s->wkappa); free(s->wkind);
    s->nw = nw;
    s->wx = dup_or_zero(NULL, nw * D); s->wn = dup_or_zero(NULL, nw * D); s->wgap = dup_or_zero(NULL, nw);
    s->wv = dup_or_zero(NULL, nw * D); s->wt = dup_or_zero(NULL, nw); s->wkappa = dup_or_zero(NULL, nw);
    s->wkind = (int32_t*)calloc((size_t)(nw > 0 ? nw : 1), sizeof(int32_t));
    int64_t o = 0;
    for (int q = 0; q < ngroups; ++q) {
        int64_t n = g[q].n;
        memcpy(s->wx + o * D, g[q].x, sizeof(double) * (size_t)(n * D));
        memcpy(s->wn + o * D, g[q].normal, sizeof(double) * (size_t)(n * D));
        memcpy(s->wgap + o, g[q].gap, sizeof(double) * (size_t)n);
        if (g[q].kind == ESPH_WALL_STATIC_VELOCITY && g[q].v) memcpy(s->wv + o * D, g[q].v, sizeof(double) * (size_t)(n * D));
        if (g[q].kind == ESPH_WALL_THERMOSTATIC) {
            if (g[q].t) memcpy(s->wt + o, g[q].t, sizeof(double) * (size_t)n);
            if (g[q].kappa) memcpy(s->wkappa + o, g[q].kappa, sizeof(double) * (size_t)n);
        }
        for (int64_t i = 0; i < n; ++i) s->wkind[o + i] = g[q].kind;
        o += n;
    }
    return 0;
}

int eo_scene_set_fluid(eo_scene* s, const esph_fluid_arrays* f) {
    const int D = s->d.dim;
    const int64_t n = f->n;
    free_fluid(s);
    s->n = n;
    s->x = dup_or_zero(f->x, n * D); s->v = dup_or_zero(f->v, n * D); s->dv = dup_or_zero(NULL, n * D);
    s->rho = dup_or_zero(f->rho, n); s->drho = dup_or_zero(NULL, n); s->p = dup_or_zero(f->p, n);
    s->c = dup_or_zero(f->c, n); s->m = dup_or_zero(f->mass, n); s->gap = dup_or_zero(f->gap, n);
    s->t = dup_or_zero(f->t, n); s->dtt = dup_or_zero(NULL, n); s->kappa = dup_or_zero(f->kappa, n);
    s->cp = dup_or_zero(f->cp, n); s->sw = dup_or_zero(NULL, n); s->ss = dup_or_zero(NULL, n);
    s->id = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) s->id[i] = i;
    s->fcap = n > 0 ? n : 1;
    s->next_id = n;
    s->step = 0;
    return 0;
}

static void* grow(void* p, int64_t n, size_t elem) {
    void* q = realloc(p, (size_t)(n > 0 ? n : 1) * elem);
    return q;
}

/* Makes room for `need` fluid particles (push! onto the fluid vector). */
static void fluid_reserve(eo_scene* s, int64_t need) {
    if (need <= s->fcap) return;
    int64_t cap = s->fcap;
    while (cap < need) cap *= 2;
    const int D = s->d.dim;
    double** vec[] = {&s->x, &s->v, &s->dv};
    for (int q = 0; q < 3; ++q) *vec[q] = (double*)grow(*vec[q], cap * D, sizeof(double));
    double** sc[] = {&s->rho, &s->drho, &s->p, &s->c, &s->m, &s->gap, &s->t, &s->dtt, &s->kappa, &s->cp, &s->sw, &s->ss};
    for (int q = 0; q < 12; ++q) *sc[q] = (double*)grow(*sc[q], cap, sizeof(double));
    s->id = (int64_t*)grow(s->id, cap, sizeof(int64_t));
    s->fcap = cap;
}

static void buf_free(eo_buf* b) {
    free(b->x); free(b->v); free(b->rho); free(b->p); free(b->c); free(b->m); free(b->gap); free(b->id);
    memset(b, 0, sizeof(*b));
}

static void buf_reserve(eo_buf* b, int D, int64_t need) {
    if (need <= b->cap) return;
    int64_t cap = b->cap > 0 ? b->cap : 1;
    while (cap < need) cap *= 2;
    b->x = (double*)grow(b->x, cap * D, sizeof(double));
    b->v = (double*)grow(b->v, cap * D, sizeof(double));
    b->rho = (double*)grow(b->rho, cap, sizeof(double));
    b->p = (double*)grow(b->p, cap, sizeof(double));
    b->c = (double*)grow(b->c, cap, sizeof(double));
    b->m = (double*)grow(b->m, cap, sizeof(double));
    b->gap = (double*)grow(b->gap, cap, sizeof(double));
    b->id = (int64_t*)grow(b->id, cap, sizeof(int64_t));
    b->cap = cap;
}

static void buf_set(eo_buf* b, int D, const esph_fluid_arrays* f) {
    buf_free(b);
    const int64_t n = f ? f->n : 0;
    buf_reserve(b, D, n > 0 ? n : 1);
    b->n = n;
    if (n > 0) {
        memcpy(b->x, f->x, sizeof(double) * (size_t)(n * D));
        memcpy(b->v, f->v, sizeof(double) * (size_t)(n * D));
        memcpy(b->rho, f->rho, sizeof(double) * (size_t)n);
        memcpy(b->p, f->p, sizeof(double) * (size_t)n);
        memcpy(b->c, f->c, sizeof(double) * (size_t)n);
        memcpy(b->m, f->mass, sizeof(double) * (size_t)n);
        memcpy(b->gap, f->gap, sizeof(double) * (size_t)n);
    }
    for (int64_t i = 0; i < n; ++i) b->id[i] = i;
    b->next_id = n;
}

int eo_scene_set_buffers(eo_scene* s, const esph_fluid_arrays* inlet, const esph_fluid_arrays* outlet) {
    buf_set(&s->inlet, s->d.dim, inlet);
    buf_set(&s->outlet, s->d.dim, outlet);
    return 0;
}

static void bpairs_free(eo_bpairs* q) {
    free(q->i); free(q->j); free(q->r); free(q->W); free(q->dW); free(q->g); free(q->vv);
    memset(q, 0, sizeof(*q));
}

/* ------------------------------------------------------------------------ */
/* Pair records: findNeighbours! / CommonNeighbour!  (Neighbour.jl:42-128)   */
/* ------------------------------------------------------------------------ */

/* relativeVelocityVector, Neighbour.jl:10-29: fluid-fluid v_i - v_j; fixed
   wall v_i; static-velocity wall v_i - v_wall. */
static void make_records(eo_scene* s, int wall, int nthreads) {
    const int D = s->d.dim;
    int64_t *start, *nbr;
    if (!wall)
        neighbour_csr(D, s->k.influence_radius, s->n, s->x, s->n, s->x, 1, 0, nthreads, &start, &nbr);
    else
        neighbour_csr(D, s->k.influence_radius, s->n, s->x, s->nw, s->wx, 0, 0, nthreads, &start, &nbr);
    int64_t m = start[s->n];
    int64_t* pi = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int64_t* pj = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    double* r = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* W = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* dW = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* gv = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m * D : 1));
    double* vv = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m * D : 1));
    double* rv = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m * D : 1));
    const double* yx = wall ? s->wx : s->x;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < s->n; ++i) {
        for (int64_t q = start[i]; q < start[i + 1]; ++q) {
            int64_t j = nbr[q];
            pi[q] = i;
            pj[q] = j;
            double d2 = dist2(D, s->x + i * D, yx + j * D);
            double rr = sqrt(d2); /* r from the neighbour search */
            r[q] = rr;
            double nrm2 = 0.0;
            for (int k = 0; k < D; ++k) {
                double a = s->x[i * D + k] - yx[j * D + k]; /* r_vec = x_i - x_j */
                rv[q * D + k] = a;
                nrm2 = nrm2 + a * a;
                double vj = wall ? (s->wkind[j] == ESPH_WALL_STATIC_VELOCITY ? s->wv[j * D + k] : 0.0)
                                 : s->v[j * D + k];
                if (wall && s->wkind[j] != ESPH_WALL_STATIC_VELOCITY)
                    vv[q * D + k] = s->v[i * D + k]; /* alias of v_i */
                else
                    vv[q * D + k] = s->v[i * D + k] - vj;
            }
            W[q] = eo_kernel_value(&s->k, rr);
            dW[q] = eo_kernel_gradient(&s->k, rr);
            /* kernel_gradient * normalize(r_vec), NaN components -> 0 (:53-59) */
            double nrm = sqrt(nrm2);
            double inv = 1.0 / nrm;
            for (int k = 0; k < D; ++k) {
                double gk = dW[q] * (rv[q * D + k] * inv);
                if (nrm == 0.0 || isnan(gk)) gk = 0.0;
                gv[q * D + k] = gk;
            }
        }
    }
    free(start);
    free(nbr);
    if (!wall) {
        s->nff = m; s->ffi = pi; s->ffj = pj; s->ffr = r; s->ffW = W; s->ffdW = dW; s->ffg = gv; s->ffvv = vv; s->ffrv = rv;
    } else {
        s->nfw = m; s->fwi = pi; s->fwj = pj; s->fwr = r; s->fwW = W; s->fwdW = dW; s->fwg = gv; s->fwvv = vv; s->fwrv = rv;
    }
}

/* ------------------------------------------------------------------------ */
/* Particle actions                                                          */
/* ------------------------------------------------------------------------ */

static inline void atomic_add(double* p, double v) {
#pragma omp atomic
    *p += v;
}

static inline double ipow(double x, int n) {
    /* x^n for integer n >= 0 (UpdatePressure.jl:13-14 uses x^gamma) */
    double y = 1.0;
    while (n > 0) {
        if (n & 1) y *= x;
        x *= x;
        n >>= 1;
    }
    return y;
}

/* updateDensity! (UpdateDensity.jl:8-12) + updatePressure! (UpdatePressure.jl:8-16) */
static void update_density_pressure(eo_scene* s, int64_t i) {
    s->rho[i] += s->drho[i] * s->d.dt;
    s->drho[i] = 0.0;
    double rr = s->rho[i] / s->d.rho_0;
    s->p[i] = s->b * (ipow(rr, s->d.gamma) - 1.0) + s->d.p_0;
    s->c[i] = s->d.c_0 * sqrt(ipow(rr, s->d.gamma - 1));
}

/* Pair records fluid i -- buffer particle j for the inlet/outlet solver
   (findNeighbours!(…, fluid, inlet/outlet, …), FluidInletOutletWallSolve.jl:41-61). */
static void make_buffer_records(eo_scene* s, const eo_buf* b, eo_bpairs* out, int nthreads) {
    const int D = s->d.dim;
    bpairs_free(out);
    int64_t *start, *nbr;
    neighbour_csr(D, s->k.influence_radius, s->n, s->x, b->n, b->x, 0, 0, nthreads, &start, &nbr);
    const int64_t m = start[s->n];
    out->n = m;
    out->i = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    out->j = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    out->r = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    out->W = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    out->dW = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    out->g = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m * D : 1));
    out->vv = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m * D : 1));
    for (int64_t i = 0; i < s->n; ++i)
        for (int64_t q = start[i]; q < start[i + 1]; ++q) {
            const int64_t j = nbr[q];
            out->i[q] = i;
            out->j[q] = j;
            const double d2 = dist2(D, s->x + i * D, b->x + j * D);
            const double rr = sqrt(d2);
            out->r[q] = rr;
            out->W[q] = eo_kernel_value(&s->k, rr);
            out->dW[q] = eo_kernel_gradient(&s->k, rr);
            double rv[3], nrm2 = 0.0;
            for (int k = 0; k < D; ++k) {
                rv[k] = s->x[i * D + k] - b->x[j * D + k];
                nrm2 = nrm2 + rv[k] * rv[k];
                out->vv[q * D + k] = s->v[i * D + k] - b->v[j * D + k];
            }
            const double nrm = sqrt(nrm2), inv = 1.0 / nrm;
            for (int k = 0; k < D; ++k) {
                double gk = out->dW[q] * (rv[k] * inv);
                if (nrm == 0.0 || isnan(gk)) gk = 0.0;
                out->g[q * D + k] = gk;
            }
        }
    free(start);
    free(nbr);
}

/* continuity!(fluid_i, buffer_j, …) (Continuity.jl:8-17); the buffer's own
   drho is reset afterwards (FluidInletOutletWallSolve.jl:91-99), so only
   the fluid side is kept. */
static void buffer_continuity(eo_scene* s, const eo_buf* b, const eo_bpairs* q) {
    const int D = s->d.dim;
    for (int64_t e = 0; e < q->n; ++e) {
        const int64_t i = q->i[e], j = q->j[e];
        double drho = 0.0;
        for (int k = 0; k < D; ++k) drho = drho + q->vv[e * D + k] * q->g[e * D + k];
        s->drho[i] += drho * b->m[j];
    }
}

/* pressureForce!/viscosityForce!(fluid_i, buffer_j, …), the fluid-fluid forms
   (PressureForce.jl:8-21, ViscosityForce.jl:8-25); the buffer side is reset. */
static void buffer_forces(eo_scene* s, const eo_buf* b, const eo_bpairs* q) {
    const int D = s->d.dim;
    const double h = s->k.h;
    for (int64_t e = 0; e < q->n; ++e) {
        const int64_t i = q->i[e], j = q->j[e];
        const double* g = q->g + e * D;
        const double* vv = q->vv + e * D;
        const double r = q->r[e], W = q->W[e], dW = q->dW[e];
        double p_rho_2 = s->p[i] / (s->rho[i] * s->rho[i]) + b->p[j] / (b->rho[j] * b->rho[j]);
        p_rho_2 += fabs(p_rho_2) * 0.01 * W / eo_kernel_value(&s->k, (s->gap[i] + b->gap[j]) / 2.0);
        for (int k = 0; k < D; ++k) s->dv[i * D + k] += -b->m[j] * p_rho_2 * g[k];
        const double sum_rho = s->rho[i] + b->rho[j];
        const double mu_i = s->nu_0 * s->rho[i];
        const double mu_j = s->nu_0 * b->rho[j];
        const double vf = 4.0 * (mu_i + mu_j) * r * dW / (sum_rho * sum_rho) / (r * r + 0.01 * (h * h));
        for (int k = 0; k < D; ++k) s->dv[i * D + k] += b->m[j] * vf * vv[k];
    }
}

/* The buffer velocity profile (user_defined_function of the examples):
   PoiseuilleFlow2D.jl:28-35 (uxTheory(y, t)), PoiseuilleFlow3D.jl:32-34. */
static double profile_u(const eo_scene* s, const double* x, double t) {
    const esph_scene_desc* d = &s->d;
    const double f = d->profile_f, nu = d->profile_nu, l = d->profile_l;
    if (d->profile == ESPH_PROFILE_CHANNEL_2D) {
        const double PI = 3.14159265358979323846;
        const double y = x[1] - d->profile_center[1];
        double ux = f / 2.0 / nu * y * (l - y);
        for (int n = 0; n <= d->profile_terms; ++n) {
            const double o = (double)(2 * n + 1);
            ux -= 4.0 * f * (l * l) / nu / (PI * PI * PI) / (o * o * o) * sin(PI * y / l * o) *
                  exp(-(o * o) * (PI * PI) * nu * t / (l * l));
        }
        return ux;
    }
    if (d->profile == ESPH_PROFILE_PIPE_3D) {
        const double y = x[1] - d->profile_center[1], z = x[2] - d->profile_center[2];
        return f / 4.0 / nu * (l * l - y * y - z * z);
    }
    return 0.0;
}

static void buf_push(eo_buf* dst, int D, const double* x, const double* v, double rho, double p, double c,
                     double m, double gap) {
    buf_reserve(dst, D, dst->n + 1);
    const int64_t w = dst->n++;
    for (int k = 0; k < D; ++k) {
        dst->x[w * D + k] = x[k];
        dst->v[w * D + k] = v[k];
    }
    dst->rho[w] = rho; dst->p[w] = p; dst->c[w] = c; dst->m[w] = m; dst->gap[w] = gap;
    dst->id[w] = dst->next_id++;
}

/* Inlet/outlet migration and the buffer velocity update after a step
   (FluidInletOutletWallSolve.jl:196-262): collect the three index sets, push
   inlet copies onto the fluid and fluid copies onto the outlet, delete (in
   order) the fluid particles that left and the outlet particles that exited,
   move the entered inlet particles back, then set buffer velocities at
   t = step * dt. */
static void migrate(eo_scene* s) {
    const int D = s->d.dim, a = s->d.io_axis;
    eo_buf* in = &s->inlet;
    eo_buf* out = &s->outlet;
    int64_t n_in = 0, n_fo = 0, n_ex = 0;
    int64_t* ids_in = (int64_t*)malloc(sizeof(int64_t) * (size_t)(in->n > 0 ? in->n : 1));
    int64_t* ids_fo = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n > 0 ? s->n : 1));
    int64_t* ids_ex = (int64_t*)malloc(sizeof(int64_t) * (size_t)(out->n > 0 ? out->n : 1));
    for (int64_t i = 0; i < in->n; ++i)
        if (in->x[i * D + a] > s->d.inlet_enter) ids_in[n_in++] = i;
    for (int64_t i = 0; i < s->n; ++i)
        if (s->x[i * D + a] > s->d.outlet_enter) ids_fo[n_fo++] = i;
    for (int64_t i = 0; i < out->n; ++i)
        if (out->x[i * D + a] > s->d.outlet_exit) ids_ex[n_ex++] = i;
    /* push inlet copies onto the fluid (deepcopy: dv and drho are zero) */
    fluid_reserve(s, s->n + n_in);
    for (int64_t e = 0; e < n_in; ++e) {
        const int64_t i = ids_in[e], w = s->n++;
        for (int k = 0; k < D; ++k) {
            s->x[w * D + k] = in->x[i * D + k];
            s->v[w * D + k] = in->v[i * D + k];
            s->dv[w * D + k] = 0.0;
        }
        s->rho[w] = in->rho[i]; s->drho[w] = 0.0; s->p[w] = in->p[i]; s->c[w] = in->c[i];
        s->m[w] = in->m[i]; s->gap[w] = in->gap[i];
        s->t[w] = 0.0; s->dtt[w] = 0.0; s->kappa[w] = 0.0; s->cp[w] = 0.0; s->sw[w] = 0.0; s->ss[w] = 0.0;
        s->id[w] = s->next_id++;
    }
    /* push fluid copies onto the outlet */
    for (int64_t e = 0; e < n_fo; ++e) {
        const int64_t i = ids_fo[e];
        buf_push(out, D, s->x + i * D, s->v + i * D, s->rho[i], s->p[i], s->c[i], s->m[i], s->gap[i]);
    }
    /* deleteat!(fluid, ids_fo), stable */
    if (n_fo > 0) {
        int64_t w = 0, e = 0;
        for (int64_t i = 0; i < s->n; ++i) {
            if (e < n_fo && ids_fo[e] == i) { ++e; continue; }
            if (w != i) {
                for (int k = 0; k < D; ++k) {
                    s->x[w * D + k] = s->x[i * D + k];
                    s->v[w * D + k] = s->v[i * D + k];
                    s->dv[w * D + k] = s->dv[i * D + k];
                }
                s->rho[w] = s->rho[i]; s->drho[w] = s->drho[i]; s->p[w] = s->p[i]; s->c[w] = s->c[i];
                s->m[w] = s->m[i]; s->gap[w] = s->gap[i]; s->id[w] = s->id[i];
            }
            ++w;
        }
        s->n = w;
    }
    /* deleteat!(outlet, ids_ex), stable */
    if (n_ex > 0) {
        int64_t w = 0, e = 0;
        for (int64_t i = 0; i < out->n; ++i) {
            if (e < n_ex && ids_ex[e] == i) { ++e; continue; }
            if (w != i) {
                for (int k = 0; k < D; ++k) {
                    out->x[w * D + k] = out->x[i * D + k];
                    out->v[w * D + k] = out->v[i * D + k];
                }
                out->rho[w] = out->rho[i]; out->p[w] = out->p[i]; out->c[w] = out->c[i];
                out->m[w] = out->m[i]; out->gap[w] = out->gap[i]; out->id[w] = out->id[i];
            }
            ++w;
        }
        out->n = w;
    }
    /* inlet_into_calculation_domain_modify: move back along the axis */
    for (int64_t e = 0; e < n_in; ++e) in->x[ids_in[e] * D + a] -= s->d.inlet_shift;
    free(ids_in);
    free(ids_fo);
    free(ids_ex);
    /* user_defined_function(t = step * dt) */
    if (s->d.profile != ESPH_PROFILE_NONE) {
        const double t = (double)s->step * s->d.dt;
        for (int64_t i = 0; i < in->n; ++i) in->v[i * D + a] = profile_u(s, in->x + i * D, t);
        for (int64_t i = 0; i < out->n; ++i) {
            if (s->d.outlet_profile) out->v[i * D + a] = profile_u(s, out->x + i * D, t);
            for (int k = 0; k < D; ++k)
                if (k != a) out->v[i * D + k] = 0.0;
        }
    }
}

static void step_once(eo_scene* s, int nthreads) {
    const int D = s->d.dim;
    const double h = s->k.h;
    const double dt = s->d.dt;
    const int solver = s->d.solver;
    free_pairs(s);
    /* neighbours (eachStep! :21-35 of FluidWallSolve.jl and the equivalents) */
    make_records(s, 0, nthreads);
    if (s->nw > 0) make_records(s, 1, nthreads);
    const int io = solver == ESPH_SOLVER_FLUID_INLET_OUTLET_WALL;
    if (io) {
        make_buffer_records(s, &s->inlet, &s->fi, nthreads);
        make_buffer_records(s, &s->outlet, &s->fo, nthreads);
    }

    /* continuity! over FF pairs, Continuity.jl:8-17 */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t q = 0; q < s->nff; ++q) {
        int64_t i = s->ffi[q], j = s->ffj[q];
        double drho = 0.0;
        for (int k = 0; k < D; ++k) drho = drho + s->ffvv[q * D + k] * s->ffg[q * D + k];
        atomic_add(&s->drho[i], drho * s->m[j]);
        atomic_add(&s->drho[j], drho * s->m[i]);
    }
    if (io) {
        buffer_continuity(s, &s->inlet, &s->fi);
        buffer_continuity(s, &s->outlet, &s->fo);
    }
    /* updateDensity! + updatePressure! */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < s->n; ++i) update_density_pressure(s, i);

    /* FF momentum (and conduction) */
    const int thermal = s->thermal;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t q = 0; q < s->nff; ++q) {
        int64_t i = s->ffi[q], j = s->ffj[q];
        const double* g = s->ffg + q * D;
        const double* vv = s->ffvv + q * D;
        double r = s->ffr[q], W = s->ffW[q], dW = s->ffdW[q];
        /* pressureForce! FF, PressureForce.jl:8-21 */
        double p_rho_2 = s->p[i] / (s->rho[i] * s->rho[i]) + s->p[j] / (s->rho[j] * s->rho[j]);
        p_rho_2 += fabs(p_rho_2) * 0.01 * W / eo_kernel_value(&s->k, (s->gap[i] + s->gap[j]) / 2.0);
        for (int k = 0; k < D; ++k) {
            atomic_add(&s->dv[i * D + k], -s->m[j] * p_rho_2 * g[k]);
            atomic_add(&s->dv[j * D + k], s->m[i] * p_rho_2 * g[k]);
        }
        /* viscosityForce! FF, ViscosityForce.jl:8-25 */
        double sum_rho = s->rho[i] + s->rho[j];
        double mu_i = s->nu_0 * s->rho[i];
        double mu_j = s->nu_0 * s->rho[j];
        double vf = 4.0 * (mu_i + mu_j) * r * dW / (sum_rho * sum_rho) / (r * r + 0.01 * (h * h));
        for (int k = 0; k < D; ++k) {
            atomic_add(&s->dv[i * D + k], s->m[j] * vf * vv[k]);
            atomic_add(&s->dv[j * D + k], -s->m[i] * vf * vv[k]);
        }
        /* thermalConduction! FF, ThermalConduction.jl:8-22 */
        if (thermal) {
            double keq = 2.0 * s->kappa[i] * s->kappa[j] / (s->kappa[i] + s->kappa[j]);
            double dq = 2.0 * keq * r * dW / s->rho[i] / s->rho[j] / (r * r + 0.01 * (h * h)) * (s->t[i] - s->t[j]);
            atomic_add(&s->dtt[i], dq * s->m[j] / s->cp[i]);
            atomic_add(&s->dtt[j], -dq * s->m[i] / s->cp[j]);
        }
    }

    /* FW interactions.  Which actions run depends on the solver:
       FluidWallSolve.jl:68-79 wallForce! + viscosityForce!(wall);
       FluidVelocitySolve.jl:60-82 and FluidThermostaticWallSolve.jl:74-103,
       FluidCompulsiveThermostaticWallSolve.jl:87-129 add pressureForce!(wall);
       thermostatic walls of the thermal solvers add thermalConduction!. */
    /* the inlet/outlet solver has no wall pressure term (FluidInletOutletWallSolve.jl:121-122) */
    const int wall_pressure = solver != ESPH_SOLVER_FLUID_WALL && !io;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t q = 0; q < s->nfw; ++q) {
        int64_t i = s->fwi[q], j = s->fwj[q];
        const double* g = s->fwg + q * D;
        const double* vv = s->fwvv + q * D;
        const double* rv = s->fwrv + q * D;
        const double* nj = s->wn + j * D;
        double r = s->fwr[q], W = s->fwW[q], dW = s->fwdW[q];
        /* wallForce!, WallForce.jl:8-30 */
        {
            double dotrn = 0.0, vn = 0.0;
            for (int k = 0; k < D; ++k) dotrn = dotrn + rv[k] * nj[k];
            double psi = fabs(dotrn);
            double a = r * r - psi * psi;
            double xi = sqrt(a > 0.0 ? a : 0.0);
            double qq = psi / s->wgap[j];
            if (!(qq > 1.0 || xi > s->wgap[j])) {
                double p_xi = fabs((1.0 + cos(3.14159265358979323846 * xi / s->wgap[j])) / 2.0);
                for (int k = 0; k < D; ++k) vn = vn + vv[k] * nj[k];
                double beta = vn > 0.0 ? 0.0 : 1.0;
                double r_psi = (0.01 * (s->c[i] * s->c[i]) + beta * s->c[i] * fabs(vn)) / h * fabs(1.0 - qq) / sqrt(qq);
                double wf = r_psi * p_xi;
                for (int k = 0; k < D; ++k) atomic_add(&s->dv[i * D + k], wf * nj[k]);
            }
        }
        /* pressureForce! FW, PressureForce.jl:23-35 */
        if (wall_pressure) {
            double p_rho_2 = s->p[i] / (s->rho[i] * s->rho[i]) + s->d.p_0 / (s->d.rho_0 * s->d.rho_0);
            p_rho_2 += fabs(p_rho_2) * 0.01 * W / eo_kernel_value(&s->k, s->gap[i]);
            for (int k = 0; k < D; ++k) atomic_add(&s->dv[i * D + k], -s->m[i] * p_rho_2 * g[k]);
        }
        /* viscosityForce! FW, ViscosityForce.jl:27-46 */
        {
            double sum_rho = 2.0 * s->rho[i];
            double mu_i = s->nu_0 * s->rho[i];
            double vf = 8.0 * mu_i * r * dW / (sum_rho * sum_rho) / (r * r + 0.01 * (h * h));
            double vn = 0.0;
            for (int k = 0; k < D; ++k) vn = vn + vv[k] * nj[k];
            for (int k = 0; k < D; ++k) atomic_add(&s->dv[i * D + k], s->m[i] * vf * (vv[k] - vn * nj[k]));
        }
        /* thermalConduction! FW thermostat, ThermalConduction.jl:24-36 */
        if (thermal && s->wkind[j] == ESPH_WALL_THERMOSTATIC) {
            double dq = 1.0 / s->d.cp * (s->kappa[i] + s->wkappa[j]) * r * dW / s->rho[i] / s->d.rho_0 /
                        (r * r + 0.01 * (h * h)) * (s->t[i] - s->wt[j]);
            atomic_add(&s->dtt[i], dq * s->m[i]);
        }
    }

    if (io) {
        buffer_forces(s, &s->inlet, &s->fi);
        buffer_forces(s, &s->outlet, &s->fo);
    }

    /* integration: updateTemperature!, updateVelocity!(…, body_force),
       updatePosition! (UpdateTemperature.jl:8-12, UpdateVelocity.jl:8-23,
       UpdatePosition.jl:8-14; FluidWallSolve.jl:82-85,
       FluidCompulsiveThermostaticWallSolve.jl:134-138) */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < s->n; ++i) {
        double bf[3];
        if (thermal) {
            s->t[i] += dt * s->dtt[i];
            s->dtt[i] = 0.0;
            for (int k = 0; k < D; ++k) bf[k] = -s->d.body_force[k] * s->d.beta * (s->t[i] - s->d.t_0);
        } else {
            for (int k = 0; k < D; ++k) bf[k] = s->d.body_force[k];
        }
        for (int k = 0; k < D; ++k) s->v[i * D + k] += bf[k] * dt;
        for (int k = 0; k < D; ++k) s->v[i * D + k] += s->dv[i * D + k] * dt;
        for (int k = 0; k < D; ++k) {
            s->x[i * D + k] += s->v[i * D + k] * dt - s->dv[i * D + k] * (dt * dt) / 2.0;
            s->dv[i * D + k] = 0.0;
        }
    }

    if (io) {
        /* buffers: dv = 0, then updatePosition! (FluidInletOutletWallSolve.jl:170-186) */
        eo_buf* bs[2] = {&s->inlet, &s->outlet};
        for (int q = 0; q < 2; ++q)
            for (int64_t i = 0; i < bs[q]->n; ++i)
                for (int k = 0; k < D; ++k)
                    bs[q]->x[i * D + k] += bs[q]->v[i * D + k] * dt - 0.0 * (dt * dt) / 2.0;
    }
    s->step += 1;
    /* reconstructScalar!(fluid, :rho_, FF pairs, kernel), ReconstructValue.jl:22-57,
       gated by isDensityReinitializedStep (FixedTimeIntervalDiscretization.jl:18-23) */
    if (s->d.density_reinitialized_interval > 0 && s->step % s->d.density_reinitialized_interval == 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
        for (int64_t q = 0; q < s->nff; ++q) {
            int64_t i = s->ffi[q], j = s->ffj[q];
            double W = s->ffW[q];
            atomic_add(&s->sw[i], W * s->m[j] / s->rho[j]);
            atomic_add(&s->ss[i], W * s->m[j] / s->rho[j] * s->rho[j]);
            atomic_add(&s->sw[j], W * s->m[i] / s->rho[i]);
            atomic_add(&s->ss[j], W * s->m[i] / s->rho[i] * s->rho[i]);
        }
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
        for (int64_t i = 0; i < s->n; ++i) {
            s->sw[i] += s->k.kernel_0 * s->m[i] / s->rho[i];
            s->ss[i] += s->k.kernel_0 * s->m[i] / s->rho[i] * s->rho[i];
            s->rho[i] = s->ss[i] / s->sw[i];
            s->sw[i] = 0.0;
            s->ss[i] = 0.0;
        }
    }
    if (io) migrate(s);
    /* check + deleteat! (FluidWallSolve.jl:73-81): stable removal */
    if (!io && s->d.check_enabled && s->d.check_interval > 0 && s->step % s->d.check_interval == 0) {
        int64_t w = 0;
        for (int64_t i = 0; i < s->n; ++i) {
            int out = 0;
            for (int k = 0; k < D; ++k)
                if (s->x[i * D + k] > s->d.check_max[k] || s->x[i * D + k] < s->d.check_min[k]) out = 1;
            for (int q = 0; q < s->d.check_nplanes && q < 8; ++q) {
                const double* pl = s->d.check_planes[q];
                double dot = pl[0] * s->x[i * D] + pl[1] * s->x[i * D + 1];
                if (D == 3) dot = dot + pl[2] * s->x[i * D + 2];
                if (dot > pl[3]) out = 1;
            }
            if (out) continue;
            if (w != i) {
                for (int k = 0; k < D; ++k) {
                    s->x[w * D + k] = s->x[i * D + k];
                    s->v[w * D + k] = s->v[i * D + k];
                    s->dv[w * D + k] = s->dv[i * D + k];
                }
                s->rho[w] = s->rho[i]; s->drho[w] = s->drho[i]; s->p[w] = s->p[i]; s->c[w] = s->c[i];
                s->m[w] = s->m[i]; s->gap[w] = s->gap[i]; s->t[w] = s->t[i]; s->dtt[w] = s->dtt[i];
                s->kappa[w] = s->kappa[i]; s->cp[w] = s->cp[i]; s->id[w] = s->id[i];
            }
            w++;
        }
        s->n = w;
    }
}

int eo_scene_step(eo_scene* s, int64_t nsteps, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    for (int64_t t = 0; t < nsteps; ++t) step_once(s, nthreads);
    free_pairs(s);
    return 0;
}

int64_t eo_scene_fluid_count(const eo_scene* s) { return s->n; }
int64_t eo_scene_step_count(const eo_scene* s) { return s->step; }

int eo_scene_get_fluid(const eo_scene* s, int32_t field, double* out) {
    const int D = s->d.dim;
    const int64_t n = s->n;
    const double* src = NULL;
    int64_t w = n;
    switch (field) {
        case ESPH_FIELD_X: src = s->x; w = n * D; break;
        case ESPH_FIELD_V: src = s->v; w = n * D; break;
        case ESPH_FIELD_RHO: src = s->rho; break;
        case ESPH_FIELD_P: src = s->p; break;
        case ESPH_FIELD_C: src = s->c; break;
        case ESPH_FIELD_MASS: src = s->m; break;
        case ESPH_FIELD_GAP: src = s->gap; break;
        case ESPH_FIELD_T: src = s->t; break;
        case ESPH_FIELD_KAPPA: src = s->kappa; break;
        case ESPH_FIELD_CP: src = s->cp; break;
        case ESPH_FIELD_ID:
            for (int64_t i = 0; i < n; ++i) out[i] = (double)s->id[i];
            return 0;
        default: return -1;
    }
    memcpy(out, src, sizeof(double) * (size_t)w);
    return 0;
}

int64_t eo_scene_group_count(const eo_scene* s, int32_t group) {
    if (group == ESPH_GROUP_INLET) return s->inlet.n;
    if (group == ESPH_GROUP_OUTLET) return s->outlet.n;
    return s->n;
}

int eo_scene_get_group(const eo_scene* s, int32_t group, int32_t field, double* out) {
    if (group == ESPH_GROUP_FLUID) return eo_scene_get_fluid(s, field, out);
    const eo_buf* b = group == ESPH_GROUP_INLET ? &s->inlet : &s->outlet;
    const int D = s->d.dim;
    const double* src = NULL;
    int64_t w = b->n;
    switch (field) {
        case ESPH_FIELD_X: src = b->x; w = b->n * D; break;
        case ESPH_FIELD_V: src = b->v; w = b->n * D; break;
        case ESPH_FIELD_RHO: src = b->rho; break;
        case ESPH_FIELD_P: src = b->p; break;
        case ESPH_FIELD_C: src = b->c; break;
        case ESPH_FIELD_MASS: src = b->m; break;
        case ESPH_FIELD_GAP: src = b->gap; break;
        case ESPH_FIELD_ID:
            for (int64_t i = 0; i < b->n; ++i) out[i] = (double)b->id[i];
            return 0;
        default: return -1;
    }
    if (w > 0) memcpy(out, src, sizeof(double) * (size_t)w);
    return 0;
}

int64_t eo_scene_neighbours(eo_scene* s, int32_t pairset, int32_t* out_pairs, int64_t cap) {
    if (pairset == ESPH_PAIRS_FLUID_FLUID)
        return eo_neighbor_list(s->d.dim, s->k.influence_radius, s->n, s->x, 0, NULL, 0, 1, out_pairs, cap);
    if (s->nw == 0) return 0;
    return eo_neighbor_list(s->d.dim, s->k.influence_radius, s->n, s->x, s->nw, s->wx, 0, 1, out_pairs, cap);
}

/* kernelValueInterpolation (DataIO/VTPToDataFrame.jl:77-142) over the fluid
   particles: brute force over every (probe, particle) pair in index order. */
int eo_scene_probe(const eo_scene* s, int32_t field, int64_t np, const double* pts, double none_value,
                   double* out) {
    const int D = s->d.dim;
    const double* f = NULL;
    int w = 1;
    switch (field) {
        case ESPH_FIELD_V: f = s->v; w = D; break;
        case ESPH_FIELD_RHO: f = s->rho; break;
        case ESPH_FIELD_P: f = s->p; break;
        case ESPH_FIELD_C: f = s->c; break;
        case ESPH_FIELD_T: f = s->thermal ? s->t : NULL; break;
        default: return -1;
    }
    if (!f) return -1;
    const double R = s->k.influence_radius, R2 = R * R;
    for (int64_t q = 0; q < np; ++q) {
        double sw = 0.0, sf[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < s->n; ++j) {
            const double d2 = dist2(D, pts + q * D, s->x + j * D);
            if (!(d2 <= R2)) continue;
            const double W = eo_kernel_value(&s->k, sqrt(d2));
            const double wj = W * s->m[j] / s->rho[j];
            sw += wj;
            for (int k = 0; k < w; ++k) sf[k] += wj * f[j * w + k];
        }
        for (int k = 0; k < w; ++k) out[q * w + k] = sw > 0.0 ? sf[k] / sw : none_value;
    }
    return 0;
}
