// esph_kernels.hip — gfx950 kernels of the per-step SPH hot path.
//
// One step of eachStep! (Solve/FluidWallSolve.jl:8-83 and the other
// fixed-count solvers) is, per scene:
//   k_continuity   continuity! + updateDensity! + updatePressure!
//   k_forces       pressure/viscosity/wall/thermal pair terms + integration
//   k_shepard      reconstructScalar!(…, :rho_, …) on re-initialisation steps
//   k_sort_*       counting sort of the new positions into cells (+ check)
// All pair loops are pull-style: each particle gathers over its neighbour
// cells, so no atomics touch the physics and sums are deterministic for a
// given particle order.  Every kernel walks a tile table (scene, first
// particle), so one launch serves one scene or a batch of scenes.
#include "esph_device.hpp"
#include "esph_launch.hpp"

#include <type_traits>

namespace esph {

// Workgroups go to the eight XCDs round-robin (block b runs on XCD b % 8).
// With one tile per block, xcd_tile gives each XCD one contiguous range of
// tiles, walked in order, so each XCD's L2 holds one compact, slowly moving
// region of the cell-sorted particles instead of every eighth tile.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    constexpr int X = 8;
    const int per = n / X, rem = n % X;
    const int x = b % X, k = b / X;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}
__device__ __forceinline__ int tile_index(int b, int n) {
    return (int)gridDim.x == n ? xcd_tile(b, n) : b;
}

template <typename T>
__device__ __forceinline__ T ld(const void* p, int i) {
    return static_cast<const T*>(p)[i];
}

template <int D>
__device__ __forceinline__ void load_x(const double* __restrict__ x, int j, double* o) {
    if (D == 2) {
        double2 q = reinterpret_cast<const double2*>(x)[j];
        o[0] = q.x;
        o[1] = q.y;
    } else {
        o[0] = x[3 * j];
        o[1] = x[3 * j + 1];
        o[2] = x[3 * j + 2];
    }
}

// Candidate rows of a point xi in cell c: the grid's cells are at least R /
// stencil wide, so neighbours lie within `stencil` cells on every axis.  For
// each (dy, dz) the cells of the row are x-contiguous in memory, so a row is
// one index range [j0, j1).  Rows and row ends whose nearest point is
// farther than R from xi (judged from xi's offset inside its cell, with a
// slack that keeps the test conservative under rounding and clamping) are
// skipped.  Calls row(j0, j1) for each remaining row.
template <int D, typename F>
__device__ __forceinline__ void for_each_row(const SceneDev& s, const int32_t* __restrict__ cstart,
                                             const double* xi, const int c[3], F&& row) {
    const int m = s.stencil;
    const double cs = s.cs;
    const double slack = 1e-7 * cs;
    const double lim = s.R2 * (1.0 + 1e-6);
    double p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) p[k] = k < D ? xi[k] - (s.lo[k] + (double)c[k] * cs) : 0.0;
    // distance from xi to the slab of the cell at offset o along one axis
    auto dmin = [&](int o, double pk) {
        double d = o == 0 ? 0.0 : (o > 0 ? (double)(o - 1) * cs + (cs - pk) : (double)(-o - 1) * cs + pk);
        d -= slack;
        return d > 0.0 ? d : 0.0;
    };
    const int mz = D == 3 ? m : 0;
    for (int dz = -mz; dz <= mz; ++dz) {
        const int cz = c[2] + dz;
        if (cz < 0 || cz >= s.nc[2]) continue;
        const double rz = D == 3 ? dmin(dz, p[2]) : 0.0;
        for (int dy = -m; dy <= m; ++dy) {
            const int cy = c[1] + dy;
            if (cy < 0 || cy >= s.nc[1]) continue;
            const double ry = dmin(dy, p[1]);
            const double r2 = ry * ry + rz * rz;
            if (r2 > lim) continue;
            int xl = -m, xh = m;
            while (xl < 0) {
                const double d = dmin(xl, p[0]);
                if (d * d + r2 <= lim) break;
                ++xl;
            }
            while (xh > 0) {
                const double d = dmin(xh, p[0]);
                if (d * d + r2 <= lim) break;
                --xh;
            }
            const int cx0 = c[0] + xl > 0 ? c[0] + xl : 0;
            const int cx1 = c[0] + xh < s.nc[0] - 1 ? c[0] + xh : s.nc[0] - 1;
            if (cx0 > cx1) continue;
            row(cstart[cell_linear(s, cx0, cy, cz)], cstart[cell_linear(s, cx1, cy, cz) + 1]);
        }
    }
}

// Iterates the particles of the candidate rows of xi whose distance
// satisfies the neighbour rule (FP64, exact).  The candidate loop issues the
// loads of four candidates before testing them.
template <int D, typename F>
__device__ __forceinline__ void for_each_neighbour(const SceneDev& s, const int32_t* __restrict__ cstart,
                                                   const double* __restrict__ xall, const double* xi,
                                                   const int c[3], int self, F&& f) {
    const double R2 = s.R2;
    for_each_row<D>(s, cstart, xi, c, [&](int j0, int j1) {
        for (int j = j0; j < j1; j += 4) {
            double xj[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) load_x<D>(xall, j + u < j1 ? j + u : j1 - 1, xj[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j + u;
                double rv[3];
                const double d2 = dist2_exact<D>(xi, xj[u], rv);
                if (jj < j1 && jj != self && d2 <= R2) f(jj, d2, rv);
            }
        }
    });
}

// Neighbour lists are stored in chunks of four: entries 4q..4q+3 of particle
// i are the int4 at [q * cap + i], so a wavefront reads 64 consecutive int4
// (1 KiB) per chunk row and each lane writes whole 16-byte pieces.
constexpr int kChunk = 4;
__device__ __forceinline__ size_t list_at(size_t cap, int k, int i) {
    return (((size_t)(k >> 2) * cap + (size_t)i) << 2) + (size_t)(k & 3);
}

// Candidates of particle i over the adjacent cells, tested first in float on
// the grid-relative positions xf: a candidate whose float distance is clearly
// inside or clearly outside R is decided there; only those within the
// rounding band tau of R take the FP64 rule of for_each_neighbour.  tau bounds
// the float error: 2^-21 (|xr_i|_inf + 2 R), so the decisions equal the FP64
// rule's.  Calls f(j) for every neighbour, in cell order.
template <int D, typename F>
__device__ __forceinline__ void for_each_candidate_f32(const SceneDev& s, const int32_t* __restrict__ cstart,
                                                       const void* __restrict__ xfall,
                                                       const double* __restrict__ xall, const double* xi,
                                                       const int c[3], int self, F&& f) {
    float xr[3] = {(float)(xi[0] - s.lo[0]), (float)(xi[1] - s.lo[1]), D == 3 ? (float)(xi[2] - s.lo[2]) : 0.0f};
    const double amax = fmax(fabs((double)xr[0]), fmax(fabs((double)xr[1]), fabs((double)xr[2])));
    const double tau = ldexp(amax + 2.0 * s.R, -21);
    const double rin = s.R - tau, rout = s.R + tau;
    const float lo2 = rin > 0.0 ? (float)(rin * rin) * (1.0f - 0x1p-22f) : -1.0f;
    const float hi2 = (float)(rout * rout) * (1.0f + 0x1p-22f);
    const double R2 = s.R2;
    for_each_row<D>(s, cstart, xi, c, [&](int j0, int j1) {
        for (int j = j0; j < j1; j += 4) {
            float d2f[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j + u < j1 ? j + u : j1 - 1;
                float dx, dy_, dz_ = 0.0f;
                if (D == 3) {
                    const float4 q = static_cast<const float4*>(xfall)[jj];
                    dx = xr[0] - q.x;
                    dy_ = xr[1] - q.y;
                    dz_ = xr[2] - q.z;
                } else {
                    const float2 q = static_cast<const float2*>(xfall)[jj];
                    dx = xr[0] - q.x;
                    dy_ = xr[1] - q.y;
                }
                d2f[u] = dx * dx + dy_ * dy_ + dz_ * dz_;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j + u;
                if (jj >= j1 || jj == self || !(d2f[u] <= hi2)) continue;
                bool in = d2f[u] < lo2;
                if (!in) {  // rounding band: the FP64 rule decides
                    double xj[3];
                    load_x<D>(xall, jj, xj);
                    double rv[3];
                    in = dist2_exact<D>(xi, xj, rv) <= R2;
                }
                if (in) f(jj);
            }
        }
    });
}

// Records the neighbours of particle i in the chunked list `list`; returns
// the count.  Entries fill a 4-entry register buffer flushed as one int4.
// Entries beyond `width` (a multiple of 4) are dropped and the particle
// later scans its cells.
template <int D>
__device__ __forceinline__ int build_list(const SceneDev& s, const int32_t* __restrict__ cstart,
                                          const void* __restrict__ xfall, const double* __restrict__ xall,
                                          const double* xi, const int c[3], int i, int self,
                                          int32_t* __restrict__ list, int width) {
    int cnt = 0;
    int4 buf = make_int4(0, 0, 0, 0);
    const size_t cap = (size_t)s.cap;
    for_each_candidate_f32<D>(s, cstart, xfall, xall, xi, c, self, [&](int j) {
        const int r = cnt & 3;
        buf.x = r == 0 ? j : buf.x;
        buf.y = r == 1 ? j : buf.y;
        buf.z = r == 2 ? j : buf.z;
        buf.w = r == 3 ? j : buf.w;
        if (r == 3 && cnt < width) *reinterpret_cast<int4*>(list + list_at(cap, cnt & ~3, i)) = buf;
        ++cnt;
    });
    if ((cnt & 3) != 0 && cnt < width) *reinterpret_cast<int4*>(list + list_at(cap, cnt & ~3, i)) = buf;
    return cnt;
}

// Iterates the recorded neighbours of particle i with the same (j, d2, rv)
// arguments as for_each_neighbour; the next entry's index and position are
// loaded while the current one is processed.
template <int D, typename F>
__device__ __forceinline__ void for_each_listed(const SceneDev& s, const double* __restrict__ xall,
                                                const double* xi, int i, const int32_t* __restrict__ list,
                                                int cnt, F&& f) {
    if (cnt <= 0) return;
    const size_t cap = (size_t)s.cap;
    int4 ch = *reinterpret_cast<const int4*>(list + list_at(cap, 0, i));
    int jn = ch.x;
    double xn[3];
    load_x<D>(xall, jn, xn);
    for (int k = 0; k < cnt; ++k) {
        const int j = jn;
        double xj[3] = {xn[0], xn[1], D == 3 ? xn[2] : 0.0};
        if (k + 1 < cnt) {
            const int r = (k + 1) & 3;
            if (r == 0) ch = *reinterpret_cast<const int4*>(list + list_at(cap, k + 1, i));
            jn = r == 0 ? ch.x : r == 1 ? ch.y : r == 2 ? ch.z : ch.w;
            load_x<D>(xall, jn, xn);
        }
        double rv[3];
        const double d2 = dist2_exact<D>(xi, xj, rv);
        f(j, d2, rv);
    }
}

// Pairs of particle i: from the step's list, or (SCAN, for particles whose
// list overflowed) by scanning the cells.  Both visit the identical pair set.
template <int D, bool SCAN, typename F>
__device__ __forceinline__ void for_pairs(const SceneDev& s, const int32_t* __restrict__ cstart,
                                          const double* __restrict__ xall, const double* xi, int i, int self,
                                          const int32_t* __restrict__ list, int cnt, F&& f) {
    if (SCAN) {
        int c[3];
        cell_of<D>(s, xi, c);
        for_each_neighbour<D>(s, cstart, xall, xi, c, self, f);
    } else {
        for_each_listed<D>(s, xall, xi, i, list, cnt, f);
    }
}

// The SCAN instances walk the step's overflow list with a small fixed grid:
// kOvfTiles tiles of kBlock entries per scene, striding over the list.
// Calls f(i) for the particle(s) of this thread: the tile slot (main
// kernels) or, in the SCAN instances, every kOvfTiles*kBlock-th entry of the
// step's overflow list.
template <bool SCAN, typename F>
__device__ __forceinline__ void for_work(const SceneDev& s, const SceneDyn& dyn, const Tile& tile, F&& f) {
    const int q0 = tile.begin + (int)threadIdx.x;
    if (SCAN) {
        for (int q = q0; q < dyn.novf; q += kOvfTiles * kBlock) f(s.ovf[q]);
    } else if (q0 < dyn.n_live) {
        f(q0);
    }
}

__device__ __forceinline__ bool overflowed(const SceneDev& s, int i) {
    return s.nnb[i] > s.kff || (s.nw > 0 && s.nnbw[i] > s.kfw);
}

// Unpacked neighbour record (see FluidSet).
template <typename T>
struct Rec {
    T v[3];
    T m, rho, pr2, gap, Tt, kap;
};

template <int D, typename T>
__device__ __forceinline__ void load_A(const void* A, int j, T* v, T& m, T& gap) {
    using V = typename Vec4<T>::type;
    const V a = static_cast<const V*>(A)[j];
    v[0] = a.x;
    v[1] = a.y;
    if (D == 3) {
        v[2] = a.z;
        m = a.w;
    } else {
        m = a.z;
        gap = a.w;
    }
}

template <int D, typename T, bool THERMAL>
__device__ __forceinline__ void load_rec(const FluidSet& S, int j, Rec<T>& r) {
    using V = typename Vec4<T>::type;
    load_A<D, T>(S.A, j, r.v, r.m, r.gap);
    const V b = static_cast<const V*>(S.B)[j];
    r.rho = b.x;
    r.pr2 = b.y;
    if (D == 3) {
        r.gap = b.z;
        r.Tt = b.w;
        if (THERMAL) r.kap = ld<T>(S.kap, j);
    } else {
        r.Tt = b.z;
        r.kap = b.w;
    }
}

// Record fields written by the sort (A, B.zw, kap) for particle p.
template <typename T>
__device__ __forceinline__ void store_sorted_rec(const FluidSet& O, int D, int p, const double* v, double m,
                                                 double gap, double Tt, double kap, bool thermal) {
    T* A = static_cast<T*>(O.A) + 4 * (size_t)p;
    T* B = static_cast<T*>(O.B) + 4 * (size_t)p;
    if (D == 3) {
        A[0] = (T)v[0]; A[1] = (T)v[1]; A[2] = (T)v[2]; A[3] = (T)m;
        B[2] = (T)gap;
        B[3] = (T)Tt;
        if (thermal) static_cast<T*>(O.kap)[p] = (T)kap;
    } else {
        A[0] = (T)v[0]; A[1] = (T)v[1]; A[2] = (T)m; A[3] = (T)gap;
        B[2] = (T)Tt;
        B[3] = (T)kap;
    }
}

// ---------------------------------------------------------------------------
// continuity! (Interaction/Continuity.jl:8-17), pull form:
//   drho_i = sum_j m_j (v_i - v_j) . gradW_ij,  gradW_ij = W'(r) r_ij / r
// then updateDensity! (SelfAction/UpdateDensity.jl:8-12) and updatePressure!
// (SelfAction/UpdatePressure.jl:8-16).  Writes rho, p, c and the record's
// (rho, p/rho^2).  This pass also records the step's neighbour lists
// (positions of the step start), which the force and Shepard passes reuse.
// ---------------------------------------------------------------------------
template <int D, int K, typename T, bool SCAN>
__global__ __launch_bounds__(kBlock) void k_continuity(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles, int ntiles) {
    using KQ = KernelQ<T, K>;
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        for_work<SCAN>(s, dyn, tile, [&](int i) {
        const FluidSet& S = s.set[dyn.parity];
        if (S.kind && S.kind[i] != KIND_FLUID) {
            // inlet/outlet buffers keep rho and p (drho is reset,
            // FluidInletOutletWallSolve.jl:91-99) and need no lists
            s.nnb[i] = 0;
            if (s.nw > 0) s.nnbw[i] = 0;
            return;
        }
        const T inv_h = (T)s.inv_h, kA = (T)s.kA, kB = (T)s.kB;
        double xi[3];
        load_x<D>(S.x, i, xi);
        T vi[3], m_i, gap_i;
        load_A<D, T>(S.A, i, vi, m_i, gap_i);
        int cnt = 0;
        if (!SCAN) {
            int c[3];
            cell_of<D>(s, xi, c);
            cnt = build_list<D>(s, s.cell_start, S.xf, S.x, xi, c, i, i, s.nbr, s.kff);
            s.nnb[i] = cnt;
            bool over = cnt > s.kff;
            if (s.nw > 0) {
                const int cw = build_list<D>(s, s.wcell_start, s.wxf, s.wx, xi, c, i, -1, s.wnbr, s.kfw);
                s.nnbw[i] = cw;
                over = over || cw > s.kfw;
            }
            if (over) {
                // finished by the SCAN instance of this kernel
                s.ovf[atomicAdd(&s.dyn->novf, 1)] = i;
                atomicAdd(&s.dyn->overflow, 1);
                return;
            }
        }
        T drho = T(0);
        for_pairs<D, SCAN>(s, s.cell_start, S.x, xi, i, i, s.nbr, cnt, [&](int j, double d2, const double* rv) {
            T vj[3], m_j, g_j;
            load_A<D, T>(S.A, j, vj, m_j, g_j);
            const T r = psqrt((T)d2);
            T W, dW;
            KQ::wdw(r * inv_h, kA, kB, W, dW);
            // normalize(0) is NaN and the reference zeroes it (Neighbour.jl:54-59)
            const T sg = d2 == 0.0 ? T(0) : dW * prcp(r);
            T dot = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) dot += (vi[k] - vj[k]) * (T)rv[k];
            drho += m_j * sg * dot;
        });
        {
#pragma clang fp contract(off)
            double rho = S.rho[i];
            rho += (double)drho * s.dt;
            const double rr = rho / s.rho0;
            const double p = s.b * (ipow(rr, s.gamma) - 1.0) + s.p0;
            const double cc = s.c0 * sqrt(ipow(rr, s.gamma - 1));
            S.rho[i] = rho;
            S.p[i] = p;
            S.c[i] = cc;
            T* B = static_cast<T*>(S.B) + 4 * (size_t)i;
            B[0] = (T)rho;
            B[1] = (T)(p / (rho * rho));
        }
        });
    }
}

// ---------------------------------------------------------------------------
// Momentum (and heat) pass + integration, pull form over the step's lists.
//  FF: pressureForce! (PressureForce.jl:8-21), viscosityForce!
//      (ViscosityForce.jl:8-25), thermalConduction! (ThermalConduction.jl:8-22)
//  FW: wallForce! (WallForce.jl:8-30), pressureForce! wall
//      (PressureForce.jl:23-35, solvers other than FluidWallSolve),
//      viscosityForce! wall (ViscosityForce.jl:27-46), thermalConduction!
//      thermostat (ThermalConduction.jl:24-36)
//  then updateTemperature!, updateVelocity!(…, body force), updatePosition!
//  (UpdateTemperature.jl:8-12, UpdateVelocity.jl:8-23, UpdatePosition.jl:8-14).
// The pair terms are the reference's, regrouped so that each pair needs at
// most two reciprocals:
//   viscosity  4 (mu_i + mu_j) r W' / (rho_i + rho_j)^2 / (r^2 + 0.01 h^2)
//              = 4 nu0 r W' / ((rho_i + rho_j) (r^2 + 0.01 h^2))
//   conduction 2 k_eq r W' (T_i - T_j) / rho_i / rho_j / (r^2 + .01 h^2) m_j / cp_i
//              = 4 k_i k_j r W' (T_i - T_j) m_j / ((k_i + k_j) rho_j (r^2 + .01 h^2)) / (rho_i cp_i)
// and 1 / W(gap) is taken per particle when both gaps are equal.
// New x, v, T go to xn, vn, Tn in the current particle order.
// ---------------------------------------------------------------------------
// Buffer particles of the inlet/outlet solver: dv = 0, then updatePosition!
// (FluidInletOutletWallSolve.jl:170-186): x += v dt - 0 dt^2 / 2.
template <int D>
__device__ __forceinline__ void buffer_move(const SceneDev& s, const FluidSet& S, int i) {
#pragma clang fp contract(off)
    const double dt = s.dt;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double v = S.v[i * D + k];
        s.vn[i * D + k] = v;
        s.xn[i * D + k] = S.x[i * D + k] + (v * dt - 0.0 * (dt * dt) / 2.0);
    }
}

template <int D, int K, typename T, bool THERMAL, bool SCAN>
__device__ __forceinline__ void forces_body(const SceneDev& s, const FluidSet& S, int i) {
    using KQ = KernelQ<T, K>;
    if (S.kind && S.kind[i] != KIND_FLUID) {
        buffer_move<D>(s, S, i);
        return;
    }
    const T inv_h = (T)s.inv_h, kA = (T)s.kA, kB = (T)s.kB;
    const T eta2 = (T)(0.01 * (s.h * s.h));
    const T nu0 = (T)s.nu0;
    double xi[3];
    load_x<D>(S.x, i, xi);
    Rec<T> me;
    load_rec<D, T, THERMAL>(S, i, me);
    const T rho_i = me.rho, pr2_i = me.pr2, gap_i = me.gap, m_i = me.m, T_i = me.Tt, kap_i = me.kap;
    T W_gap, dW_gap;
    KQ::wdw(gap_i * inv_h, kA, kB, W_gap, dW_gap);
    const T inv_Wgap_i = prcp(W_gap);
    T dv[3] = {T(0), T(0), T(0)};
    T dT = T(0);
    for_pairs<D, SCAN>(s, s.cell_start, S.x, xi, i, i, s.nbr, s.nnb[i], [&](int j, double d2, const double* rv) {
        Rec<T> o;
        load_rec<D, T, THERMAL>(S, j, o);
        const T r = psqrt((T)d2);
        T W, dW;
        KQ::wdw(r * inv_h, kA, kB, W, dW);
        const T sg = d2 == 0.0 ? T(0) : dW * prcp(r);
        // pressure, with the 1% correction W(r) / W((gap_i + gap_j) / 2)
        T P = pr2_i + o.pr2;
        T inv_Wg = inv_Wgap_i;
        if (o.gap != gap_i) {
            T w2, d2w;
            KQ::wdw((gap_i + o.gap) * T(0.5) * inv_h, kA, kB, w2, d2w);
            inv_Wg = prcp(w2);
        }
        P += fabs(P) * T(0.01) * W * inv_Wg;
        // viscosity
        const T r2e = r * r + eta2;
        const T vf = T(4) * nu0 * r * dW * prcp((rho_i + o.rho) * r2e);
        const T ps = -P * sg;
#pragma unroll
        for (int k = 0; k < D; ++k) dv[k] += o.m * (ps * (T)rv[k] + vf * (me.v[k] - o.v[k]));
        if (THERMAL)
            dT += T(4) * kap_i * o.kap * r * dW * (T_i - o.Tt) * o.m * prcp((kap_i + o.kap) * o.rho * r2e);
    });
    if (THERMAL) dT *= prcp(rho_i * (T)S.cp[i]);
    if (s.nw > 0) {
        using V = typename Vec4<T>::type;
        const bool wall_pressure = (s.flags & FLAG_WALL_PRESSURE) != 0;
        const T c_i = (T)S.c[i];
        const T P0 = (T)(s.p0 / (s.rho0 * s.rho0));
        const T inv_rho_i = prcp(rho_i);
        const T th_w = (T)(1.0 / (s.cp * s.rho0)) * inv_rho_i * m_i;
        T dTw = T(0);
        for_pairs<D, SCAN>(s, s.wcell_start, s.wx, xi, i, -1, s.wnbr, s.nnbw[i], [&](int j, double d2, const double* rv) {
            const V wa = static_cast<const V*>(s.wA)[j];
            const V wb = static_cast<const V*>(s.wB)[j];
            T n[3], vw[3], gap_j, Tw, kw;
            n[0] = wa.x;
            n[1] = wa.y;
            vw[0] = wb.x;
            vw[1] = wb.y;
            if (D == 3) {
                n[2] = wa.z;
                gap_j = wa.w;
                vw[2] = wb.z;
                Tw = wb.w;
                kw = THERMAL ? ld<T>(s.wkap, j) : T(0);
            } else {
                gap_j = wa.z;
                Tw = wa.w;
                kw = wb.z;
            }
            const T r = psqrt((T)d2);
            T W, dW;
            KQ::wdw(r * inv_h, kA, kB, W, dW);
            const T sg = d2 == 0.0 ? T(0) : dW * prcp(r);
            T vv[3];
            T vn = T(0), rn = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                vv[k] = me.v[k] - vw[k];
                vn += vv[k] * n[k];
                rn += (T)rv[k] * n[k];
            }
            // wallForce! (Rogers & Dalrymple 2008)
            {
                const T psi = fabs(rn);
                const T a = r * r - psi * psi;
                const T xi_ = psqrt(a > T(0) ? a : T(0));
                const T q = psi / gap_j;
                if (!(q > T(1) || xi_ > gap_j)) {
                    const T p_xi = fabs((T(1) + cos(T(3.14159265358979323846) * xi_ / gap_j)) * T(0.5));
                    const T beta = vn > T(0) ? T(0) : T(1);
                    const T wf = (T(0.01) * (c_i * c_i) + beta * c_i * fabs(vn)) * inv_h * fabs(T(1) - q) *
                                 prsqrt(q) * p_xi;
#pragma unroll
                    for (int k = 0; k < D; ++k) dv[k] += wf * n[k];
                }
            }
            const T r2e = r * r + eta2;
            const T vf = T(2) * nu0 * r * dW * inv_rho_i * prcp(r2e);
            T ps = T(0);
            if (wall_pressure) {
                T P = pr2_i + P0;
                P += fabs(P) * T(0.01) * W * inv_Wgap_i;
                ps = -P * sg;
            }
#pragma unroll
            for (int k = 0; k < D; ++k) dv[k] += m_i * (ps * (T)rv[k] + vf * (vv[k] - vn * n[k]));
            if (THERMAL && s.wkind[j] == ESPH_WALL_THERMOSTATIC)
                dTw += (kap_i + kw) * r * dW * (T_i - Tw) * prcp(r2e);
        });
        if (THERMAL) dT += dTw * th_w;
    }
    {
#pragma clang fp contract(off)
        const double dt = s.dt;
        double bf[3];
        if (THERMAL) {
            const double Tn = S.T[i] + dt * (double)dT;
            s.Tn[i] = Tn;
#pragma unroll
            for (int k = 0; k < D; ++k) bf[k] = -s.g[k] * s.beta * (Tn - s.t0);
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) bf[k] = s.g[k];
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double v = S.v[i * D + k];
            v += bf[k] * dt;
            const double dvk = (double)dv[k];
            v += dvk * dt;
            s.vn[i * D + k] = v;
            s.xn[i * D + k] = xi[k] + (v * dt - dvk * (dt * dt) / 2.0);
        }
    }
}

template <int D, int K, typename T, bool THERMAL, bool SCAN>
__global__ __launch_bounds__(kBlock) void k_forces(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, int ntiles) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (((s.flags & FLAG_THERMAL) != 0) != THERMAL) continue;
        for_work<SCAN>(s, dyn, tile, [&](int i) {
            if (!SCAN && overflowed(s, i)) return;
            forces_body<D, K, T, THERMAL, SCAN>(s, s.set[dyn.parity], i);
        });
    }
}

// ---------------------------------------------------------------------------
// reconstructScalar!(fluid, :rho_, FF pairs, kernel) (GroupAction/
// ReconstructValue.jl:22-57): Shepard filter over the step-start pairs with
// the post-continuity density, in binary64 in every form.
// ---------------------------------------------------------------------------
template <int D, int K, bool SCAN>
__global__ __launch_bounds__(kBlock) void k_shepard(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, int ntiles) {
    using KQ = KernelQ<double, K>;
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (s.kreinit <= 0 || (dyn.step + 1) % s.kreinit != 0) continue;
        for_work<SCAN>(s, dyn, tile, [&](int i) {
        if (!SCAN && overflowed(s, i)) return;
        const FluidSet& S = s.set[dyn.parity];
        const int8_t* kind = S.kind;
        if (kind && kind[i] != KIND_FLUID) return;
        double xi[3];
        load_x<D>(S.x, i, xi);
        double sw = 0.0, ss = 0.0;
        for_pairs<D, SCAN>(s, s.cell_start, S.x, xi, i, i, s.nbr, s.nnb[i], [&](int j, double d2, const double*) {
            if (kind && kind[j] != KIND_FLUID) return;  // fluid-fluid pairs only
            double W, dW;
            KQ::wdw(sqrt(d2) * s.inv_h, s.kA, s.kB, W, dW);
            const double rho_j = S.rho[j];
            const double a = W * S.m[j] / rho_j;
            sw += a;
            ss += a * rho_j;
        });
        {
#pragma clang fp contract(off)
            const double rho_i = S.rho[i];
            const double a = s.W0 * S.m[i] / rho_i;
            sw += a;
            ss += a * rho_i;
            s.rhon[i] = ss / sw;
        }
        });
    }
}

// ---------------------------------------------------------------------------
// Counting sort of the step's new positions into cells.  Deterministic: the
// order inside a cell is the particles' previous order.  The check predicate
// (FluidWallSolve.jl:89-97, deleteat! is stable) sends deleted particles to
// the tail bucket `ncell`, which is dropped.
// mode 0: end of a step; mode 1: upload (rho from rhon, no check).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool check_due(const SceneDev& s, const SceneDyn& d) {
    return (s.flags & FLAG_CHECK) && s.kcheck > 0 && (d.step + 1) % s.kcheck == 0;
}
__device__ __forceinline__ bool shepard_due(const SceneDev& s, const SceneDyn& d) {
    return s.kreinit > 0 && (d.step + 1) % s.kreinit == 0;
}

__global__ __launch_bounds__(kBlock) void k_sort_clear(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ ctiles, int ntiles) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = ctiles[t];
        const SceneDev& s = scenes[tile.scene];
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            const int c = tile.begin + q * kBlock + (int)threadIdx.x;
            if (c <= s.ncell) s.count[c] = 0;
        }
        if (tile.begin == 0 && threadIdx.x == 0) s.dyn->outside = 0;
    }
}

__device__ __forceinline__ int cell_key(const SceneDev& s, int D, const double* x) {
    const int cx = cell_coord(x[0], s.lo[0], s.inv_cs, s.nc[0]);
    const int cy = cell_coord(x[1], s.lo[1], s.inv_cs, s.nc[1]);
    const int cz = D == 3 ? cell_coord(x[2], s.lo[2], s.inv_cs, s.nc[2]) : 0;
    return cell_linear(s, cx, cy, cz);
}

// Spawned copies valid this step (the rest exceeded the capacity).
__device__ __forceinline__ int spawn_count(const SceneDev& s, const SceneDyn& d) {
    const int room = s.cap - d.n_live;
    return d.nspawn < room ? d.nspawn : room;
}

// Slot of this lane's key in count[key]: lanes of one wavefront holding the
// same key in consecutive lanes (the usual case: particles are sorted by
// cell) share one atomic; the run head adds the run length and each lane
// takes head result + its offset in the run.  Works for any active mask.
__device__ __forceinline__ int wave_counted_slot(int32_t* count, int key) {
    const unsigned lane = threadIdx.x & 63u;
    const uint64_t active = __ballot(1);
    const int prev = __shfl(key, (int)((lane + 63u) & 63u), 64);
    // run head: first active lane, or the previous lane inactive, or a different key
    const bool prev_active = lane > 0 && ((active >> (lane - 1)) & 1ull);
    const bool head = !prev_active || prev != key;
    const uint64_t heads = __ballot(head) & active;
    // this lane's run head and the next head after it
    const uint64_t upto = heads & ((lane == 63u) ? ~0ull : ((2ull << lane) - 1ull));
    const int my_head = 63 - __clzll((long long)upto);
    // run end: the first lane after this one that is a head or inactive
    const uint64_t above = (lane == 63u) ? 0ull : ~((2ull << lane) - 1ull);
    const uint64_t stop = (heads | ~active) & above;
    const int run_end = stop ? __ffsll((long long)stop) - 1 : 64;
    int base = 0;
    if (head) base = atomicAdd(&count[key], run_end - (int)lane);
    base = __shfl(base, my_head, 64);
    return base + ((int)lane - my_head);
}

__global__ __launch_bounds__(kBlock) void k_sort_bin(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const int D = s.dim;
        double x[3] = {0.0, 0.0, 0.0};
        bool finite = true;
        for (int k = 0; k < D; ++k) {
            x[k] = s.xn[i * D + k];
            finite = finite && isfinite(x[k]);
        }
        int key;
        bool dead = false;
        if (mode == 0 && check_due(s, dyn)) {
            for (int k = 0; k < D; ++k)
                if (x[k] > s.check_max[k] || x[k] < s.check_min[k]) dead = true;
        }
        // inlet/outlet migration (FluidInletOutletWallSolve.jl:196-262)
        int spawn_key = -1;
        if (mode == 0 && (s.flags & FLAG_IO)) {
            const FluidSet& S = s.set[dyn.parity];
            const int kd = S.kind[i], a = s.io_axis;
            if (kd == KIND_OUTLET && x[a] > s.outlet_exit) dead = true;
            if (kd == KIND_FLUID && x[a] > s.outlet_enter) s.leave_id[atomicAdd(&s.dyn->nleave, 1)] = S.id[i];
            if (kd == KIND_INLET && x[a] > s.inlet_enter) {
                spawn_key = cell_key(s, D, x);  // the fluid copy stays where the inlet particle is
                x[a] -= s.inlet_shift;          // the inlet particle moves back
            }
        }
        key = dead ? s.ncell : cell_key(s, D, x);
        if (!finite) atomicOr(&s.dyn->err, 1);
        {
            // particles beyond the grid share its border cells (clamped):
            // counted, since many of them make those cells crowded
            bool far = false;
            for (int k = 0; k < D; ++k) {
                const double f = floor((x[k] - s.lo[k]) * s.inv_cs);
                far = far || !(f >= -1.0 && f <= (double)s.nc[k]);
            }
            if (far && !dead) atomicAdd(&s.dyn->outside, 1);
        }
        s.key[i] = key;
        s.slot[i] = wave_counted_slot(s.count, key);
        if (spawn_key >= 0) {
            const int e = atomicAdd(&s.dyn->nspawn, 1);
            if (dyn.n_live + e < s.cap) {
                s.spawn_src[e] = i;
                s.spawn_id[e] = s.set[dyn.parity].id[i];
                s.spawn_key[e] = spawn_key;
                s.spawn_slot[e] = atomicAdd(&s.count[spawn_key], 1);
            } else {
                atomicOr(&s.dyn->err, 2);  // capacity exhausted: the copy is dropped
            }
        }
    }
}

// Inclusive scan over a wavefront (64 lanes).
__device__ __forceinline__ int wave_inclusive_scan(int x) {
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int wsum[kBlock / 64];
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int incl = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        off += w < wave ? wsum[w] : 0;
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// Phase 1: scan of kCellTile counts per tile; tile total to bsum.
__global__ __launch_bounds__(kBlock) void k_sort_scan1(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ ctiles, int ntiles) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = ctiles[t];
        const SceneDev& s = scenes[tile.scene];
        const int c0 = tile.begin + (int)threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = c0 + q <= s.ncell ? s.count[c0 + q] : 0;
            sum += v[q];
        }
        int total;
        int run = block_exclusive_scan(sum, &total);
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            if (c0 + q <= s.ncell) s.cell_start[c0 + q] = run;
            run += v[q];
        }
        if (threadIdx.x == 0) s.bsum[tile.begin / kCellTile] = total;
    }
}

// Phase 2: one block per scene scans its tile totals (exclusive, in place).
__global__ __launch_bounds__(kBlock) void k_sort_scan2(const SceneDev* __restrict__ scenes, int nscenes) {
    for (int sc = blockIdx.x; sc < nscenes; sc += gridDim.x) {
        const SceneDev& s = scenes[sc];
        int carry = 0;
        for (int base = 0; base < s.cell_tiles; base += kBlock) {
            const int q = base + (int)threadIdx.x;
            const int v = q < s.cell_tiles ? s.bsum[q] : 0;
            int total;
            const int ex = block_exclusive_scan(v, &total);
            if (q < s.cell_tiles) s.bsum[q] = carry + ex;
            carry += total;
        }
    }
}

// Phase 3: add tile offsets; the last cell also writes the grand total.
__global__ __launch_bounds__(kBlock) void k_sort_scan3(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ ctiles, int ntiles) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = ctiles[t];
        const SceneDev& s = scenes[tile.scene];
        const int off = s.bsum[tile.begin / kCellTile];
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            const int c = tile.begin + (int)threadIdx.x * kScanItems + q;
            if (c <= s.ncell) {
                const int st = s.cell_start[c] + off;
                s.cell_start[c] = st;
                if (c == s.ncell) s.cell_start[c + 1] = st + s.count[c];
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_place(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const int i = tile.begin + (int)threadIdx.x;
        const SceneDyn dyn = *s.dyn;
        if (i < dyn.n_live) {
            s.tmp[s.cell_start[s.key[i]] + s.slot[i]] = i;
        } else if (mode == 0 && (s.flags & FLAG_IO) && i < dyn.n_live + spawn_count(s, dyn)) {
            // a fluid copy of an entering inlet particle; its rank among this
            // step's copies follows the inlet order (push! in that order)
            const int e = i - dyn.n_live, ns = spawn_count(s, dyn), my = s.spawn_id[e];
            int r = 0;
            for (int q = 0; q < ns; ++q) r += s.spawn_id[q] < my ? 1 : 0;
            s.spawn_order[r] = e;
            s.tmp[s.cell_start[s.spawn_key[e]] + s.spawn_slot[e]] = dyn.n_live + r;
        }
    }
}

// Velocity profile of the inlet/outlet buffers (the examples'
// user_defined_function): PoiseuilleFlow2D.jl:28-35, PoiseuilleFlow3D.jl:32-34.
__device__ double profile_u(const SceneDev& s, const double* x, double t) {
#pragma clang fp contract(off)
    const double f = s.profile_f, nu = s.profile_nu, l = s.profile_l;
    if (s.profile == ESPH_PROFILE_CHANNEL_2D) {
        const double PI = 3.14159265358979323846;
        const double y = x[1] - s.profile_c[1];
        double ux = f / 2.0 / nu * y * (l - y);
        for (int n = 0; n <= s.profile_terms; ++n) {
            const double o = (double)(2 * n + 1);
            ux -= 4.0 * f * (l * l) / nu / (PI * PI * PI) / (o * o * o) * sin(PI * y / l * o) *
                  exp(-(o * o) * (PI * PI) * nu * t / (l * l));
        }
        return ux;
    }
    if (s.profile == ESPH_PROFILE_PIPE_3D) {
        const double y = x[1] - s.profile_c[1], z = x[2] - s.profile_c[2];
        return f / 4.0 / nu * (l * l - y * y - z * z);
    }
    return 0.0;
}

// Final position = cell start + number of cell members with a smaller
// entry index (previous order; a step's fluid copies of inlet particles come
// after every existing particle, in inlet order); gathers every field into
// the other set.  With the inlet/outlet solver it also applies the step's
// migration (kinds and group sequence numbers) and the buffer velocities.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const SceneDev* __restrict__ scenes,
                                                         const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const bool io = mode == 0 && (s.flags & FLAG_IO);
        const int ns = io ? spawn_count(s, dyn) : 0;
        const int q = tile.begin + (int)threadIdx.x;
        if (q >= dyn.n_live + ns) continue;
        const bool spawn = q >= dyn.n_live;
        int src, key, sp_rank = 0;
        if (spawn) {
            sp_rank = q - dyn.n_live;
            const int e = s.spawn_order[sp_rank];
            src = s.spawn_src[e];
            key = s.spawn_key[e];
        } else {
            src = q;
            key = s.key[q];
            if (key == s.ncell) continue;  // deleted
        }
        const int b = s.cell_start[key], e = s.cell_start[key + 1];
        int rank = 0;
        for (int w = b; w < e; ++w) rank += s.tmp[w] < q ? 1 : 0;
        const int p = b + rank;
        const FluidSet& S = s.set[dyn.parity];
        const FluidSet& O = s.set[dyn.parity ^ 1];
        const int D = s.dim;
        const bool thermal = (s.flags & FLAG_THERMAL) != 0;
        double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < D; ++k) {
            x[k] = s.xn[src * D + k];
            v[k] = s.vn[src * D + k];
        }
        int8_t kd = S.kind ? S.kind[src] : (int8_t)KIND_FLUID;
        const bool src_fluid = kd == KIND_FLUID;
        int32_t id = S.id[src];
        if (io) {
            const int a = s.io_axis;
            if (spawn) {
                kd = KIND_FLUID;  // push!(fluid, deepcopy(inlet particle))
                id = dyn.next_fluid_id + sp_rank;
            } else if (kd == KIND_INLET && x[a] > s.inlet_enter) {
                x[a] -= s.inlet_shift;  // inlet_into_calculation_domain_modify
            } else if (kd == KIND_FLUID && x[a] > s.outlet_enter) {
                int r = 0;  // push! onto the outlet in fluid order
                for (int l = 0; l < dyn.nleave; ++l) r += s.leave_id[l] < id ? 1 : 0;
                kd = KIND_OUTLET;
                id = dyn.next_outlet_id + r;
            }
            if (kd != KIND_FLUID && s.profile != ESPH_PROFILE_NONE) {
                const double tt = (double)(dyn.step + 1) * s.dt;
                if (kd == KIND_INLET || s.outlet_profile) v[a] = profile_u(s, x, tt);
                if (kd == KIND_OUTLET)
                    for (int k = 0; k < D; ++k)
                        if (k != a) v[k] = 0.0;
            }
        }
        for (int k = 0; k < D; ++k) {
            O.x[p * D + k] = x[k];
            O.v[p * D + k] = v[k];
        }
        if (D == 3)
            static_cast<float4*>(O.xf)[p] = make_float4((float)(x[0] - s.lo[0]), (float)(x[1] - s.lo[1]),
                                                        (float)(x[2] - s.lo[2]), 0.0f);
        else
            static_cast<float2*>(O.xf)[p] = make_float2((float)(x[0] - s.lo[0]), (float)(x[1] - s.lo[1]));
        const bool rho_new = mode == 1 || (shepard_due(s, dyn) && src_fluid && !spawn);
        const double rho = rho_new ? s.rhon[src] : S.rho[src];
        const double pp = S.p[src];
        O.rho[p] = rho;
        O.p[p] = pp;
        O.c[p] = S.c[src];
        const double m = S.m[src], gap = S.gap[src];
        O.m[p] = m;
        O.gap[p] = gap;
        O.id[p] = id;
        if (O.kind) O.kind[p] = kd;
        double Tt = 0.0, kap = 0.0;
        if (thermal) {
            Tt = s.Tn[src];
            kap = S.kappa[src];
            O.T[p] = Tt;
            O.kappa[p] = kap;
            O.cp[p] = S.cp[src];
        }
        if (s.prec == ESPH_FP32_BULK) {
            store_sorted_rec<float>(O, D, p, v, m, gap, Tt, kap, thermal);
            if (kd != KIND_FLUID) {  // buffers never pass the continuity pass
                float* B = static_cast<float*>(O.B) + 4 * (size_t)p;
                B[0] = (float)rho;
                B[1] = (float)(pp / (rho * rho));
            }
        } else {
            store_sorted_rec<double>(O, D, p, v, m, gap, Tt, kap, thermal);
            if (kd != KIND_FLUID) {
                double* B = static_cast<double*>(O.B) + 4 * (size_t)p;
                B[0] = rho;
                B[1] = pp / (rho * rho);
            }
        }
    }
}

__global__ void k_sort_finish(const SceneDev* __restrict__ scenes, int nscenes, int mode) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= nscenes) return;
    const SceneDev& s = scenes[sc];
    SceneDyn* d = s.dyn;
    if (mode == 0 && (s.flags & FLAG_IO)) {
        const int room = s.cap - d->n_live;  // n_live is still the step's input count here
        d->next_fluid_id += d->nspawn < room ? d->nspawn : room;
        d->next_outlet_id += d->nleave;
    }
    d->n_live = s.cell_start[s.ncell];
    d->parity ^= 1;
    d->novf = 0;
    d->nspawn = 0;
    d->nleave = 0;
    if (mode == 0) d->step += 1;
}

// ---------------------------------------------------------------------------
// Pair export (neighbour sets for checking and for the standalone
// neighbour-list entry point).  pairset 0: FF pairs with id_i < id_j;
// pairset 1: (fluid id, wall id).
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_pairs(const SceneDev* __restrict__ scenes,
                                                  const Tile* __restrict__ tiles, int ntiles, int pairset,
                                                  int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                                                  int32_t* __restrict__ out) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[dyn.parity];
        const int8_t* kind = S.kind;  // inlet/outlet solver: fluid particles only
        int64_t w = offs ? offs[i] : 0;
        int cnt = 0;
        if (kind && kind[i] != KIND_FLUID) {
            if (counts) counts[i] = 0;
            continue;
        }
        double xi[3];
        load_x<D>(S.x, i, xi);
        int c[3];
        cell_of<D>(s, xi, c);
        const int idi = S.id[i];
        if (pairset == ESPH_PAIRS_FLUID_FLUID) {
            for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i, [&](int j, double, const double*) {
                const int idj = S.id[j];
                if (kind && kind[j] != KIND_FLUID) return;
                if (idj > idi) {
                    if (out) { out[2 * w] = idi; out[2 * w + 1] = idj; ++w; }
                    ++cnt;
                }
            });
        } else if (s.nw > 0) {
            for_each_neighbour<D>(s, s.wcell_start, s.wx, xi, c, -1, [&](int j, double, const double*) {
                if (out) { out[2 * w] = idi; out[2 * w + 1] = s.wid[j]; ++w; }
                ++cnt;
            });
        }
        if (counts) counts[i] = cnt;
    }
}

// kernelValueInterpolation (DataIO/VTPToDataFrame.jl:77-142) over the
// resident fluid: one probe point per thread, the probe's 3^D cells of the
// current cell list, FP64 sums of W m_j / rho_j and W m_j / rho_j f_j in cell
// order; none_value where no fluid particle is within R.
template <int D, int K>
__global__ __launch_bounds__(kBlock) void k_probe(const SceneDev* __restrict__ sd, int64_t np,
                                                  const double* __restrict__ pts, int field, double none_value,
                                                  double* __restrict__ out) {
#pragma clang fp contract(off)
    const SceneDev& s = *sd;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const SceneDyn dyn = *s.dyn;
    const FluidSet& S = s.set[dyn.parity];
    const double* f = S.rho;
    int w = 1;
    if (field == ESPH_FIELD_V) { f = S.v; w = D; }
    else if (field == ESPH_FIELD_P) f = S.p;
    else if (field == ESPH_FIELD_C) f = S.c;
    else if (field == ESPH_FIELD_T) f = S.T;
    double xi[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < D; ++k) xi[k] = pts[q * D + k];
    int c[3];
    cell_of<D>(s, xi, c);
    double sw = 0.0, sf[3] = {0.0, 0.0, 0.0};
    for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, -1, [&](int j, double d2, const double*) {
        if (S.kind && S.kind[j] != KIND_FLUID) return;
        const double W = SmoothKernelFn<double, K>::value(sqrt(d2), s.h, s.sigma);
        const double wj = W * S.m[j] / S.rho[j];
        sw += wj;
        for (int k = 0; k < w; ++k) sf[k] += wj * f[(size_t)j * w + k];
    });
    for (int k = 0; k < w; ++k) out[q * w + k] = sw > 0.0 ? sf[k] / sw : none_value;
}

// Kernel values on the device (checks of the device math).
template <int K>
__global__ void k_kernel_eval(int64_t n, const double* __restrict__ r, double* __restrict__ w,
                              double* __restrict__ dw, double h, double sigma) {
    using SK = SmoothKernelFn<double, K>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] = SK::value(r[i], h, sigma);
    dw[i] = SK::gradient(r[i], h, sigma);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define ESPH_DISPATCH_K(KVAR, BODY)                                                 \
    switch (KVAR) {                                                                 \
        case ESPH_CUBIC_SPLINE: { constexpr int K = ESPH_CUBIC_SPLINE; BODY; } break; \
        case ESPH_GAUSSIAN: { constexpr int K = ESPH_GAUSSIAN; BODY; } break;         \
        case ESPH_WENDLAND_C2: { constexpr int K = ESPH_WENDLAND_C2; BODY; } break;   \
        case ESPH_WENDLAND_C4: { constexpr int K = ESPH_WENDLAND_C4; BODY; } break;   \
        default: return hipErrorInvalidValue;                                       \
    }

static inline int grid_for(int ntiles) { return ntiles > 0 ? ntiles : 1; }

template <typename L>
static hipError_t launch_both(const Bucket& b, hipStream_t st, L&& launch) {
    if (b.ntiles <= 0) return hipSuccess;
    launch(std::false_type{}, dim3(grid_for(b.ntiles)), b.tiles, b.ntiles);
    launch(std::true_type{}, dim3(kOvfTiles * b.nscenes), b.ovf_tiles, kOvfTiles * b.nscenes);
    return hipGetLastError();
}

hipError_t launch_continuity(const Bucket& b, hipStream_t st) {
    dim3 blk(kBlock);
    hipError_t e = hipSuccess;
    ESPH_DISPATCH_K(b.kernel, {
        e = launch_both(b, st, [&](auto scan, dim3 g, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            if (b.dim == 2 && b.prec == ESPH_FP64) hipLaunchKernelGGL((k_continuity<2, K, double, S>), g, blk, 0, st, b.scenes, tl, nt);
            else if (b.dim == 2) hipLaunchKernelGGL((k_continuity<2, K, float, S>), g, blk, 0, st, b.scenes, tl, nt);
            else if (b.prec == ESPH_FP64) hipLaunchKernelGGL((k_continuity<3, K, double, S>), g, blk, 0, st, b.scenes, tl, nt);
            else hipLaunchKernelGGL((k_continuity<3, K, float, S>), g, blk, 0, st, b.scenes, tl, nt);
        });
    });
    return e;
}

template <bool TH>
static hipError_t launch_forces_t(const Bucket& b, hipStream_t st) {
    dim3 blk(kBlock);
    hipError_t e = hipSuccess;
    ESPH_DISPATCH_K(b.kernel, {
        e = launch_both(b, st, [&](auto scan, dim3 g, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            if (b.dim == 2 && b.prec == ESPH_FP64) hipLaunchKernelGGL((k_forces<2, K, double, TH, S>), g, blk, 0, st, b.scenes, tl, nt);
            else if (b.dim == 2) hipLaunchKernelGGL((k_forces<2, K, float, TH, S>), g, blk, 0, st, b.scenes, tl, nt);
            else if (b.prec == ESPH_FP64) hipLaunchKernelGGL((k_forces<3, K, double, TH, S>), g, blk, 0, st, b.scenes, tl, nt);
            else hipLaunchKernelGGL((k_forces<3, K, float, TH, S>), g, blk, 0, st, b.scenes, tl, nt);
        });
    });
    return e;
}

hipError_t launch_forces(const Bucket& b, hipStream_t st) {
    if (b.has_thermal) {
        hipError_t e = launch_forces_t<true>(b, st);
        if (e != hipSuccess) return e;
    }
    if (b.has_plain) return launch_forces_t<false>(b, st);
    return hipSuccess;
}

hipError_t launch_shepard(const Bucket& b, hipStream_t st) {
    dim3 blk(kBlock);
    hipError_t e = hipSuccess;
    ESPH_DISPATCH_K(b.kernel, {
        e = launch_both(b, st, [&](auto scan, dim3 g, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            if (b.dim == 2) hipLaunchKernelGGL((k_shepard<2, K, S>), g, blk, 0, st, b.scenes, tl, nt);
            else hipLaunchKernelGGL((k_shepard<3, K, S>), g, blk, 0, st, b.scenes, tl, nt);
        });
    });
    return e;
}

hipError_t launch_sort(const SortPlan& p, int mode, hipStream_t st) {
    dim3 blk(kBlock);
    hipLaunchKernelGGL(k_sort_clear, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_bin, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_scan1, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_scan2, dim3(grid_for(p.nscenes)), blk, 0, st, p.scenes, p.nscenes);
    hipLaunchKernelGGL(k_sort_scan3, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_place, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_scatter, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_finish, dim3((p.nscenes + 63) / 64), dim3(64), 0, st, p.scenes, p.nscenes, mode);
    return hipGetLastError();
}

hipError_t launch_pairs(const Bucket& b, int pairset, int32_t* counts, const int64_t* offs, int32_t* out,
                        hipStream_t st) {
    if (b.ntiles <= 0) return hipSuccess;
    dim3 g(grid_for(b.ntiles)), blk(kBlock);
    if (b.dim == 2)
        hipLaunchKernelGGL(k_pairs<2>, g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
    else
        hipLaunchKernelGGL(k_pairs<3>, g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
    return hipGetLastError();
}

hipError_t launch_probe(const Bucket& b, int64_t np, const double* pts, int field, double none_value, double* out,
                        hipStream_t st) {
    if (np <= 0) return hipSuccess;
    dim3 g((unsigned)((np + kBlock - 1) / kBlock)), blk(kBlock);
    if (b.dim == 2) {
        ESPH_DISPATCH_K(b.kernel, {
            hipLaunchKernelGGL((k_probe<2, K>), g, blk, 0, st, b.scenes, np, pts, field, none_value, out);
        });
    } else {
        ESPH_DISPATCH_K(b.kernel, {
            hipLaunchKernelGGL((k_probe<3, K>), g, blk, 0, st, b.scenes, np, pts, field, none_value, out);
        });
    }
    return hipGetLastError();
}

hipError_t launch_kernel_eval(int kernel, int64_t n, const double* r, double* w, double* dw, double h,
                              double sigma, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), blk(256);
    ESPH_DISPATCH_K(kernel, { hipLaunchKernelGGL(k_kernel_eval<K>, g, blk, 0, st, n, r, w, dw, h, sigma); });
    return hipGetLastError();
}

}  // namespace esph
