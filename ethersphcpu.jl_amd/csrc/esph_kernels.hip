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

namespace esph {

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

// Iterates the particles of cells adjacent to cell c (3 x-contiguous ranges
// in 2D, 9 in 3D) whose distance to xi satisfies the neighbour rule.
template <int D, typename F>
__device__ __forceinline__ void for_each_neighbour(const SceneDev& s, const int32_t* __restrict__ cstart,
                                                   const double* __restrict__ xall, const double* xi,
                                                   const int c[3], int self, F&& f) {
    const int zlo = D == 3 ? -1 : 0, zhi = D == 3 ? 1 : 0;
    const int cx0 = c[0] > 0 ? c[0] - 1 : 0;
    const int cx1 = c[0] + 1 < s.nc[0] ? c[0] + 1 : s.nc[0] - 1;
    for (int dz = zlo; dz <= zhi; ++dz) {
        const int cz = c[2] + dz;
        if (cz < 0 || cz >= s.nc[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int cy = c[1] + dy;
            if (cy < 0 || cy >= s.nc[1]) continue;
            const int j0 = cstart[cell_linear(s, cx0, cy, cz)];
            const int j1 = cstart[cell_linear(s, cx1, cy, cz) + 1];
            for (int j = j0; j < j1; ++j) {
                if (j == self) continue;
                double xj[3], rv[3];
                load_x<D>(xall, j, xj);
                const double d2 = dist2_exact<D>(xi, xj, rv);
                if (d2 <= s.R2) f(j, d2, rv);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// continuity! (Interaction/Continuity.jl:8-17), pull form:
//   drho_i = sum_j m_j (v_i - v_j) . gradW_ij
// then updateDensity! (SelfAction/UpdateDensity.jl:8-12) and updatePressure!
// (SelfAction/UpdatePressure.jl:8-16).  Writes rho, p, c and p/rho^2.
// ---------------------------------------------------------------------------
template <int D, int K, typename T>
__global__ __launch_bounds__(kBlock) void k_continuity(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles, int ntiles) {
    using SK = SmoothKernelFn<T, K>;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[dyn.parity];
        const T h = (T)s.h, sigma = (T)s.sigma;
        double xi[3];
        load_x<D>(S.x, i, xi);
        T vi[3];
#pragma unroll
        for (int k = 0; k < D; ++k) vi[k] = ld<T>(S.vs, i * D + k);
        int c[3];
        cell_of<D>(s, xi, c);
        T drho = T(0);
        for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i, [&](int j, double d2, const double* rv) {
            const double r = sqrt(d2);
            const T dW = SK::gradient((T)r, h, sigma);
            const T inv = T(1) / (T)r;
            T dot = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                T g = dW * ((T)rv[k] * inv);
                if (r == 0.0) g = T(0);  // normalize(0) is NaN -> 0 (Neighbour.jl:54-59)
                dot = dot + (vi[k] - ld<T>(S.vs, j * D + k)) * g;
            }
            drho += dot * ld<T>(S.ms, j);
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
            static_cast<T*>(S.pr2s)[i] = (T)(p / (rho * rho));
            if (sizeof(T) != sizeof(double)) static_cast<T*>(S.rhos)[i] = (T)rho;
        }
    }
}

// ---------------------------------------------------------------------------
// Momentum (and heat) pass + integration.
//  FF: pressureForce! (PressureForce.jl:8-21), viscosityForce!
//      (ViscosityForce.jl:8-25), thermalConduction! (ThermalConduction.jl:8-22)
//  FW: wallForce! (WallForce.jl:8-30), pressureForce! wall
//      (PressureForce.jl:23-35, solvers other than FluidWallSolve),
//      viscosityForce! wall (ViscosityForce.jl:27-46), thermalConduction!
//      thermostat (ThermalConduction.jl:24-36)
//  then updateTemperature!, updateVelocity!(…, body force), updatePosition!
//  (UpdateTemperature.jl:8-12, UpdateVelocity.jl:8-23, UpdatePosition.jl:8-14).
// New x, v, T go to xn, vn, Tn in the current particle order.
// ---------------------------------------------------------------------------
template <int D, int K, typename T>
__global__ __launch_bounds__(kBlock) void k_forces(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, int ntiles) {
    using SK = SmoothKernelFn<T, K>;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[dyn.parity];
        const bool thermal = (s.flags & FLAG_THERMAL) != 0;
        const T h = (T)s.h, sigma = (T)s.sigma;
        const T eta2 = (T)(0.01 * (s.h * s.h));
        const T nu0 = (T)s.nu0;
        double xi[3];
        load_x<D>(S.x, i, xi);
        T vi[3];
#pragma unroll
        for (int k = 0; k < D; ++k) vi[k] = ld<T>(S.vs, i * D + k);
        const T rho_i = sizeof(T) == sizeof(double) ? (T)S.rho[i] : ld<T>(S.rhos, i);
        const T pr2_i = ld<T>(S.pr2s, i);
        const T gap_i = ld<T>(S.gaps, i);
        const T m_i = ld<T>(S.ms, i);
        const T mu_i = nu0 * rho_i;
        T T_i = T(0), kap_i = T(0), cp_i = T(1);
        if (thermal) {
            T_i = ld<T>(S.Ts, i);
            kap_i = ld<T>(S.kappas, i);
            cp_i = ld<T>(S.cps, i);
        }
        int c[3];
        cell_of<D>(s, xi, c);
        T dv[3] = {T(0), T(0), T(0)};
        T dT = T(0);
        for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i, [&](int j, double d2, const double* rv) {
            const double rd = sqrt(d2);
            const T r = (T)rd;
            const T W = SK::value(r, h, sigma);
            const T dW = SK::gradient(r, h, sigma);
            const T inv = T(1) / r;
            T g[3];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                g[k] = dW * ((T)rv[k] * inv);
                if (rd == 0.0) g[k] = T(0);
            }
            const T m_j = ld<T>(S.ms, j);
            const T rho_j = ld<T>(S.rhos, j);
            // pressure
            T P = pr2_i + ld<T>(S.pr2s, j);
            P += fabs(P) * T(0.01) * W / SK::value((gap_i + ld<T>(S.gaps, j)) / T(2), h, sigma);
            // viscosity
            const T sum_rho = rho_i + rho_j;
            const T mu_j = nu0 * rho_j;
            const T vf = T(4) * (mu_i + mu_j) * r * dW / (sum_rho * sum_rho) / (r * r + eta2);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                dv[k] += -m_j * P * g[k];
                dv[k] += m_j * vf * (vi[k] - ld<T>(S.vs, j * D + k));
            }
            if (thermal) {
                const T kap_j = ld<T>(S.kappas, j);
                const T keq = T(2) * kap_i * kap_j / (kap_i + kap_j);
                const T dq = T(2) * keq * r * dW / rho_i / rho_j / (r * r + eta2) * (T_i - ld<T>(S.Ts, j));
                dT += dq * m_j / cp_i;
            }
        });
        if (s.nw > 0) {
            const bool wall_pressure = (s.flags & FLAG_WALL_PRESSURE) != 0;
            const T c_i = (T)S.c[i];
            const T P0 = (T)(s.p0 / (s.rho0 * s.rho0));
            const T Wgap_i = SK::value(gap_i, h, sigma);
            const T rho0 = (T)s.rho0;
            const T inv_cpm = (T)(1.0 / s.cp);
            for_each_neighbour<D>(s, s.wcell_start, s.wx, xi, c, -1, [&](int j, double d2, const double* rv) {
                const double rd = sqrt(d2);
                const T r = (T)rd;
                const T W = SK::value(r, h, sigma);
                const T dW = SK::gradient(r, h, sigma);
                const T inv = T(1) / r;
                T g[3], n[3], vv[3];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    g[k] = dW * ((T)rv[k] * inv);
                    if (rd == 0.0) g[k] = T(0);
                    n[k] = ld<T>(s.wns, j * D + k);
                    vv[k] = vi[k] - ld<T>(s.wvs, j * D + k);
                }
                const T gap_j = ld<T>(s.wgaps, j);
                T vn = T(0), rn = T(0);
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    vn = vn + vv[k] * n[k];
                    rn = rn + (T)rv[k] * n[k];
                }
                // wallForce! (Rogers & Dalrymple 2008)
                {
                    const T psi = fabs(rn);
                    const T a = r * r - psi * psi;
                    const T xi_ = sqrt(a > T(0) ? a : T(0));
                    const T q = psi / gap_j;
                    if (!(q > T(1) || xi_ > gap_j)) {
                        const T p_xi = fabs((T(1) + cos(T(3.14159265358979323846) * xi_ / gap_j)) / T(2));
                        const T beta = vn > T(0) ? T(0) : T(1);
                        const T r_psi = (T(0.01) * (c_i * c_i) + beta * c_i * fabs(vn)) / h * fabs(T(1) - q) / sqrt(q);
                        const T wf = r_psi * p_xi;
#pragma unroll
                        for (int k = 0; k < D; ++k) dv[k] += wf * n[k];
                    }
                }
                if (wall_pressure) {
                    T P = pr2_i + P0;
                    P += fabs(P) * T(0.01) * W / Wgap_i;
#pragma unroll
                    for (int k = 0; k < D; ++k) dv[k] += -m_i * P * g[k];
                }
                {
                    const T sum_rho = T(2) * rho_i;
                    const T vf = T(8) * mu_i * r * dW / (sum_rho * sum_rho) / (r * r + eta2);
#pragma unroll
                    for (int k = 0; k < D; ++k) dv[k] += m_i * vf * (vv[k] - vn * n[k]);
                }
                if (thermal && s.wkind[j] == ESPH_WALL_THERMOSTATIC) {
                    const T dq = inv_cpm * (kap_i + ld<T>(s.wkappas, j)) * r * dW / rho_i / rho0 / (r * r + eta2) *
                                 (T_i - ld<T>(s.wTs, j));
                    dT += dq * m_i;
                }
            });
        }
        {
#pragma clang fp contract(off)
            const double dt = s.dt;
            double bf[3];
            if (thermal) {
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
}

// ---------------------------------------------------------------------------
// reconstructScalar!(fluid, :rho_, FF pairs, kernel) (GroupAction/
// ReconstructValue.jl:22-57): Shepard filter over the step-start pairs with
// the post-continuity density, summed in binary64 in every form.
// ---------------------------------------------------------------------------
template <int D, int K>
__global__ __launch_bounds__(kBlock) void k_shepard(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, int ntiles) {
    using SK = SmoothKernelFn<double, K>;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (s.kreinit <= 0 || (dyn.step + 1) % s.kreinit != 0) continue;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[dyn.parity];
        double xi[3];
        load_x<D>(S.x, i, xi);
        int c[3];
        cell_of<D>(s, xi, c);
        double sw = 0.0, ss = 0.0;
        for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i, [&](int j, double d2, const double*) {
            const double W = SK::value(sqrt(d2), s.h, s.sigma);
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
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = ctiles[t];
        const SceneDev& s = scenes[tile.scene];
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            const int c = tile.begin + q * kBlock + (int)threadIdx.x;
            if (c <= s.ncell) s.count[c] = 0;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_bin(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
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
        if (dead) {
            key = s.ncell;
        } else {
            int c[3];
            c[0] = cell_coord(x[0], s.lo[0], s.inv_cs, s.nc[0]);
            c[1] = cell_coord(x[1], s.lo[1], s.inv_cs, s.nc[1]);
            c[2] = D == 3 ? cell_coord(x[2], s.lo[2], s.inv_cs, s.nc[2]) : 0;
            key = cell_linear(s, c[0], c[1], c[2]);
        }
        if (!finite) atomicOr(&s.dyn->err, 1);
        s.key[i] = key;
        s.slot[i] = atomicAdd(&s.count[key], 1);
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
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
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
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
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
                                                       const Tile* __restrict__ tiles, int ntiles) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= s.dyn->n_live) continue;
        s.tmp[s.cell_start[s.key[i]] + s.slot[i]] = i;
    }
}

// Final position = cell start + number of cell members with a smaller
// previous index; gathers every field of particle i into the other set.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const SceneDev* __restrict__ scenes,
                                                         const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const int key = s.key[i];
        if (key == s.ncell) continue;  // deleted
        const int b = s.cell_start[key], e = s.cell_start[key + 1];
        int rank = 0;
        for (int q = b; q < e; ++q) rank += s.tmp[q] < i ? 1 : 0;
        const int p = b + rank;
        const FluidSet& S = s.set[dyn.parity];
        const FluidSet& O = s.set[dyn.parity ^ 1];
        const int D = s.dim;
        const bool f32 = s.prec == ESPH_FP32_BULK;
        for (int k = 0; k < D; ++k) {
            const double x = s.xn[i * D + k];
            const double v = s.vn[i * D + k];
            O.x[p * D + k] = x;
            O.v[p * D + k] = v;
            if (f32) static_cast<float*>(O.vs)[p * D + k] = (float)v;
        }
        const bool rho_new = mode == 1 || shepard_due(s, dyn);
        const double rho = rho_new ? s.rhon[i] : S.rho[i];
        O.rho[p] = rho;
        O.p[p] = S.p[i];
        O.c[p] = S.c[i];
        const double m = S.m[i], gap = S.gap[i];
        O.m[p] = m;
        O.gap[p] = gap;
        O.id[p] = S.id[i];
        if (f32) {
            static_cast<float*>(O.ms)[p] = (float)m;
            static_cast<float*>(O.gaps)[p] = (float)gap;
            static_cast<float*>(O.rhos)[p] = (float)rho;
        }
        if (s.flags & FLAG_THERMAL) {
            const double T = s.Tn[i], kap = S.kappa[i], cp = S.cp[i];
            O.T[p] = T;
            O.kappa[p] = kap;
            O.cp[p] = cp;
            if (f32) {
                static_cast<float*>(O.Ts)[p] = (float)T;
                static_cast<float*>(O.kappas)[p] = (float)kap;
                static_cast<float*>(O.cps)[p] = (float)cp;
            }
        }
    }
}

__global__ void k_sort_finish(const SceneDev* __restrict__ scenes, int nscenes, int mode) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= nscenes) return;
    const SceneDev& s = scenes[sc];
    SceneDyn* d = s.dyn;
    d->n_live = s.cell_start[s.ncell];
    d->parity ^= 1;
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
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[dyn.parity];
        double xi[3];
        load_x<D>(S.x, i, xi);
        int c[3];
        cell_of<D>(s, xi, c);
        const int idi = S.id[i];
        int64_t w = offs ? offs[i] : 0;
        int cnt = 0;
        if (pairset == ESPH_PAIRS_FLUID_FLUID) {
            for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i, [&](int j, double, const double*) {
                const int idj = S.id[j];
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

hipError_t launch_continuity(const Bucket& b, hipStream_t st) {
    if (b.ntiles <= 0) return hipSuccess;
    dim3 g(grid_for(b.ntiles)), blk(kBlock);
    ESPH_DISPATCH_K(b.kernel, {
        if (b.dim == 2 && b.prec == ESPH_FP64) hipLaunchKernelGGL((k_continuity<2, K, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else if (b.dim == 2) hipLaunchKernelGGL((k_continuity<2, K, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else if (b.prec == ESPH_FP64) hipLaunchKernelGGL((k_continuity<3, K, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else hipLaunchKernelGGL((k_continuity<3, K, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
    });
    return hipGetLastError();
}

hipError_t launch_forces(const Bucket& b, hipStream_t st) {
    if (b.ntiles <= 0) return hipSuccess;
    dim3 g(grid_for(b.ntiles)), blk(kBlock);
    ESPH_DISPATCH_K(b.kernel, {
        if (b.dim == 2 && b.prec == ESPH_FP64) hipLaunchKernelGGL((k_forces<2, K, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else if (b.dim == 2) hipLaunchKernelGGL((k_forces<2, K, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else if (b.prec == ESPH_FP64) hipLaunchKernelGGL((k_forces<3, K, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else hipLaunchKernelGGL((k_forces<3, K, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
    });
    return hipGetLastError();
}

hipError_t launch_shepard(const Bucket& b, hipStream_t st) {
    if (b.ntiles <= 0) return hipSuccess;
    dim3 g(grid_for(b.ntiles)), blk(kBlock);
    ESPH_DISPATCH_K(b.kernel, {
        if (b.dim == 2) hipLaunchKernelGGL((k_shepard<2, K>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
        else hipLaunchKernelGGL((k_shepard<3, K>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles);
    });
    return hipGetLastError();
}

hipError_t launch_sort(const SortPlan& p, int mode, hipStream_t st) {
    dim3 blk(kBlock);
    hipLaunchKernelGGL(k_sort_clear, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_bin, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_scan1, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_scan2, dim3(grid_for(p.nscenes)), blk, 0, st, p.scenes, p.nscenes);
    hipLaunchKernelGGL(k_sort_scan3, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_place, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles);
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

hipError_t launch_kernel_eval(int kernel, int64_t n, const double* r, double* w, double* dw, double h,
                              double sigma, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), blk(256);
    ESPH_DISPATCH_K(kernel, { hipLaunchKernelGGL(k_kernel_eval<K>, g, blk, 0, st, n, r, w, dw, h, sigma); });
    return hipGetLastError();
}

}  // namespace esph
