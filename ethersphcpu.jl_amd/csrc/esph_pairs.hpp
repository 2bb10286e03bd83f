// esph_pairs.hpp — pair kernels of one step (continuity, forces +
// integration, Shepard) and their launchers, templated on the dimension;
// esph_pairs2.hip and esph_pairs3.hip instantiate them for 2D and 3D.
#pragma once

#include "esph_common.hpp"

namespace esph {
// ---------------------------------------------------------------------------
// continuity! (Interaction/Continuity.jl:8-17), pull form:
//   drho_i = sum_j m_j (v_i - v_j) . gradW_ij,  gradW_ij = W'(r) r_ij / r
// then updateDensity! (SelfAction/UpdateDensity.jl:8-12) and updatePressure!
// (SelfAction/UpdatePressure.jl:8-16).  Writes rho, p, c and the record's
// (rho, p/rho^2).  It walks the candidate lists (every particle within
// Rv = R (1 + skin)) that each sort's last launch builds (k_rebuild; brick
// walks build them here, right after a sort), as all pair passes do until
// the next sort, applying the exact rule d^2 <= R^2 to each candidate.
// Candidates beyond R are weighted by zero rather than branched around: the
// sums equal those over the neighbours alone, and the wavefront stays
// convergent.
// ---------------------------------------------------------------------------
// One particle of the continuity pass.  dry (tuner runs): the step's lists
// as they are, overflowed particles left out, the new density to
// SceneDev::rhon only.  BRICK (FLAG_BRICK scenes, 3D FP32_BULK): a rebuild
// also records the candidates' LDS slots in the brick region (bm), and the
// walk reads the candidates from the staged region (bst; null when the
// region exceeded kBrickCap: from device memory, as the other scenes).
template <int D, int K, typename T, bool SCAN, int DEPTH, bool BRICK = false>
__device__ __forceinline__ void continuity_particle(const SceneDev& s, const SceneDyn& dyn, int i, int dry,
                                                    BrickMap* bm = nullptr, const BrickStage* bst = nullptr) {
    using KQ = KernelQ<T, K>;
    constexpr bool FASTR = sizeof(T) == 4 && K != ESPH_GAUSSIAN;  // pair_r
    const FluidSet& S = s.set[cur_set(dyn)];
    if (S.kind && S.kind[i] != KIND_FLUID) {
        // inlet/outlet buffers keep rho and p (drho is reset,
        // FluidInletOutletWallSolve.jl:91-99) and need no lists
        if (dry) return;
        s.nvnb[i] = 0;
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
        if constexpr (BRICK) {
            // right after a sort: candidate lists within Rv from the cells,
            // with the candidates' LDS slots
            cnt = (dyn.rebuild || !s.verlet) && !dry ? rebuild_lists<D>(s, S, xi, i, bm) : s.nvnb[i];
        } else {
            cnt = s.nvnb[i];  // built after each sort (k_rebuild)
        }
        if (dry && overflowed(s, i)) return;
        if (overflowed(s, i)) {
            // more candidates than the list holds (rare): this thread scans
            // the particle's cells instead (same pair set)
            atomicAdd(&s.dyn->overflow, 1);
            continuity_particle<D, K, T, true, 1>(s, dyn, i, dry);
            return;
        }
    }
    // candidates beyond R (the lists hold every particle within Rv)
    // contribute exactly zero: weighted by 0 instead of branched around
    const double R2 = s.R2;
    T drho = T(0);
    auto load = [&](int j) {
        ARec<T> a;
        load_A<D, T>(S.A, j, a.v, a.m, a.gap);
        return a;
    };
    // one candidate: in = 1 for a neighbour under the exact rule, else 0;
    // ih = 1 / h in the units of d2 and rv (QScale)
    auto pair = [&](T ih, T in, T d2, const T* rv, const ARec<T>& a) {
        const T* vj = a.v;
        const T m_j = a.m;
        T r, ri;
        pair_r<FASTR>(d2, r, ri);
        T W, dW;
        KQ::wdw(r * ih, kA, kB, W, dW);
        // normalize(0) is NaN and the reference zeroes it (Neighbour.jl:54-59);
        // with FASTR W'(0) ri is already 0 (pair_r)
        const T sg = FASTR ? dW * ri : (d2 == T(0) ? T(0) : dW * ri);
        T dot = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) dot += (vi[k] - vj[k]) * rv[k];
        drho += m_j * sg * dot * in;
    };
    if constexpr (sizeof(T) == 4 && !SCAN) {
        // FP32_BULK: geometry from the quantized positions (FP64 rule near R)
        const T ihq = QScale<T>(s, true).inv_h;
        auto qbody = [&](int, bool in, float d2, const float* rv, const ARec<T>& a, int32_t) {
            pair(ihq, in ? T(1) : T(0), d2, rv, a);
        };
        if constexpr (BRICK) {
            if (bst) {
                for_pairs_brick<DEPTH>(s, S, xi, i, cnt, *bst,
                                       [&](bool in, float d2, const float* rv, const ARec<float>& a, int32_t w) {
                    qbody(0, in, d2, rv, a, w);
                });
            } else {
                // a region beyond the staging (rare): one gather at a time,
                // which keeps the kernel's registers to the brick walk's
                for_pairs_q<D, 1, false>(s, S, xi, i, s.vnbr, cnt, load, qbody);
            }
        } else {
            for_pairs_q<D, DEPTH>(s, S, xi, i, s.vnbr, cnt, load, qbody);
        }
    } else {
        for_pairs_d<D, SCAN, DEPTH, true, D == 3, sizeof(T) == 4>(s, s.cell_start, S.x, xi, i, i, s.vnbr, cnt, load,
                                    [&](int, double d2, const double* rv, const ARec<T>& a) {
            T rvt[3] = {(T)rv[0], (T)rv[1], D == 3 ? (T)rv[2] : T(0)};
            pair(inv_h, d2 <= R2 ? T(1) : T(0), (T)d2, rvt, a);
        });
    }
    {
#pragma clang fp contract(off)
        double rho = S.rho[i];
        rho += (double)drho * s.dt;
        if (dry) {
            s.rhon[i] = rho;
            return;
        }
        const double rr = rho / s.rho0;
        const double p = s.b * (ipow(rr, s.gamma) - 1.0) + s.p0;
        const double cc = s.c0 * sqrt(ipow(rr, s.gamma - 1));
        S.rho[i] = rho;
        S.p[i] = p;
        S.c[i] = cc;
        T* B = static_cast<T*>(S.B) + 4 * (size_t)i;
        B[0] = (T)rho;
        B[1] = (T)(p / (rho * rho));
        // the density the force pass's two-record walk reads (FluidSet::xq)
        if (D == 3 && S.xq) reinterpret_cast<float*>(S.xq)[4 * (size_t)i + 3] = (float)(rho - s.rho0);
    }
}

template <int D, int K, typename T, bool SCAN, int DEPTH, int BS>
__global__ __launch_bounds__(SCAN ? kBlock : BS) void k_continuity(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles, int ntiles, int dry) {
    constexpr int SUB = SCAN ? 1 : kBlock / BS;  // blocks per tile
    const int nv = ntiles * SUB;
    for (int vb = blockIdx.x; vb < nv; vb += gridDim.x) {
        const int v = tile_index(vb, nv);
        const Tile tile = tiles[v / SUB];
        const int first = tile.begin + (v % SUB) * BS;
        const SceneDev& s = scenes[tile.scene];
        if (!SCAN && (s.flags & FLAG_BRICK)) continue;  // k_continuity_brick
        const SceneDyn dyn = *s.dyn;
        for_work<SCAN>(s, dyn, tile, first, [&](int i) { continuity_particle<D, K, T, SCAN, DEPTH>(s, dyn, i, dry); });
    }
}

// Bricks of the FLAG_BRICK scenes listed in bt (one entry per scene), XCD by
// XCD: the scene's bricks (blist, in brick order) are cut into eight
// contiguous ranges, and workgroup g (on XCD g % 8) takes entries g / 8,
// g / 8 + gridDim.x / 8, ... of range g % 8, so each XCD's L2 serves one
// compact part of the scene.  For each: the region's tables, its staging
// (unless it exceeds kBrickCap), then f(s, dyn, spill) with every thread.
template <typename F>
__device__ __forceinline__ void for_each_brick(const SceneDev* __restrict__ scenes, const Tile* __restrict__ bt,
                                               int nbt, BrickTables& tb, int4* lq, float4* la, F&& f) {
    for (int e = 0; e < nbt; ++e) {
        const SceneDev& s = scenes[bt[e].scene];
        if (!(s.flags & FLAG_BRICK)) continue;
        const SceneDyn dyn = *s.dyn;
        const FluidSet& S = s.set[cur_set(dyn)];
        const int nb = dyn.nbrick;
        const int per = (nb + 7) >> 3;
        const int xg = (int)(blockIdx.x & 7u), g8 = (int)(gridDim.x >> 3);
        for (int k = (int)(blockIdx.x >> 3); k < per; k += g8) {
            const int q = xg * per + k;
            if (q >= nb) break;
            brick_tables(s, s.blist[q], tb);
            const bool spill = tb.rb[tb.nr] > kBrickCap;
            if (!spill)
                brick_stage(S, tb, lq, la);
            else if (threadIdx.x == 0)
                atomicAdd(&s.dyn->brick_spill, 1);
            f(s, dyn, spill);
            __syncthreads();  // the next brick rewrites the tables and the staging
        }
    }
}

// The continuity pass of FLAG_BRICK scenes (3D FP32_BULK), brick by brick.
template <int K, int DEPTH>
__global__ __launch_bounds__(kBrickThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_continuity_brick(
    const SceneDev* __restrict__ scenes, const Tile* __restrict__ bt, int nbt, int dry) {
    __shared__ int4 lq[kBrickCap];
    __shared__ float4 la[kBrickCap];
    __shared__ BrickTables tb;
    for_each_brick(scenes, bt, nbt, tb, lq, la, [&](const SceneDev& s, const SceneDyn& dyn, bool spill) {
        BrickMap bm{&tb, 0};
        const BrickStage st{lq, la};
        for_brick_particles(tb, [&](int i) {
            continuity_particle<3, K, float, false, DEPTH, true>(s, dyn, i, dry, &bm, spill ? nullptr : &st);
        });
    });
}

// ---------------------------------------------------------------------------
// Momentum (and heat) pass + integration, pull form over the candidate lists.
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
// The new state goes to the other per-step copy N in the current particle
// order: x, v, T, the quantized position and whole A and B records (B.xy keep
// this step's density and pressure until the next continuity pass).
// ---------------------------------------------------------------------------
// The float positions (FluidSet::xf) are read only by the list rebuild,
// which follows a sort, and the sort writes them for its new order; the
// per-step copies therefore carry the quantized positions alone.
// (rho: the density the next continuity pass starts from, for xq.w)
template <int D>
__device__ __forceinline__ void store_next_q(const SceneDev& s, const FluidSet& N, int i, const double* x,
                                             double rho) {
    if (N.xq) store_xq<D>(s, N, i, x, (float)(rho - s.rho0));
}

// Buffer particles of the inlet/outlet solver: dv = 0, then updatePosition!
// (FluidInletOutletWallSolve.jl:170-186): x += v dt - 0 dt^2 / 2; their
// records and temperature carry over unchanged.
template <int D, typename T>
__device__ __forceinline__ void buffer_move(const SceneDev& s, const FluidSet& S, const FluidSet& N, int i) {
#pragma clang fp contract(off)
    using V = typename Vec4<T>::type;
    const double dt = s.dt;
    double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double v = S.v[i * D + k];
        N.v[i * D + k] = v;
        x[k] = S.x[i * D + k] + (v * dt - 0.0 * (dt * dt) / 2.0);
        N.x[i * D + k] = x[k];
    }
    store_next_q<D>(s, N, i, x, S.rho[i]);
    static_cast<V*>(N.A)[i] = static_cast<const V*>(S.A)[i];
    static_cast<V*>(N.B)[i] = static_cast<const V*>(S.B)[i];
    if (N.T) N.T[i] = S.T[i];
}

// ---------------------------------------------------------------------------
// Wall terms of one fluid particle and one wall candidate: wallForce!
// (WallForce.jl:8-30, Rogers & Dalrymple 2008), pressureForce! wall
// (PressureForce.jl:23-35) and viscosityForce! wall (ViscosityForce.jl:27-46),
// added to acc; r, W'(r) and r^2 + 0.01 h^2 are returned for the thermostat
// term.  in: the exact rule d^2 <= R^2 (candidates beyond it add nothing).
// ---------------------------------------------------------------------------
template <typename T>
struct WallConst {
    T inv_h, kA, kB, eta2, nu0;
    T P0;           // p0 / rho0^2, the wall's p / rho^2
    bool pressure;  // FLAG_WALL_PRESSURE
    double R2;
};

// Windows of 64 wall candidates that wall_walk_shared gathers at a time
// (two: four take more than the 128 registers of the float force pass).
constexpr int kWallWindows = 2;
// Wavefronts with wall candidates in more lanes than this walk them lane by
// lane (per-wavefront choice of the force pass's wall walk).
constexpr int kWallShareLanes = 16;

// The particle's side of the wall terms.
template <typename T>
struct WallOwner {
    double x[3];
    T v[3];
    T c, pr2, inv_rho, m, inv_wg;  // sound speed, p / rho^2, 1 / rho, mass, 1 / W(gap)
    int32_t ex, i;                 // first position in the wavefront's sequence (wall_walk_shared), particle
};

// Wall records (SceneDev::wA, wB) of candidate j.
template <typename T>
__device__ __forceinline__ void load_wall(const SceneDev& s, int j, typename Vec4<T>::type& wa,
                                          typename Vec4<T>::type& wb) {
    using V = typename Vec4<T>::type;
    if constexpr (sizeof(V) == 16) {
        wa = ldrec<V>(s.wA, j);
        wb = ldrec<V>(s.wB, j);
    } else {
        wa = static_cast<const V*>(s.wA)[j];
        wb = static_cast<const V*>(s.wB)[j];
    }
}

// Normal, velocity, gap, temperature and conductivity of a wall record pair
// (kwv: the 3D conductivity, kept apart).
template <int D, typename T, typename V>
__device__ __forceinline__ void unpack_wall(const V& wa, const V& wb, T kwv, T* n, T* vw, T& gap_j, T& Tw, T& kw) {
    n[0] = wa.x;
    n[1] = wa.y;
    vw[0] = wb.x;
    vw[1] = wb.y;
    if (D == 3) {
        n[2] = wa.z;
        gap_j = wa.w;
        vw[2] = wb.z;
        Tw = wb.w;
        kw = kwv;
    } else {
        n[2] = T(0);
        vw[2] = T(0);
        gap_j = wa.z;
        Tw = wa.w;
        kw = wb.z;
    }
}

template <int D, int K, typename T, bool FASTR>
__device__ __forceinline__ void wall_pair(const WallConst<T>& w, const WallOwner<T>& o, bool in, double d2,
                                          const double* rv, const T* n, const T* vw, T gap_j, T* acc, T& r, T& dW,
                                          T& r2e) {
    using KQ = KernelQ<T, K>;
    T ri;
    pair_r<FASTR>((T)d2, r, ri);
    T W;
    KQ::wdw(r * w.inv_h, w.kA, w.kB, W, dW);
    // with FASTR W'(0) ri is already 0 (pair_r)
    const T sg = FASTR ? dW * ri : (d2 == 0.0 ? T(0) : dW * ri);
    T vv[3];
    T vn = T(0), rn = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        vv[k] = o.v[k] - vw[k];
        vn += vv[k] * n[k];
        rn += (T)rv[k] * n[k];
    }
    // wallForce! (Rogers & Dalrymple 2008)
    {
        const T psi = fabs(rn);
        const T a = r * r - psi * psi;
        const T xi_ = psqrt(a > T(0) ? a : T(0));
        const T q = pdiv(psi, gap_j);
        if (in && !(q > T(1) || xi_ > gap_j)) {
            const T p_xi = fabs((T(1) + cos_pi_ratio(xi_, gap_j)) * T(0.5));
            const T beta = vn > T(0) ? T(0) : T(1);
            const T wf =
                (T(0.01) * (o.c * o.c) + beta * o.c * fabs(vn)) * w.inv_h * fabs(T(1) - q) * prsqrt(q) * p_xi;
#pragma unroll
            for (int k = 0; k < D; ++k) acc[k] += wf * n[k];
        }
    }
    r2e = r * r + w.eta2;
    T vf = T(2) * w.nu0 * r * dW * o.inv_rho * prcp(r2e);
    T ps = T(0);
    if (w.pressure) {
        T P = o.pr2 + w.P0;
        P += fabs(P) * T(0.01) * W * o.inv_wg;
        ps = -P * sg;
    }
    vf = in ? vf : T(0);
    ps = in ? ps : T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) acc[k] += o.m * (ps * (T)rv[k] + vf * (vv[k] - vn * n[k]));
}

// The wall walk of a wavefront's particles with their candidates shared out
// over all its lanes (called by every lane together; cnt: the lane's wall
// candidates, 0 for lanes without a particle to advance).  A few particles
// of a wavefront lie near a wall -- along a lattice line, the first three
// next to an end wall -- and walking their lists in their own lanes (20 to
// 60 candidates each) leaves the other lanes idle: on the 1M-particle 3D dam
// break a tenth of the wall walk's lane slots did work.  Here the lists of
// lanes 0..63, in lane and list order, form one sequence of positions, cut
// into windows of 64: in window b lane p takes position b + p, finds whose
// it is (the last particle starting at or before it; own[] marks the starts
// inside the window) and that particle's inputs (slot[], the wavefront's LDS
// copy of every lane's WallOwner), adds up the candidate's terms, and a
// segmented sum over the window gives each particle the total of its
// positions there, read from its last one.  kWallWindows windows at a time:
// their list entries, then their wall records, are gathered together, so a
// group costs two memory round trips (the walk's time is those waits: the
// pass is latency-bound, and a wavefront near a wall waits for its walls
// after its fluid walk).  slot and own: the workgroup's areas, 64 and
// 64 kWallWindows entries per wavefront.
template <int D, int K, typename T, bool FASTR>
__device__ __forceinline__ void wall_walk_shared(const SceneDev& s, const WallConst<T>& wc, int cnt,
                                                 const WallOwner<T>& me, WallOwner<T>* slot, int* own, T* dv) {
    using V = typename Vec4<T>::type;
    constexpr int G = kWallWindows;
    const int lane = (int)(threadIdx.x & 63);
    // this wavefront's entries (the offset is uniform: a scalar register)
    const unsigned w0 = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    slot += w0;
    own += w0 * G;
    const int incl = wave_segmented_sum(cnt, 0);  // inclusive prefix sum over the lanes
    const int total = __builtin_amdgcn_readlane(incl, 63);
    if (total == 0) return;
    const int ex = incl - cnt;
    {
        WallOwner<T> o = me;
        o.ex = ex;
        slot[lane] = o;
    }
    const size_t cap = (size_t)s.cap;
    int carry = 0;  // whose the last position of the previous window was
    for (int b0 = 0; b0 < total; b0 += 64 * G) {
        // the starts of the group's windows: own[64 u + p] for position b0 + 64 u + p
        wave_lds_fence();
#pragma unroll
        for (int u = 0; u < G; ++u) own[64 * u + lane] = -1;
        const int rel = ex - b0;
        if (cnt > 0 && rel >= 0 && rel < 64 * G) own[rel] = lane;
        wave_lds_fence();
        int head[G], o[G], j[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const uint64_t starts = __ballot(own[64 * u + lane] >= 0);
            const uint64_t upto = starts & (lane == 63 ? ~0ull : (2ull << lane) - 1ull);
            // the window lane where this position's particle starts (0: before the window)
            head[u] = upto ? 63 - __clzll((long long)upto) : 0;
            const int m = own[64 * u + head[u]];
            o[u] = upto ? m : carry;
            carry = __builtin_amdgcn_readlane(o[u], 63);
        }
        // list entries, then wall records, of every window of the group
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int g = b0 + 64 * u + lane;
            j[u] = -1;
            if (g < total) {
                const WallOwner<T>& ow = slot[o[u]];
                j[u] = s.wnbr[list_at(cap, g - ow.ex, ow.i)];
            }
        }
        double xj[G][3];
        V wa[G], wb[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (j[u] >= 0) {
                load_x<D>(s.wx, j[u], xj[u]);
                load_wall<T>(s, j[u], wa[u], wb[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int b = b0 + 64 * u;
            if (b >= total) break;
            T acc[3] = {T(0), T(0), T(0)};
            if (j[u] >= 0) {
                const WallOwner<T> ow = slot[o[u]];
                double rv[3];
                const double d2 = dist2_exact<D>(ow.x, xj[u], rv);
                T n[3], vw[3], gap_j, Tw, kw;
                unpack_wall<D, T>(wa[u], wb[u], T(0), n, vw, gap_j, Tw, kw);
                T r, dW, r2e;
                wall_pair<D, K, T, FASTR>(wc, ow, d2 <= wc.R2, d2, rv, n, vw, gap_j, acc, r, dW, r2e);
            }
#pragma unroll
            for (int k = 0; k < D; ++k) acc[k] = wave_segmented_sum(acc[k], head[u]);
            // this lane's particle: the sum at its last position in the window
            const bool mine = cnt > 0 && ex < b + 64 && ex + cnt > b;
            const int last = mine ? min(ex + cnt - b, 64) - 1 : lane;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const T t = __shfl(acc[k], last, 64);
                if (mine) dv[k] += t;
            }
        }
    }
}

// Returns the particle's squared displacement from its position when the
// candidate lists were built (0 when the lists are rebuilt every step).
// full_records: the next copy's records were last written in another
// particle order (the first step after a sort), so B is written whole;
// otherwise B.zw there still hold this order's gap (and, without heat, T) and
// B.xy is rewritten by the next continuity pass before any read.
// BRICK (FLAG_BRICK scenes): the fluid walk reads the candidates from the
// staged brick region bst (null when the region exceeded kBrickCap: from
// device memory, as the other scenes).
// WAVE: called by every lane of the wavefront together (act: the lane has a
// particle i to advance; the others take part in the shared wall walk
// only), with the workgroup's LDS areas for wall_walk_shared.
template <int D, int K, typename T, bool THERMAL, bool SCAN, int DEPTH, bool EOSW, bool BRICK = false,
          bool WAVE = false>
__device__ __forceinline__ double forces_body(const SceneDev& s, const FluidSet& S, const FluidSet& N, int i,
                                              bool full_records, const BrickStage* bst = nullptr, bool act = true,
                                              WallOwner<T>* wslot = nullptr, int* wown = nullptr) {
    using KQ = KernelQ<T, K>;
    constexpr bool FASTR = sizeof(T) == 4 && K != ESPH_GAUSSIAN;  // pair_r
    static_assert(!WAVE || (!THERMAL && !SCAN && !BRICK), "wavefront form: plain list walks");
    bool live = act;
    if (live && S.kind && S.kind[i] != KIND_FLUID) {
        buffer_move<D, T>(s, S, N, i);
        live = false;
    }
    if constexpr (!WAVE) {
        if (!live) return 0.0;
    } else if (!live) {
        i = 0;  // in range (capacity >= 1); nothing of it is walked or written
    }
    // candidates of the fluid walk
    const int nv = live ? s.nvnb[i] : 0;
    // (wavefront form) wall candidates, loaded ahead: the wall walk after the
    // fluid walk starts from them
    int cw = 0;
    if constexpr (WAVE) {
        if (s.nw > 0 && live) cw = s.nnbw[i];
    }
    const T inv_h = (T)s.inv_h, kA = (T)s.kA, kB = (T)s.kB;
    const T eta2 = (T)(0.01 * (s.h * s.h));
    const T nu0 = (T)s.nu0;
    double xi[3];
    load_x<D>(S.x, i, xi);
    Rec<T> me;
    load_rec<D, T, THERMAL>(S, i, me);
    const T rho_i = me.rho, gap_i = me.gap, m_i = me.m, T_i = me.Tt, kap_i = me.kap;
    T pr2_i = me.pr2;  // the EOS form's own value in the two-record walk below
    T W_gap, dW_gap;
    KQ::wdw(gap_i * inv_h, kA, kB, W_gap, dW_gap);
    const T inv_Wgap_i = prcp(W_gap);
    const T c01_i = T(0.01) * inv_Wgap_i;
    T dv[3] = {T(0), T(0), T(0)};
    T dT = T(0);
    // candidates beyond R contribute exactly zero (weighted, not branched)
    const double R2 = s.R2;
    // gather groups: the positions with the A records, then the B records
    // (RecLoaderAB); the float form in 3D takes the positions quantized,
    // like the continuity pass.  The wall walk below: binary64, the
    // positions then the records; float, both together
    constexpr bool SPLIT = sizeof(T) == 8;
    // one candidate: in = 1 for a neighbour under the exact rule, else 0;
    // zero: coincident positions (normalize(0), Neighbour.jl:54-59);
    // uniform: the scene's gaps are all equal, so W((gap_i + gap_j) / 2) = W(gap_i)
    // g: the length constants in the units of d2 and rv (QScale)
    auto ff = [&](T in, bool zero, T d2, const T* rv, const Rec<T>& o, auto uniform, const QScale<T>& g) {
        T r, ri;
        pair_r<FASTR>(d2, r, ri);
        T W, dW;
        KQ::wdw(r * g.inv_h, kA, kB, W, dW);
        // with FASTR W'(0) ri is already 0 (pair_r)
        const T sg = FASTR ? dW * ri : (zero ? T(0) : dW * ri);
        // pressure, with the 1% correction W(r) / W((gap_i + gap_j) / 2)
        T P = pr2_i + o.pr2;
        if (sizeof(T) == 4 && decltype(uniform)::value) {
            // one gap: 0.01 / W(gap_i) is the particle's constant
            P += fabs(P) * W * c01_i;
        } else {
            T inv_Wg = inv_Wgap_i;
            if (!decltype(uniform)::value && o.gap != gap_i) {
                T w2, d2w;
                KQ::wdw((gap_i + o.gap) * T(0.5) * inv_h, kA, kB, w2, d2w);
                inv_Wg = prcp(w2);
            }
            P += fabs(P) * T(0.01) * W * inv_Wg;
        }
        // viscosity
        const T r2e = r * r + g.eta2;
        const T vf = T(4) * nu0 * g.len * r * dW * prcp((rho_i + o.rho) * r2e);
        const T ps = -P * sg;
        // in is 0 or 1, so it weights the pair exactly through the mass
        const T mw = o.m * in;
        if (sizeof(T) == 4) {
            // the two terms as two multiply-adds per component
            const T a = mw * ps, b = mw * vf;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                dv[k] += a * rv[k];
                dv[k] += b * (me.v[k] - o.v[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) dv[k] += mw * (ps * rv[k] + vf * (me.v[k] - o.v[k]));
        }
        if (THERMAL)
            dT += T(4) * g.len * kap_i * o.kap * r * dW * (T_i - o.Tt) * o.m * prcp((kap_i + o.kap) * o.rho * r2e) * in;
    };
    const QScale<T> gq(s, true), gm(s, false);
    if constexpr (sizeof(T) == 4 && !SCAN && D == 3 && !THERMAL && EOSW) {
        {
            // two records per candidate: the quantized position with
            // rho - rho0 (FluidSet::xq), and the velocity with the mass (A);
            // p / rho^2 from the equation of state, for the candidates and,
            // alike, for this particle
            float rho_e;
            const bool lin = s.gamma == 1 && s.p0f == 0.0f;
            const bool tait = !lin && s.gamma == kTaitGamma;
            const float d_i = reinterpret_cast<const float*>(S.xq)[4 * (size_t)i + 3];
            pr2_i = lin ? eos_pr2<true>(s, d_i, rho_e) : eos_pr2<false>(s, d_i, rho_e);
            auto loadA = [&](int j) {
                ARec<T> a;
                load_A<D, T>(S.A, j, a.v, a.m, a.gap);
                return a;
            };
            // the walk in three forms, chosen once per particle: the linear
            // equation of state (gamma 1, p0 0: no sum per candidate), gamma
            // 7 (the sum unrolled) or any gamma (a loop per candidate)
            auto walk = [&](auto form) {
                constexpr int FORM = decltype(form)::value;  // 1 linear, kTaitGamma, 0 any
                auto body = [&](int, bool in, float d2, const float* rv, const ARec<T>& a, int32_t w) {
                    Rec<T> o;
                    o.v[0] = a.v[0];
                    o.v[1] = a.v[1];
                    o.v[2] = a.v[2];
                    o.m = a.m;
                    o.pr2 = eos_pr2<FORM == 1, FORM == kTaitGamma ? kTaitGamma : 0>(s, __int_as_float(w), o.rho);
                    ff(in ? T(1) : T(0), d2 == 0.0f, d2, rv, o, std::true_type{}, gq);
                };
                if constexpr (BRICK) {
                    if (bst)
                        for_pairs_brick<DEPTH>(s, S, xi, i, nv, *bst,
                                               [&](bool in, float d2, const float* rv, const ARec<float>& a, int32_t w) {
                            body(0, in, d2, rv, a, w);
                        });
                    else  // a region beyond the staging (rare): one gather at a time
                        for_pairs_q<D, 1, false>(s, S, xi, i, s.vnbr, nv, loadA, body);
                } else {
                    for_pairs_q<D, DEPTH, false>(s, S, xi, i, s.vnbr, nv, loadA, body);
                }
            };
            if (lin)
                walk(std::integral_constant<int, 1>{});
            else if (tait)
                walk(std::integral_constant<int, kTaitGamma>{});
            else
                walk(std::integral_constant<int, 0>{});
        }
    } else if constexpr (sizeof(T) == 4 && !SCAN && D == 3) {
        // FP32_BULK in 3D: geometry from the quantized positions (FP64 rule
        // near R).  The 2D ensemble measured 4% slower with them (1.90
        // against 1.83 ms per batched step) and keeps the binary64 ones.
        for_pairs_q<D, DEPTH, false>(s, S, xi, i, s.vnbr, nv, RecLoaderAB<D, T, THERMAL>{S},
                                     [&](int, bool in, float d2, const float* rv, const Rec<T>& o, int32_t) {
            ff(in ? T(1) : T(0), d2 == 0.0f, d2, rv, o, std::false_type{}, gq);
        });
    } else {
        for_pairs_d<D, SCAN, DEPTH, false, false, sizeof(T) == 4>(s, s.cell_start, S.x, xi, i, i, s.vnbr, nv,
                                           RecLoaderAB<D, T, THERMAL>{S},
                                           [&](int, double d2, const double* rv, const Rec<T>& o) {
            const T rvt[3] = {(T)rv[0], (T)rv[1], D == 3 ? (T)rv[2] : T(0)};
            ff(d2 <= R2 ? T(1) : T(0), d2 == 0.0, (T)d2, rvt, o, std::false_type{}, gm);
        });
    }
    if (THERMAL) dT *= prcp(rho_i * (T)S.cp[i]);
    if (s.nw > 0) {
        using V = typename Vec4<T>::type;
        WallConst<T> wc;
        wc.inv_h = inv_h;
        wc.kA = kA;
        wc.kB = kB;
        wc.eta2 = eta2;
        wc.nu0 = nu0;
        wc.P0 = (T)(s.p0 / (s.rho0 * s.rho0));
        wc.pressure = (s.flags & FLAG_WALL_PRESSURE) != 0;
        wc.R2 = s.R2;
        WallOwner<T> own;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            own.x[k] = xi[k];
            own.v[k] = me.v[k];
        }
        own.c = (T)S.c[i];
        own.pr2 = pr2_i;
        own.inv_rho = prcp(rho_i);
        own.m = m_i;
        own.inv_wg = inv_Wgap_i;
        own.ex = 0;
        own.i = i;
        bool lanes = true;
        if constexpr (WAVE) {
            // shared out when few lanes have wall candidates (next to an end
            // wall); a wavefront along a floor or side wall, with candidates
            // in most lanes, walks them lane by lane, its gathers four deep
            lanes = (s.flags & FLAG_WALL_LANES) != 0 || __popcll(__ballot(cw > 0)) > kWallShareLanes;
        }
        if (!lanes) {
            if constexpr (WAVE) wall_walk_shared<D, K, T, FASTR>(s, wc, cw, own, wslot, wown, dv);
        } else if (live) {
            const T th_w = (T)(1.0 / (s.cp * s.rho0)) * own.inv_rho * m_i;
            T dTw = T(0);
            struct WRec {
                V wa, wb;
                T kw;
            };
            auto wload = [&](int j) {
                WRec w;
                load_wall<T>(s, j, w.wa, w.wb);
                w.kw = (D == 3 && THERMAL) ? ld<T>(s.wkap, j) : T(0);
                return w;
            };
            // (brick walks: at most two wall gathers in flight, which keeps the
            // four-deep fluid walk within the registers of four waves per SIMD)
            constexpr int WDEPTH = BRICK && DEPTH > 2 ? 2 : DEPTH;
            for_pairs_d<D, SCAN, WDEPTH, SPLIT, false, sizeof(T) == 4>(s, s.wcell_start, s.wx, xi, i, -1, s.wnbr,
                                                                    WAVE ? cw : s.nnbw[i], wload,
                                                [&](int j, double d2, const double* rv, const WRec& w) {
                // candidates within Rv: the exact rule selects the terms (no branch)
                const bool in = d2 <= wc.R2;
                T n[3], vw[3], gap_j, Tw, kw;
                unpack_wall<D, T>(w.wa, w.wb, w.kw, n, vw, gap_j, Tw, kw);
                T r, dW, r2e;
                wall_pair<D, K, T, FASTR>(wc, own, in, d2, rv, n, vw, gap_j, dv, r, dW, r2e);
                if (THERMAL && in && s.wkind[j] == ESPH_WALL_THERMOSTATIC)
                    dTw += (kap_i + kw) * r * dW * (T_i - Tw) * prcp(r2e);
            });
            if (THERMAL) dT += dTw * th_w;
        }
    }
    if constexpr (WAVE) {
        if (!live) return 0.0;
    }
    {
#pragma clang fp contract(off)
        const double dt = s.dt;
        double bf[3];
        double Tn = 0.0;
        if (THERMAL) {
            Tn = S.T[i] + dt * (double)dT;
            N.T[i] = Tn;
#pragma unroll
            for (int k = 0; k < D; ++k) bf[k] = -s.g[k] * s.beta * (Tn - s.t0);
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) bf[k] = s.g[k];
        }
        double disp2 = 0.0;
        double xnew[3] = {0.0, 0.0, 0.0}, vnew[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double v = S.v[i * D + k];
            v += bf[k] * dt;
            const double dvk = (double)dv[k];
            v += dvk * dt;
            N.v[i * D + k] = v;
            vnew[k] = v;
            const double xk = xi[k] + (v * dt - dvk * (dt * dt) / 2.0);
            N.x[i * D + k] = xk;
            xnew[k] = xk;
            if (s.verlet) {
                const double e = xk - s.xref[i * D + k];
                disp2 += e * e;
            }
        }
        store_next_q<D>(s, N, i, xnew, (double)rho_i);
        // records of the new state; float records hold T - t0 (see k_sort_scatter)
        using V = typename Vec4<T>::type;
        const T Tr = THERMAL ? (T)(sizeof(T) == 4 ? Tn - s.t0 : Tn) : me.Tt;
        V a, b;
        a.x = (T)vnew[0];
        a.y = (T)vnew[1];
        b.x = rho_i;
        b.y = pr2_i;
        if (D == 3) {
            a.z = (T)vnew[2];
            a.w = m_i;
            b.z = gap_i;
            b.w = Tr;
        } else {
            a.z = m_i;
            a.w = gap_i;
            b.z = Tr;
            b.w = kap_i;
        }
        static_cast<V*>(N.A)[i] = a;
        if (THERMAL || full_records) static_cast<V*>(N.B)[i] = b;
        return disp2;
    }
}

// Wavefront maximum of the squared displacements (rounded up to float;
// float bits of non-negative values and NaN order as unsigned integers),
// then one atomic per wavefront only when it raises the scene's running
// maximum: a contended atomic per wavefront would serialise the whole pass
// on one address.  Every lane of the wavefront calls it.
__device__ __forceinline__ void record_disp(const SceneDev& s, double dmax) {
    uint32_t bits = __float_as_uint(__double2float_ru(dmax));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t other = __shfl_xor(bits, o, 64);
        bits = other > bits ? other : bits;
    }
    if ((threadIdx.x & 63) == 0 && bits != 0 &&
        bits > __hip_atomic_load(&s.dyn->disp2_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&s.dyn->disp2_bits, bits);
}

// Float forces without heat are held to 4 waves per SIMD (128 VGPRs): the
// gather pipeline otherwise takes a few registers more and falls to 3.
// EOSW: the instance for scenes with FLAG_EOS_WALK (3D FP32_BULK without
// heat), whose fluid walk gathers two records per candidate; the other
// instances skip those scenes and it skips the rest.
template <int D, int K, typename T, bool THERMAL, bool SCAN, int DEPTH, int BS, bool EOSW = false>
__global__ __launch_bounds__(SCAN ? kBlock : BS)
__attribute__((amdgpu_waves_per_eu(EOSW && !SCAN ? kEosWalkWaves
                                   : std::is_same<T, float>::value && !THERMAL && !SCAN ? 4 : 1))) void k_forces(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, int ntiles) {
    constexpr int SUB = SCAN ? 1 : kBlock / BS;  // blocks per tile
    // the wavefront form of the body (3D list walks without heat): LDS for
    // the shared wall walk, 64 entries per wavefront.  2D scenes walk their
    // walls lane by lane: the wavefront form takes the 2D float pass from 94
    // registers (five waves per SIMD) to 106 (four), and on the 256-scene 2D
    // ensemble the shared walk gained nothing (1.678 against 1.685 ms per
    // batched step)
    constexpr bool WAVE = !SCAN && !THERMAL && D == 3;
    __shared__ WallOwner<T> wslot[WAVE ? BS : 1];
    __shared__ int wown[WAVE ? BS * kWallWindows : 1];
    const int nv = ntiles * SUB;
    for (int vb = blockIdx.x; vb < nv; vb += gridDim.x) {
        const int v = tile_index(vb, nv);
        const Tile tile = tiles[v / SUB];
        const int first = tile.begin + (v % SUB) * BS;
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (((s.flags & FLAG_THERMAL) != 0) != THERMAL) continue;
        if (((s.flags & FLAG_EOS_WALK) != 0) != EOSW) continue;
        if (!SCAN && (s.flags & FLAG_BRICK)) continue;  // k_forces_brick
        double dmax = 0.0;
        if constexpr (WAVE) {
            // every lane of the wavefront runs the body together (the wall
            // walk shares candidates out over the lanes); a particle whose
            // candidates overflowed its list (rare) then scans its cells on
            // its own
            const int i = first + (int)threadIdx.x;
            const bool act = i < dyn.n_live;
            const FluidSet &S = s.set[cur_set(dyn)], &N = s.set[next_set(dyn)];
            const bool ovf = act && overflowed(s, i);
            dmax = forces_body<D, K, T, THERMAL, false, DEPTH, EOSW, false, true>(
                s, S, N, i, dyn.rebuild != 0, nullptr, act && !ovf, wslot, wown);
            if (ovf) dmax = forces_body<D, K, T, THERMAL, true, 1, EOSW>(s, S, N, i, dyn.rebuild != 0);
        } else {
            for_work<SCAN>(s, dyn, tile, first, [&](int i) {
                const FluidSet &S = s.set[cur_set(dyn)], &N = s.set[next_set(dyn)];
                // a particle whose candidates overflowed its list (rare) scans its cells
                const double d2 = !SCAN && overflowed(s, i)
                                      ? forces_body<D, K, T, THERMAL, true, 1, EOSW>(s, S, N, i, dyn.rebuild != 0)
                                      : forces_body<D, K, T, THERMAL, SCAN, DEPTH, EOSW>(s, S, N, i,
                                                                                       dyn.rebuild != 0);
                dmax = d2 > dmax || d2 != d2 ? d2 : dmax;  // NaN propagates
            });
        }
        if (s.verlet) record_disp(s, dmax);
    }
}

// The force pass of FLAG_BRICK scenes (3D FP32_BULK, FLAG_EOS_WALK), brick
// by brick.
template <int K, int DEPTH>
__global__ __launch_bounds__(kBrickThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_forces_brick(
    const SceneDev* __restrict__ scenes, const Tile* __restrict__ bt, int nbt) {
    __shared__ int4 lq[kBrickCap];
    __shared__ float4 la[kBrickCap];
    __shared__ BrickTables tb;
    for_each_brick(scenes, bt, nbt, tb, lq, la, [&](const SceneDev& s, const SceneDyn& dyn, bool spill) {
        const BrickStage st{lq, la};
        double dmax = 0.0;
        for_brick_particles(tb, [&](int i) {
            const FluidSet &S = s.set[cur_set(dyn)], &N = s.set[next_set(dyn)];
            const double d2 = overflowed(s, i)
                                  ? forces_body<3, K, float, false, true, 1, true>(s, S, N, i, dyn.rebuild != 0)
                                  : forces_body<3, K, float, false, false, DEPTH, true, true>(
                                        s, S, N, i, dyn.rebuild != 0, spill ? nullptr : &st);
            dmax = d2 > dmax || d2 != d2 ? d2 : dmax;  // NaN propagates
        });
        if (s.verlet) record_disp(s, dmax);
    });
}

// ---------------------------------------------------------------------------
// reconstructScalar!(fluid, :rho_, FF pairs, kernel) (GroupAction/
// ReconstructValue.jl:22-57): Shepard filter over the step-start pairs with
// the post-continuity density, in binary64 in every form.
// ---------------------------------------------------------------------------
// One particle of the Shepard pass (k_shepard).
template <int D, int K, bool SCAN, int DEPTH>
__device__ __forceinline__ void shepard_particle(const SceneDev& s, const SceneDyn& dyn, int i) {
    using KQ = KernelQ<double, K>;
    const FluidSet& S = s.set[cur_set(dyn)];
    const int8_t* kind = S.kind;
    if (kind && kind[i] != KIND_FLUID) return;
    double xi[3];
    load_x<D>(S.x, i, xi);
    double sw = 0.0, ss = 0.0;
    struct SRec {
        double m, rho;
        int8_t kind;
    };
    auto load = [&](int j) {
        SRec r;
        r.m = S.m[j];
        r.rho = S.rho[j];
        r.kind = kind ? kind[j] : (int8_t)KIND_FLUID;
        return r;
    };
    const double R2 = s.R2;
    for_pairs_d<D, SCAN, DEPTH, true, D == 3>(s, s.cell_start, S.x, xi, i, i, s.vnbr, s.nvnb[i], load,
                                [&](int, double d2, const double*, const SRec& o) {
        // weight 1 for neighbours under the exact rule among fluid particles
        const double in = d2 <= R2 && o.kind == KIND_FLUID ? 1.0 : 0.0;
        double W, dW;
        KQ::wdw(sqrt(d2) * s.inv_h, s.kA, s.kB, W, dW);
        const double rho_j = o.rho;
        const double a = W * o.m / rho_j * in;
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

// force (tuner runs): whether or not a re-initialisation is due.
template <int D, int K, bool SCAN, int DEPTH, int BS>
__global__ __launch_bounds__(SCAN ? kBlock : BS) void k_shepard(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, int ntiles, int force) {
    constexpr int SUB = SCAN ? 1 : kBlock / BS;  // blocks per tile
    const int nv = ntiles * SUB;
    for (int vb = blockIdx.x; vb < nv; vb += gridDim.x) {
        const int v = tile_index(vb, nv);
        const Tile tile = tiles[v / SUB];
        const int first = tile.begin + (v % SUB) * BS;
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (!force && (s.kreinit <= 0 || (dyn.step + 1) % s.kreinit != 0)) continue;
        // the sort moves rhon into rho only when a scheduled pass ran (a
        // tuner run or a skipped pass leaves the densities as they are)
        if (!force && threadIdx.x == 0) s.dyn->shepard_ran = 1;
        for_work<SCAN>(s, dyn, tile, first, [&](int i) {
            // a particle whose candidates overflowed its list (rare) scans its cells
            if (!SCAN && overflowed(s, i))
                shepard_particle<D, K, true, 1>(s, dyn, i);
            else
                shepard_particle<D, K, SCAN, DEPTH>(s, dyn, i);
        });
    }
}


// The pair passes run in workgroups of bs_ threads (the tuner's choice of
// 64, 128 or 256 for the pass), kBlock / bs_ of them per particle tile; the
// overflow (SCAN) instances always run kBlock-thread workgroups.
#define ESPH_PAIR_LAUNCH(KERNEL_T, ...)                                                          \
    do {                                                                                        \
        if constexpr (S) {                                                                      \
            constexpr int BSZ = kBlock;                                                         \
            hipLaunchKernelGGL((KERNEL_T), dim3(nt), dim3(BSZ), 0, st, __VA_ARGS__);            \
        } else {                                                                                \
            ESPH_DISPATCH_BS(bs_, {                                                             \
                hipLaunchKernelGGL((KERNEL_T), dim3(nt * (kBlock / BSZ)), dim3(BSZ), 0, st,      \
                                   __VA_ARGS__);                                                \
            });                                                                                 \
        }                                                                                       \
    } while (0)

// Workgroups of the brick kernels: eight ranges of a scene's bricks, one per
// XCD, kBrickGrid / 8 workgroups each (each workgroup one brick per scene
// while a scene has at most kBrickGrid bricks holding particles).
constexpr int kBrickGrid = 4096;

template <int D>
hipError_t launch_continuity_d(const Bucket& b, hipStream_t st) {
    hipError_t e = hipSuccess;
    const int bs_ = b.cbs, dry = b.tuning ? 1 : 0;
    if (b.ntiles <= 0) return e;
    ESPH_DISPATCH_K(b.kernel, {
        auto launch = [&](auto scan, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            ESPH_DISPATCH_DEPTH(S ? 1 : b.cdepth, {
                if (b.prec == ESPH_FP64)
                    ESPH_PAIR_LAUNCH(SINGLE_ARG(k_continuity<D, K, double, S, DP, BSZ>), b.scenes, tl, nt, dry);
                else
                    ESPH_PAIR_LAUNCH(SINGLE_ARG(k_continuity<D, K, float, S, DP, BSZ>), b.scenes, tl, nt, dry);
            });
        };
        if (b.has_tile) launch(std::false_type{}, b.tiles, b.ntiles);
        if constexpr (D == 3) {
            if (b.has_brick && b.prec == ESPH_FP32_BULK) {
                ESPH_DISPATCH_DEPTH(b.cdepth, {
                    hipLaunchKernelGGL((k_continuity_brick<K, DP>), dim3(kBrickGrid), dim3(kBrickThreads), 0, st,
                                       b.scenes, b.btiles, b.nbtiles, dry);
                });
            }
        }
        e = hipGetLastError();
    });
    return e;
}

template <int D, bool TH, bool EW = false>
hipError_t launch_forces_t(const Bucket& b, hipStream_t st, bool main = true) {
    hipError_t e = hipSuccess;
    const int bs_ = b.bs;
    ESPH_DISPATCH_K(b.kernel, {
        e = launch_both(b, st, [&](auto scan, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            if (!S && !main) return;
            ESPH_DISPATCH_DEPTH(S ? 1 : b.depth, {
                if constexpr (EW)
                    ESPH_PAIR_LAUNCH(SINGLE_ARG(k_forces<D, K, float, TH, S, DP, BSZ, true>), b.scenes, tl, nt);
                else if (b.prec == ESPH_FP64)
                    ESPH_PAIR_LAUNCH(SINGLE_ARG(k_forces<D, K, double, TH, S, DP, BSZ>), b.scenes, tl, nt);
                else
                    ESPH_PAIR_LAUNCH(SINGLE_ARG(k_forces<D, K, float, TH, S, DP, BSZ>), b.scenes, tl, nt);
            });
        });
    });
    return e;
}

// One force-pass instance per kind of scene in the bucket: thermal, plain,
// (3D FP32_BULK) the two-record walk of FLAG_EOS_WALK scenes, and their brick
// walk (FLAG_BRICK).
template <int D>
hipError_t launch_forces_d(const Bucket& b, hipStream_t st) {
    hipError_t e = hipSuccess;
    if (b.has_thermal) e = launch_forces_t<D, true>(b, st);
    if (e == hipSuccess && b.has_plain) e = launch_forces_t<D, false>(b, st);
    if constexpr (D == 3) {
        if (b.prec == ESPH_FP32_BULK) {
            if (e == hipSuccess && b.has_eos) e = launch_forces_t<D, false, true>(b, st);
            if (e == hipSuccess && b.has_brick && b.ntiles > 0) {
                ESPH_DISPATCH_K(b.kernel, {
                    ESPH_DISPATCH_DEPTH(b.depth, {
                        hipLaunchKernelGGL((k_forces_brick<K, DP>), dim3(kBrickGrid), dim3(kBrickThreads), 0, st,
                                           b.scenes, b.btiles, b.nbtiles);
                    });
                });
                e = hipGetLastError();
            }
        }
    }
    return e;
}

template <int D>
hipError_t launch_shepard_d(const Bucket& b, hipStream_t st) {
    hipError_t e = hipSuccess;
    const int bs_ = b.sbs, force = b.tuning ? 1 : 0;
    ESPH_DISPATCH_K(b.kernel, {
        e = launch_both(b, st, [&](auto scan, const Tile* tl, int nt) {
            constexpr bool S = decltype(scan)::value;
            ESPH_DISPATCH_DEPTH(S ? 1 : b.sdepth, {
                ESPH_PAIR_LAUNCH(SINGLE_ARG(k_shepard<D, K, S, DP, BSZ>), b.scenes, tl, nt, force);
            });
        });
    });
    return e;
}

}  // namespace esph
