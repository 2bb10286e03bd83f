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
#include "esph_common.hpp"

namespace esph {
// ---------------------------------------------------------------------------
// Counting sort of the step's new positions into cells.  Deterministic: the
// order inside a cell is the particles' previous order.  The check predicate
// (FluidWallSolve.jl:73-81, deleteat! is stable) sends deleted particles to
// the tail bucket `ncell`, which is dropped.
// mode 0: end of a step; mode 1: upload (rho from rhon, no check).
// The cell counts are zero on entry: the allocation clears them and the
// scan's last phase clears them again once read.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_key(const SceneDev& s, int D, const double* x) {
    const int cx = cell_coord(x[0], s.lo[0], s.inv_cs[0], s.nc[0]);
    const int cy = cell_coord(x[1], s.lo[1], s.inv_cs[1], s.nc[1]);
    const int cz = D == 3 ? cell_coord(x[2], s.lo[2], s.inv_cs[2], s.nc[2]) : 0;
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

// Whether this step ends with a sort: always at upload, when the candidate
// lists are rebuilt every step, on check steps, for the inlet/outlet solver,
// and when some particle has moved (Rv - R) / 2 since the lists were built
// (or a position is not finite).  Otherwise the step commits in place.
// Reads only state that is final once the force pass has run.
__device__ __forceinline__ int sort_decision(const SceneDev& s, const SceneDyn& d, int mode) {
    const float disp2 = __uint_as_float(d.disp2_bits);
    const bool moved = !((double)disp2 <= s.disp2_limit);
    return (mode != 0 || !s.verlet || (s.flags & FLAG_IO) || check_due(s, d) || moved) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_sort_bin(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        // every block takes the step's decision itself; the scene's first
        // tile records it for the later sort kernels
        const int sorted = sort_decision(s, dyn, mode);
        if (tile.begin == 0 && threadIdx.x == 0) s.dyn->sorted = sorted;
        if (!sorted) continue;  // a step without sort
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const int D = s.dim;
        double x[3] = {0.0, 0.0, 0.0};
        bool finite = true;
        const double* xnew = s.set[next_set(dyn)].x;  // the step's (or the upload's) positions
        for (int k = 0; k < D; ++k) {
            x[k] = xnew[i * D + k];
            finite = finite && isfinite(x[k]);
        }
        int key;
        bool dead = false;
        if (mode == 0 && check_due(s, dyn)) {
            for (int k = 0; k < D; ++k)
                if (x[k] > s.check_max[k] || x[k] < s.check_min[k]) dead = true;
            for (int q = 0; q < s.check_nplanes; ++q) {
#pragma clang fp contract(off)
                double dot = s.check_planes[q][0] * x[0] + s.check_planes[q][1] * x[1];
                if (D == 3) dot = dot + s.check_planes[q][2] * x[2];
                if (dot > s.check_planes[q][3]) dead = true;
            }
        }
        // inlet/outlet migration (FluidInletOutletWallSolve.jl:196-262)
        int spawn_key = -1;
        if (mode == 0 && (s.flags & FLAG_IO)) {
            const FluidSet& S = s.set[cur_set(dyn)];
            const int kd = S.kind[i], a = s.io_axis;
            if (kd == KIND_OUTLET && x[a] > s.outlet_exit) dead = true;
            if (kd == KIND_FLUID && x[a] > s.outlet_enter) s.leave_id[atomicAdd(&s.dyn->nleave, 1)] = S.id[i];
            if (kd == KIND_INLET && x[a] > s.inlet_enter) {
                spawn_key = cell_key(s, D, x);  // the fluid copy stays where the inlet particle is
                x[a] -= s.inlet_shift;          // the inlet particle moves back
            }
        }
        key = dead ? s.ncell : cell_key(s, D, x);
        if (!dead && (s.flags & FLAG_BRICK)) s.bflag[brick_of_cell(s, key)] = 1;
        if (!finite) atomicOr(&s.dyn->err, 1);
        {
            // particles beyond the grid share its border cells (clamped):
            // counted, since many of them make those cells crowded
            bool far = false;
            for (int k = 0; k < D; ++k) {
                const double f = floor((x[k] - s.lo[k]) * s.inv_cs[k]);
                far = far || !(f >= -1.0 && f <= (double)s.nc[k]);
            }
            if (far && !dead) atomicAdd(&s.dyn->outside_acc, 1);
        }
        s.key[i] = key;
        s.slot[i] = wave_counted_slot(s.count, key);
        if (spawn_key >= 0) {
            const int e = atomicAdd(&s.dyn->nspawn, 1);
            if (dyn.n_live + e < s.cap) {
                s.spawn_src[e] = i;
                s.spawn_id[e] = s.set[cur_set(dyn)].id[i];
                s.spawn_key[e] = spawn_key;
                s.spawn_slot[e] = atomicAdd(&s.count[spawn_key], 1);
            } else {
                atomicOr(&s.dyn->err, 2);  // capacity exhausted: the copy is dropped
            }
        }
    }
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

// Exclusive scan of a scene's tile totals (bsum, in place) by one block.
__device__ __forceinline__ void scan_tile_totals(const SceneDev& s) {
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

// The bricks marked by this sort's binning, in brick order, into blist
// (FLAG_BRICK), and the marks cleared; by one block.  Each thread takes a
// contiguous run of 16-byte words of the marks (one byte per brick, 0 or 1,
// the array padded with zeros to a multiple of 16 * kBlock bytes).
__device__ __forceinline__ void compact_bricks(const SceneDev& s) {
    const int nw = ((s.nbricks + 16 * kBlock - 1) / (16 * kBlock)) * kBlock;  // 16-byte words
    const int per = nw / kBlock;
    uint4* w = reinterpret_cast<uint4*>(s.bflag) + (size_t)threadIdx.x * per;
    int c = 0;
    for (int q = 0; q < per; ++q) {
        const uint4 v = w[q];
        c += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
    }
    int total;
    int off = block_exclusive_scan(c, &total);
    const int b0 = (int)threadIdx.x * per * 16;
    for (int q = 0; q < per; ++q) {
        const uint4 v = w[q];
        const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if ((u[k >> 2] >> (8 * (k & 3))) & 0xffu) s.blist[off++] = b0 + 16 * q + k;
        w[q] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (threadIdx.x == 0) s.dyn->nbrick = total;
}

// Phases 1 and 2: scan of kCellTile counts per tile into cell_local, tile
// total to bsum; the last of a scene's tiles to finish (a ticket counter, no
// waiting) then scans the scene's tile totals.  Release: each block's total is fenced before its
// ticket; acquire: the last block fences after its ticket, before reading.
__global__ __launch_bounds__(kBlock) void k_sort_scan1(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ ctiles, int ntiles) {
    __shared__ int last;
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = ctiles[t];
        const SceneDev& s = scenes[tile.scene];
        if (!s.dyn->sorted) continue;  // a step without sort
        // each wavefront scans a quarter of the tile, a row of 256 cells
        // (one 16-byte word per lane, 1 KiB per load) at a time, carrying
        // the running sum; count and cell_start are allocated in whole tiles
        constexpr int W = kScanItems / 4;  // words per lane
        const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
        const int wbase = tile.begin + wave * (kCellTile / (kBlock / 64));
        int4 v[W];
        int ex[W];
        int carry = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int c = wbase + 256 * k + 4 * lane;
            int4 w = *reinterpret_cast<const int4*>(s.count + c);
            // cells past ncell count nothing
            w.x = c <= s.ncell ? w.x : 0;
            w.y = c + 1 <= s.ncell ? w.y : 0;
            w.z = c + 2 <= s.ncell ? w.z : 0;
            w.w = c + 3 <= s.ncell ? w.w : 0;
            v[k] = w;
            const int sum = w.x + w.y + w.z + w.w;
            const int incl = wave_inclusive_scan(sum);
            ex[k] = carry + incl - sum;
            carry += __builtin_amdgcn_readlane(incl, 63);
        }
        __shared__ int wtot[kBlock / 64];
        if (lane == 0) wtot[wave] = carry;
        __syncthreads();
        int woff = 0, total = 0;
#pragma unroll
        for (int q = 0; q < kBlock / 64; ++q) {
            woff += q < wave ? wtot[q] : 0;
            total += wtot[q];
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int c = wbase + 256 * k + 4 * lane;
            int4 o;
            o.x = woff + ex[k];
            o.y = o.x + v[k].x;
            o.z = o.y + v[k].y;
            o.w = o.z + v[k].z;
            // tile-local prefixes; k_sort_offsets_place adds the tile offsets
            if (c <= s.ncell) *reinterpret_cast<int4*>(s.cell_local + c) = o;
        }
        if (threadIdx.x == 0) {
            s.bsum[tile.begin / kCellTile] = total;
            __threadfence();
            last = atomicAdd(&s.dyn->scan_done, 1) == s.cell_tiles - 1;
        }
        __syncthreads();
        if (last) {
            __threadfence();
            scan_tile_totals(s);
            if (s.flags & FLAG_BRICK) compact_bricks(s);
            if (threadIdx.x == 0) s.dyn->scan_done = 0;
        }
        __syncthreads();
    }
}

// Phase 3 and placement in one launch: blocks [0, nct) turn the cell tiles'
// local prefixes (cell_local, from k_sort_scan1) into cell_start by adding
// their tile offsets, write the grand total after the last cell and clear
// the counts for the next sort; blocks [nct, nct + npt) place each particle
// of a particle tile at cell start + slot in tmp, with the cell start taken
// from cell_local and the tile offsets directly (neither part reads what the
// other writes).  The block of each scene's first cell tile also copies the
// scene's step state to SceneDev::snap for k_sort_scatter, sort or not.
__device__ __forceinline__ int cell_start_of(const SceneDev& s, int c) {
    return s.cell_local[c] + s.bsum[c / kCellTile];
}

__global__ __launch_bounds__(kBlock) void k_sort_offsets_place(const SceneDev* __restrict__ scenes,
                                                               const Tile* __restrict__ ctiles, int nct,
                                                               const Tile* __restrict__ ptiles, int npt,
                                                               int mode) {
    for (int tb = blockIdx.x; tb < nct + npt; tb += gridDim.x) {
        if (tb < nct) {
            const Tile tile = ctiles[tb];
            const SceneDev& s = scenes[tile.scene];
            if (tile.begin == 0 && threadIdx.x == 0) *s.snap = *s.dyn;
            if (!s.dyn->sorted) continue;  // a step without sort
            const int off = s.bsum[tile.begin / kCellTile];
            // 16-byte words across the block (1 KiB per wavefront load)
#pragma unroll 4
            for (int q = (int)threadIdx.x; q < kCellTile / 4; q += kBlock) {
                const int c = tile.begin + 4 * q;
                if (c <= s.ncell) {
                    const int4 w = *reinterpret_cast<const int4*>(s.cell_local + c);
                    const int4 n = *reinterpret_cast<const int4*>(s.count + c);
                    const int4 o = make_int4(w.x + off, w.y + off, w.z + off, w.w + off);
                    if (c + 3 <= s.ncell) {
                        *reinterpret_cast<int4*>(s.cell_start + c) = o;
                        *reinterpret_cast<int4*>(s.count + c) = make_int4(0, 0, 0, 0);  // ready for the next sort
                        if (c + 3 == s.ncell) s.cell_start[s.ncell + 1] = o.w + n.w;
                    } else {
                        // the word holding the last cell: the grand total follows it
                        const int ev[4] = {o.x, o.y, o.z, o.w}, m[4] = {n.x, n.y, n.z, n.w};
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            if (c + k <= s.ncell) {
                                s.cell_start[c + k] = ev[k];
                                s.count[c + k] = 0;
                                if (c + k == s.ncell) s.cell_start[c + k + 1] = ev[k] + m[k];
                            }
                        }
                    }
                }
            }
        } else {
            const Tile tile = ptiles[tb - nct];
            const SceneDev& s = scenes[tile.scene];
            if (!s.dyn->sorted) continue;  // a step without sort
            const int i = tile.begin + (int)threadIdx.x;
            const SceneDyn dyn = *s.dyn;
            if (i < dyn.n_live) {
                s.tmp[cell_start_of(s, s.key[i]) + s.slot[i]] = i;
            } else if (mode == 0 && (s.flags & FLAG_IO) && i < dyn.n_live + spawn_count(s, dyn)) {
                // a fluid copy of an entering inlet particle; its rank among this
                // step's copies follows the inlet order (push! in that order)
                const int ei = i - dyn.n_live, ns = spawn_count(s, dyn), my = s.spawn_id[ei];
                int r = 0;
                for (int q = 0; q < ns; ++q) r += s.spawn_id[q] < my ? 1 : 0;
                s.spawn_order[r] = ei;
                s.tmp[cell_start_of(s, s.spawn_key[ei]) + s.spawn_slot[ei]] = dyn.n_live + r;
            }
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

// The step's end for one scene (one thread): the new particle count and
// order, or the switch to the step's outputs, and the per-step counters
// cleared.
__device__ __forceinline__ void sort_finish(const SceneDev& s, int mode) {
    SceneDyn* d = s.dyn;
    if (d->sorted) {
        if (mode == 0 && (s.flags & FLAG_IO)) {
            const int room = s.cap - d->n_live;  // n_live is still the step's input count here
            d->next_fluid_id += d->nspawn < room ? d->nspawn : room;
            d->next_outlet_id += d->nleave;
        }
        d->n_live = s.cell_start[s.ncell];
        d->outside = d->outside_acc;
        d->outside_acc = 0;
        d->parity ^= 1;  // the sort wrote the other order, per-step copy dpar
        d->rebuild = 1;
        d->nsorts += 1;
    } else {
        d->dpar ^= 1;  // the step's outputs become the current state
        d->rebuild = 0;
    }
    d->novf = 0;
    d->shepard_ran = 0;
    d->nspawn = 0;
    d->nleave = 0;
    d->disp2_bits = 0;
    if (mode == 0) d->step += 1;
}

// Final position = cell start + number of cell members with a smaller
// entry index (previous order; a step's fluid copies of inlet particles come
// after every existing particle, in inlet order); gathers every field into
// the other set.  With the inlet/outlet solver it also applies the step's
// migration (kinds and group sequence numbers) and the buffer velocities.
__device__ __forceinline__ void scatter_particle(const SceneDev& s, const SceneDyn& dyn, const Tile& tile,
                                                 int mode) {
    {
        if (!dyn.sorted) {
            // a step without sort keeps the particle order: its outputs are
            // already in the other per-step copy (sort_finish switches to it);
            // only a Shepard step's densities move, into the order's rho
            const int i = tile.begin + (int)threadIdx.x;
            if (dyn.shepard_ran && i < dyn.n_live) s.set[cur_set(dyn)].rho[i] = s.rhon[i];
            return;
        }
        const bool io = mode == 0 && (s.flags & FLAG_IO);
        const int ns = io ? spawn_count(s, dyn) : 0;
        const int q = tile.begin + (int)threadIdx.x;
        if (q >= dyn.n_live + ns) return;
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
            if (key == s.ncell) return;  // deleted
        }
        const int b = s.cell_start[key], e = s.cell_start[key + 1];
        int rank = 0;
        for (int w = b; w < e; ++w) rank += s.tmp[w] < q ? 1 : 0;
        const int p = b + rank;
        // from the step's outputs (N) and the current order's other arrays
        // (S) into the other order, per-step copy dpar (free once the step's
        // outputs exist)
        const FluidSet& S = s.set[cur_set(dyn)];
        const FluidSet& N = s.set[next_set(dyn)];
        const FluidSet& O = s.set[2 * (dyn.parity ^ 1) + dyn.dpar];
        const int D = s.dim;
        const bool thermal = (s.flags & FLAG_THERMAL) != 0;
        double x[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < D; ++k) {
            x[k] = N.x[src * D + k];
            v[k] = N.v[src * D + k];
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
        for (int k = 0; k < D; ++k) s.xref[p * D + k] = x[k];  // the candidate lists are rebuilt from here
        const bool rho_new = mode == 1 || (dyn.shepard_ran && src_fluid && !spawn);
        const double rho = rho_new ? s.rhon[src] : S.rho[src];
        if (O.xq) {
            // 3D: w = rho - rho0 (FluidSet::xq)
            int32_t qx[3];
            quantize(s.lo, s.inv_qu, D, x, qx);
            if (D == 3)
                static_cast<int4*>(O.xq)[p] = make_int4(qx[0], qx[1], qx[2], __float_as_int((float)(rho - s.rho0)));
            else
                static_cast<int2*>(O.xq)[p] = make_int2(qx[0], qx[1]);
        }
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
            Tt = N.T[src];
            kap = S.kappa[src];
            O.T[p] = Tt;
            O.kappa[p] = kap;
            O.cp[p] = S.cp[src];
        }
        if (s.prec == ESPH_FP32_BULK) {
            // float records hold T - t0: the conduction terms use only
            // differences, which keep their precision near the reference
            // temperature this way
            store_sorted_rec<float>(O, D, p, v, m, gap, Tt - s.t0, kap, thermal);
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

// The scatter of every particle tile, and the step's end (sort_finish) by
// the scene's first tile.  The tiles read the step state from the copy
// taken before this launch (SceneDev::snap), so that end can be written
// while other tiles still run; the kernels after this one see the new state.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const SceneDev* __restrict__ scenes,
                                                         const Tile* __restrict__ tiles, int ntiles, int mode) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.snap;
        scatter_particle(s, dyn, tile, mode);
        if (tile.begin == 0 && threadIdx.x == 0) sort_finish(s, mode);
    }
}

// Candidate lists of the new order after a sort (every fluid particle within
// Rv = R (1 + skin) of each fluid particle, and the wall particles within Rv
// where the cell is near a wall), for every scene but the brick-walk ones,
// whose continuity pass builds them with the candidates' LDS slots.  A
// launch of its own: without the pair arithmetic beside it, the build takes
// fewer registers than the continuity pass and keeps more wavefronts in
// flight to cover its gathers.  Steps without a sort keep the lists.
__global__ __launch_bounds__(kBlock) void k_rebuild(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, int ntiles) {
    for (int tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
        const int t = tile_index(tb, ntiles);
        const Tile tile = tiles[t];
        const SceneDev& s = scenes[tile.scene];
        const SceneDyn dyn = *s.dyn;
        if (!dyn.rebuild || (s.flags & FLAG_BRICK)) continue;
        const int i = tile.begin + (int)threadIdx.x;
        if (i >= dyn.n_live) continue;
        const FluidSet& S = s.set[cur_set(dyn)];
        if (S.kind && S.kind[i] != KIND_FLUID) {
            // inlet/outlet buffers walk no pairs
            s.nvnb[i] = 0;
            if (s.nw > 0) s.nnbw[i] = 0;
            continue;
        }
        double xi[3];
        if (s.dim == 3) {
            load_x<3>(S.x, i, xi);
            rebuild_lists<3>(s, S, xi, i);
        } else {
            load_x<2>(S.x, i, xi);
            rebuild_lists<2>(s, S, xi, i);
        }
    }
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
        const FluidSet& S = s.set[cur_set(dyn)];
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

// The pair decisions of the step's own walks, exported for checking: for
// each live fluid particle i, every candidate of its list that the pair
// passes weight with 1 -- the quantized-geometry decision of the FP32_BULK
// walks (qgeom, taken by the continuity pass and the 3D force pass), the
// binary64 rule over the candidate list in FP64, and for walls the binary64
// rule over the wall candidates; particles whose lists overflowed take the
// cell scan of the SCAN instances.  Directed pairs (id_i, id_j): both orders
// of a fluid-fluid pair appear when both walks decide it.  Lists that the
// next continuity pass would rebuild (brick walks, after a sort) are rebuilt
// here the same way first.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void k_walk_pairs(const SceneDev* __restrict__ scenes,
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
        const FluidSet& S = s.set[cur_set(dyn)];
        double xi[3];
        load_x<D>(S.x, i, xi);
        // (brick-walk scenes build their lists in the continuity pass)
        if ((dyn.rebuild || !s.verlet) && (s.flags & FLAG_BRICK)) rebuild_lists<D>(s, S, xi, i);
        int64_t w = offs ? offs[i] : 0;
        int cnt = 0;
        const int idi = S.id[i];
        auto emit = [&](int idj) {
            if (out) {
                out[2 * w] = idi;
                out[2 * w + 1] = idj;
                ++w;
            }
            ++cnt;
        };
        const double R2 = s.R2;
        const bool scan = overflowed(s, i);
        if (pairset == ESPH_PAIRS_FLUID_FLUID) {
            if (scan) {
                int c[3];
                cell_of<D>(s, xi, c);
                for_each_neighbour<D>(s, s.cell_start, S.x, xi, c, i,
                                      [&](int j, double, const double*) { emit(S.id[j]); });
            } else if constexpr (sizeof(T) == 4) {
                for_pairs_q<D, 1>(s, S, xi, i, s.vnbr, s.nvnb[i], [](int) { return 0; },
                                  [&](int j, bool in, float, const float*, int, int32_t) {
                                      if (in) emit(S.id[j]);
                                  });
            } else {
                for_pairs_d<D, false, 1>(s, s.cell_start, S.x, xi, i, i, s.vnbr, s.nvnb[i], [](int) { return 0; },
                                         [&](int j, double d2, const double*, int) {
                                             if (d2 <= R2) emit(S.id[j]);
                                         });
            }
        } else if (s.nw > 0) {
            if (scan) {
                int c[3];
                cell_of<D>(s, xi, c);
                for_each_neighbour<D>(s, s.wcell_start, s.wx, xi, c, -1,
                                      [&](int j, double, const double*) { emit(s.wid[j]); });
            } else {
                for_pairs_d<D, false, 1>(s, s.wcell_start, s.wx, xi, i, -1, s.wnbr, s.nnbw[i], [](int) { return 0; },
                                         [&](int j, double d2, const double*, int) {
                                             if (d2 <= R2) emit(s.wid[j]);
                                         });
            }
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
    const FluidSet& S = s.set[cur_set(dyn)];
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

hipError_t launch_continuity(const Bucket& b, hipStream_t st) {
    return b.dim == 2 ? launch_continuity_d<2>(b, st) : launch_continuity_d<3>(b, st);
}

hipError_t launch_forces(const Bucket& b, hipStream_t st) {
    return b.dim == 2 ? launch_forces_d<2>(b, st) : launch_forces_d<3>(b, st);
}

hipError_t launch_shepard(const Bucket& b, hipStream_t st) {
    return b.dim == 2 ? launch_shepard_d<2>(b, st) : launch_shepard_d<3>(b, st);
}

hipError_t launch_sort(const SortPlan& p, int mode, hipStream_t st) {
    dim3 blk(kBlock);
    hipLaunchKernelGGL(k_sort_bin, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_scan1, dim3(grid_for(p.nctiles)), blk, 0, st, p.scenes, p.ctiles, p.nctiles);
    hipLaunchKernelGGL(k_sort_offsets_place, dim3(grid_for(p.nctiles + p.nptiles)), blk, 0, st, p.scenes, p.ctiles,
                       p.nctiles, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_sort_scatter, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles, mode);
    hipLaunchKernelGGL(k_rebuild, dim3(grid_for(p.nptiles)), blk, 0, st, p.scenes, p.ptiles, p.nptiles);
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

hipError_t launch_walk_pairs(const Bucket& b, int pairset, int32_t* counts, const int64_t* offs, int32_t* out,
                             hipStream_t st) {
    if (b.ntiles <= 0) return hipSuccess;
    dim3 g(grid_for(b.ntiles)), blk(kBlock);
    const bool f32 = b.prec == ESPH_FP32_BULK;
    if (b.dim == 2) {
        if (f32) hipLaunchKernelGGL((k_walk_pairs<2, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
        else hipLaunchKernelGGL((k_walk_pairs<2, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
    } else {
        if (f32) hipLaunchKernelGGL((k_walk_pairs<3, float>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
        else hipLaunchKernelGGL((k_walk_pairs<3, double>), g, blk, 0, st, b.scenes, b.tiles, b.ntiles, pairset, counts, offs, out);
    }
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
