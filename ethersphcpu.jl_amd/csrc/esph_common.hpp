// esph_common.hpp — device helpers shared by the gfx950 kernels of the
// per-step SPH hot path: XCD-aware tile order, cell-grid candidate rows,
// neighbour lists, packed pair records and launch dispatch.
#pragma once

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

// Gathers through a raw buffer resource (gfx9 descriptor, word 3 0x00020000)
// on the wave-uniform base of a scene's array: each lane supplies a 32-bit
// byte offset (one shift or multiply) instead of a 64-bit address (three
// instructions), and the loads count only in vmcnt.  Every lane of a
// wavefront works on one scene (a tile is one scene's), so the base is the
// same in all lanes; it is moved to scalar registers as such.  Offsets stay
// below 2^32: scenes hold fewer than 2^27 particles and walls.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t urs(const void* base) {
    const uint64_t a = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(((uint64_t)hi << 32) | lo), (short)0, -1,
                                             0x00020000);
}

// Record j of an array of 16-byte (or 8-byte) records.
template <typename V>
__device__ __forceinline__ V ldrec(const void* base, int j) {
    const __amdgpu_buffer_rsrc_t r = urs(base);
    if constexpr (sizeof(V) == 16) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)((uint32_t)j << 4), 0, 0);
        return __builtin_bit_cast(V, v);
    } else {
        static_assert(sizeof(V) == 8, "16- or 8-byte records");
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)((uint32_t)j << 3), 0, 0);
        return __builtin_bit_cast(V, v);
    }
}

// Binary64 position j (D doubles per particle).
template <int D>
__device__ __forceinline__ void load_x(const double* __restrict__ x, int j, double* o) {
    const __amdgpu_buffer_rsrc_t r = urs(x);
    if (D == 2) {
        const double2 q = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)((uint32_t)j << 4), 0, 0));
        o[0] = q.x;
        o[1] = q.y;
    } else {
        const uint32_t off = (uint32_t)j * 24u;
        const double2 q = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
        const double z = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 16u), 0, 0));
        o[0] = q.x;
        o[1] = q.y;
        o[2] = z;
    }
}

// A 4-entry candidate-list chunk, loaded non-temporally: each pass reads a
// particle's list once, and keeping those lines out of the caches leaves
// them to the positions and records the pass gathers (measured on the
// 1M-particle 3D dam break: FP32 step 0.619 -> 0.598 ms, FP64 1.078 -> 1.064).
typedef int esph_nint4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_chunk(const int32_t* p) {
    const esph_nint4 v = __builtin_nontemporal_load(reinterpret_cast<const esph_nint4*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}

// Candidate rows of a point xi in cell c: the grid's cells are at least Rv /
// stencil[k] wide along axis k, so neighbours lie within stencil[k] cells
// along it.  For
// each (dy, dz) the cells of the row are x-contiguous in memory, so a row is
// one index range [j0, j1).  Rows and row ends whose nearest point is
// farther than R from xi (judged from xi's offset inside its cell, with a
// slack that keeps the test conservative under rounding and clamping) are
// skipped.  Calls row(j0, j1) -- or row(j0, j1, cy, cz) with the row's cell
// coordinates -- for each remaining row.
template <int D, typename F>
__device__ __forceinline__ void for_each_row(const SceneDev& s, const int32_t* __restrict__ cstart,
                                             const double* xi, const int c[3], F&& row) {
    // trimmed against the candidate radius Rv: between list rebuilds the
    // particles sit up to (Rv - R) / 2 away from the cells they were sorted
    // into, and Rv covers that.  In float, with a slack (1e-5 of a cell, and
    // 1e-5 of Rv^2) far above the float roundings of these few operations on
    // values below a few cells: the trimming only ever keeps a few more
    // cells, whose particles the candidate test then rejects.
    const float lim = (float)(s.Rv2 * (1.0 + 1e-5));
    float p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) p[k] = k < D ? (float)(xi[k] - (s.lo[k] + (double)c[k] * s.cs[k])) : 0.0f;
    // distance from xi to the slab of the cell at offset o along axis k
    auto dmin = [&](int o, int k) {
        const float cs = (float)s.cs[k], pk = p[k];
        float d = o == 0 ? 0.0f : (o > 0 ? (float)(o - 1) * cs + (cs - pk) : (float)(-o - 1) * cs + pk);
        d -= 1e-5f * cs;
        return d > 0.0f ? d : 0.0f;
    };
    const int mx = s.stencil[0], my = s.stencil[1], mz = D == 3 ? s.stencil[2] : 0;
    // the row found last, walked once the next one's range is in flight
    bool pending = false;
    int pj0 = 0, pj1 = 0, pcy = 0, pcz = 0;
    auto emit = [&]() {
        if constexpr (std::is_invocable_v<F, int, int, int, int>)
            row(pj0, pj1, pcy, pcz);
        else
            row(pj0, pj1);
    };
    for (int dz = -mz; dz <= mz; ++dz) {
        const int cz = c[2] + dz;
        if (cz < 0 || cz >= s.nc[2]) continue;
        const float rz = D == 3 ? dmin(dz, 2) : 0.0f;
        for (int dy = -my; dy <= my; ++dy) {
            const int cy = c[1] + dy;
            if (cy < 0 || cy >= s.nc[1]) continue;
            const float ry = dmin(dy, 1);
            const float r2 = ry * ry + rz * rz;
            if (r2 > lim) continue;
            // the cells along x within the distance left, ext: at offset
            // o > 0 the slab is (o - 1) cs + (cs - p0) - slack away, at o < 0
            // (-o - 1) cs + p0 - slack (one square root; the float roundings
            // are far inside the slack, so no cell holding a candidate is cut)
            const float ext = __builtin_sqrtf(lim - r2), csx = (float)s.cs[0], icx = (float)s.inv_cs[0];
            const float sl = 1e-5f * csx;
            const int nl = (int)floorf((ext - p[0] + sl) * icx) + 1, nh = (int)floorf((ext + p[0] + sl) * icx);
            const int xl = -(nl < 0 ? 0 : (nl > mx ? mx : nl)), xh = nh < 0 ? 0 : (nh > mx ? mx : nh);
            const int cx0 = c[0] + xl > 0 ? c[0] + xl : 0;
            const int cx1 = c[0] + xh < s.nc[0] - 1 ? c[0] + xh : s.nc[0] - 1;
            if (cx0 > cx1) continue;
            // this row's range is loaded before the previous row is walked,
            // so the two loads' latency overlaps that walk
            const int j0 = cstart[cell_linear(s, cx0, cy, cz)], j1 = cstart[cell_linear(s, cx1, cy, cz) + 1];
            if (pending) emit();
            pj0 = j0;
            pj1 = j1;
            pcy = cy;
            pcz = cz;
            pending = true;
        }
    }
    if (pending) emit();
}

// Iterates the particles of the candidate rows of xi whose distance
// satisfies the neighbour rule (FP64, exact).  The candidate loop issues the
// loads of U candidates before testing them (U = 1 where the scan is the
// rare fallback inside a pass, to keep that pass's registers).
template <int D, int U = 4, typename F>
__device__ __forceinline__ void for_each_neighbour(const SceneDev& s, const int32_t* __restrict__ cstart,
                                                   const double* __restrict__ xall, const double* xi,
                                                   const int c[3], int self, F&& f) {
    const double R2 = s.R2;
    for_each_row<D>(s, cstart, xi, c, [&](int j0, int j1) {
        for (int j = j0; j < j1; j += U) {
            double xj[U][3];
#pragma unroll
            for (int u = 0; u < U; ++u) load_x<D>(xall, j + u < j1 ? j + u : j1 - 1, xj[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
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

// Candidates of particle i within the candidate radius Rv over the adjacent
// cells, tested first in float on the grid-relative positions xf: a candidate
// whose float distance is clearly inside or clearly outside Rv is decided
// there; only those within the rounding band tau of Rv take the FP64 rule
// (d^2 <= Rv^2).  tau bounds the float error: 2^-21 (|xr_i|_inf + 2 Rv), so the
// decisions equal the FP64 rule's.  Calls f(j) for every candidate, in cell
// order; hook(cy, cz) before the candidates of each cell row.  Used right
// after a sort, when xf and the cells are current.
struct NoRowHook {
    __device__ __forceinline__ void operator()(int, int) const {}
};

template <int D, typename F, typename H = NoRowHook>
__device__ __forceinline__ void for_each_candidate_f32(const SceneDev& s, const int32_t* __restrict__ cstart,
                                                       const void* __restrict__ xfall,
                                                       const double* __restrict__ xall, const double* xi,
                                                       const int c[3], int self, F&& f, H&& hook = H{}) {
    float xr[3] = {(float)(xi[0] - s.lo[0]), (float)(xi[1] - s.lo[1]), D == 3 ? (float)(xi[2] - s.lo[2]) : 0.0f};
    const double amax = fmax(fabs((double)xr[0]), fmax(fabs((double)xr[1]), fabs((double)xr[2])));
    const double tau = ldexp(amax + 2.0 * s.Rv, -21);
    const double rin = s.Rv - tau, rout = s.Rv + tau;
    const float lo2 = rin > 0.0 ? (float)(rin * rin) * (1.0f - 0x1p-22f) : -1.0f;
    const float hi2 = (float)(rout * rout) * (1.0f + 0x1p-22f);
    const double R2 = s.Rv2;
    for_each_row<D>(s, cstart, xi, c, [&](int j0, int j1, int cy, int cz) {
        hook(cy, cz);  // the candidates that follow lie in this row
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
// With a brick map (BM not void: FLAG_BRICK fluid lists), each entry's slot
// in the brick region's LDS staging also goes to s.vslot, in the same layout.
struct NoBrickMap {};

template <int D, typename BM = NoBrickMap>
__device__ __forceinline__ int build_list(const SceneDev& s, const int32_t* __restrict__ cstart,
                                          const void* __restrict__ xfall, const double* __restrict__ xall,
                                          const double* xi, const int c[3], int i, int self,
                                          int32_t* __restrict__ list, int width, BM* bm = nullptr) {
    constexpr bool SLOTS = !std::is_same<BM, NoBrickMap>::value;
    int cnt = 0;
    int4 buf = make_int4(0, 0, 0, 0);
    ushort4 sbuf = make_ushort4(0, 0, 0, 0);
    const size_t cap = (size_t)s.cap;
    auto put = [&](int j) {
        const int r = cnt & 3;
        buf.x = r == 0 ? j : buf.x;
        buf.y = r == 1 ? j : buf.y;
        buf.z = r == 2 ? j : buf.z;
        buf.w = r == 3 ? j : buf.w;
        if constexpr (SLOTS) {
            const unsigned short q = (unsigned short)bm->slot(j);
            sbuf.x = r == 0 ? q : sbuf.x;
            sbuf.y = r == 1 ? q : sbuf.y;
            sbuf.z = r == 2 ? q : sbuf.z;
            sbuf.w = r == 3 ? q : sbuf.w;
        }
        if (r == 3 && cnt < width) {
            *reinterpret_cast<int4*>(list + list_at(cap, cnt & ~3, i)) = buf;
            if constexpr (SLOTS) *reinterpret_cast<ushort4*>(s.vslot + list_at(cap, cnt & ~3, i)) = sbuf;
        }
        ++cnt;
    };
    if constexpr (SLOTS)
        for_each_candidate_f32<D>(s, cstart, xfall, xall, xi, c, self, put, [&](int cy, int cz) { bm->row(cy, cz); });
    else
        for_each_candidate_f32<D>(s, cstart, xfall, xall, xi, c, self, put);
    if ((cnt & 3) != 0 && cnt < width) {
        *reinterpret_cast<int4*>(list + list_at(cap, cnt & ~3, i)) = buf;
        if constexpr (SLOTS) *reinterpret_cast<ushort4*>(s.vslot + list_at(cap, cnt & ~3, i)) = sbuf;
    }
    return cnt;
}

// Rebuilds particle i's fluid and wall candidate lists from the cells (right
// after a sort, or every step when the lists are not reused) and returns the
// fluid candidate count.
template <int D, typename BM = NoBrickMap>
__device__ __forceinline__ int rebuild_lists(const SceneDev& s, const FluidSet& S, const double* xi, int i,
                                             BM* bm = nullptr) {
    int c[3];
    cell_of<D>(s, xi, c);
    const int cnt = build_list<D>(s, s.cell_start, S.xf, S.x, xi, c, i, i, s.vnbr, s.kff, bm);
    s.nvnb[i] = cnt;
    if (s.nw > 0)
        s.nnbw[i] = s.wnear[cell_linear(s, c[0], c[1], c[2])]
                        ? build_list<D>(s, s.wcell_start, s.wxf, s.wx, xi, c, i, -1, s.wnbr, s.kfw)
                        : 0;
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

// Returns j once the values p[0..D) -- positions loaded before -- have
// arrived (an empty asm statement that reads them and rewrites j; no
// instruction is emitted).  The pair walks gather a group's positions first
// and its records with the indices returned here, so the records form a
// second group of loads issued after the first has landed.  Measured on the
// 1M-particle 3D dam break, FP32 continuity: 0.226 ms against 0.265 with
// both groups issued together (the compiler then issues part of them after
// a wait that drains every gather in flight).
template <int D, typename P>
__device__ __forceinline__ int after(int j, const P* p) {
    if (D == 3)
        asm volatile("" : "+v"(j) : "v"(p[0]), "v"(p[1]), "v"(p[2]));
    else
        asm volatile("" : "+v"(j) : "v"(p[0]), "v"(p[1]));
    return j;
}

// Loaders may gather a record in two groups: operator()(j) issues the first
// part with the positions, late(j, rec) the rest once the first part has
// arrived (see RecLoaderAB).
template <typename L, typename = void>
struct has_late : std::false_type {};
template <typename L>
struct has_late<L, std::void_t<decltype(&L::late)>> : std::true_type {};

// Pipelined form of for_pairs: neighbours are taken DEPTH at a time (WIDE
// with DEPTH 4: two chunks, 8 candidates, at a time).  The positions of a
// group, then its records load(j) (see after; with SPLIT false, together),
// are issued before any of their arithmetic, so each lane keeps DEPTH
// gathers in flight.  The list chunk after the current one is fetched while
// the current one is used.  Every slot of a written chunk holds a valid
// particle index (slots past the count keep earlier entries or zero), and a
// group past the last chunk reuses an earlier one, so a group may load past
// the count; body runs for the counted entries in list order (TAILW: also,
// without a branch, for the slots past the count with d^2 = 4 R^2, which
// every body's exact rule rejects, its terms adding zero -- for bodies with
// float arithmetic, where the extra terms cost about what the branch per
// candidate does: the 2D ensemble 1.603 against 1.613 ms per batched step;
// in binary64, with its costly divisions and square roots, the FP64 3D dam
// break took 0.836 against 0.800 ms per step), so the result depends
// neither on DEPTH nor on WIDE.
template <int D, bool SCAN, int DEPTH, bool SPLIT = true, bool WIDE = false, bool TAILW = false, typename Load,
          typename Body>
__device__ __forceinline__ void for_pairs_d(const SceneDev& s, const int32_t* __restrict__ cstart,
                                            const double* __restrict__ xall, const double* xi, int i, int self,
                                            const int32_t* __restrict__ list, int cnt, Load&& load, Body&& body) {
    static_assert(DEPTH == 1 || DEPTH == 2 || DEPTH == 4, "DEPTH divides a 4-entry chunk");
    if (SCAN) {
        // the overflow fallback inside a pass: one candidate at a time
        int c[3];
        cell_of<D>(s, xi, c);
        for_each_neighbour<D, 1>(s, cstart, xall, xi, c, self, [&](int j, double d2, const double* rv) {
            auto r = load(j);
            if constexpr (has_late<std::decay_t<Load>>::value) load.late(j, r);
            body(j, d2, rv, r);
        });
        return;
    }
    if (cnt <= 0) return;
    const size_t cap = (size_t)s.cap;
    const int nq = (cnt + 3) >> 2;
    // d^2 handed to the body for a group's slots past the count: their list
    // entries are earlier candidates of the same list (or index 0, the
    // buffer's initial entry), so their data are finite, and at 4 R^2 (twice
    // R, beyond every kernel's support) every body's exact rule rejects them
    const double beyond = 4.0 * s.R2;
    if constexpr (WIDE && DEPTH == 4) {
        // two chunks (8 candidates) per group, as in for_pairs_q
        int4 c0 = ld_chunk(list + list_at(cap, 0, i));
        int4 c1 = nq > 1 ? ld_chunk(list + list_at(cap, 4, i)) : c0;
        for (int q = 0; q < nq; q += 2) {
            int4 n0 = c0, n1 = c1;
            if (q + 2 < nq) n0 = ld_chunk(list + list_at(cap, 4 * (q + 2), i));
            if (q + 3 < nq) n1 = ld_chunk(list + list_at(cap, 4 * (q + 3), i));
            const int js[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            const int m = cnt - 4 * q;  // entries of these two chunks: min(8, m)
            double xj[8][3];
            decltype(load(0)) rec[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) load_x<D>(xall, js[u], xj[u]);
#pragma unroll
            for (int u = 0; u < 8; ++u) rec[u] = load(SPLIT ? after<D>(js[u], xj[u]) : js[u]);
            if constexpr (has_late<std::decay_t<Load>>::value) {
#pragma unroll
                for (int u = 0; u < 8; ++u) load.late(js[u], rec[u]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (TAILW || u < m) {
                    double rv[3];
                    const double d2 = dist2_exact<D>(xi, xj[u], rv);
                    body(js[u], !TAILW || u < m ? d2 : beyond, rv, rec[u]);
                }
            }
            c0 = n0;
            c1 = n1;
        }
        return;
    }
    int4 cur = ld_chunk(list + list_at(cap, 0, i));
    for (int q = 0; q < nq; ++q) {
        int4 nxt = cur;
        if (q + 1 < nq) nxt = ld_chunk(list + list_at(cap, 4 * (q + 1), i));
        const int js[4] = {cur.x, cur.y, cur.z, cur.w};
        const int m = cnt - 4 * q;  // entries of this chunk: min(4, m)
#pragma unroll
        for (int g = 0; g < 4; g += DEPTH) {
            if (g < m) {
                double xj[DEPTH][3];
                decltype(load(0)) rec[DEPTH];
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) load_x<D>(xall, js[g + u], xj[u]);
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) rec[u] = load(SPLIT ? after<D>(js[g + u], xj[u]) : js[g + u]);
                if constexpr (has_late<std::decay_t<Load>>::value) {
#pragma unroll
                    for (int u = 0; u < DEPTH; ++u) load.late(js[g + u], rec[u]);
                }
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) {
                    if (TAILW || g + u < m) {
                        double rv[3];
                        const double d2 = dist2_exact<D>(xi, xj[u], rv);
                        body(js[g + u], !TAILW || g + u < m ? d2 : beyond, rv, rec[u]);
                    }
                }
            }
        }
        cur = nxt;
    }
}

// Quantized positions (FluidSet::xq, FP32_BULK).
// (w: rho - rho0 in 3D, see FluidSet::xq)
template <int D>
__device__ __forceinline__ void store_xq(const SceneDev& s, const FluidSet& N, int i, const double* x, float w) {
    int32_t q[3];
    quantize(s.lo, s.inv_qu, D, x, q);
    if (D == 3)
        static_cast<int4*>(N.xq)[i] = make_int4(q[0], q[1], q[2], __float_as_int(w));
    else
        static_cast<int2*>(N.xq)[i] = make_int2(q[0], q[1]);
}

// q[0..2]: quantized position; q[3]: the w bits in 3D (0 in 2D)
template <int D>
__device__ __forceinline__ void load_q(const void* xq, int j, int32_t* q) {
    if (D == 3) {
        const int4 v = ldrec<int4>(xq, j);
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
        q[3] = v.w;
    } else {
        const int2 v = ldrec<int2>(xq, j);
        q[0] = v.x;
        q[1] = v.y;
        q[2] = 0;
        q[3] = 0;
    }
}

// p / rho^2 of a fluid particle from d = rho - rho0 in float, the equation of
// state of updatePressure! (SelfAction/UpdatePressure.jl:8-16) written in d
// so that the small difference rho / rho0 - 1 keeps its precision:
// (rho / rho0)^gamma - 1 = e (1 + u + ... + u^(gamma-1)), e = d / rho0, u = 1 + e;
// p = b e (...) + p0.  Also returns rho = rho0 + d.
// LIN: gamma is 1 and p0 is 0 (the linear equation of state of the
// examples), so p = (b / rho0) d with no loop.  GAMMA > 0: gamma is GAMMA,
// the sum unrolled (the same multiply-adds in the same order as the loop
// over s.gamma, without a uniform branch per term).
constexpr int kTaitGamma = 7;  // the exponent of the Tait equation of state most scenes use
template <bool LIN = false, int GAMMA = 0>
__device__ __forceinline__ float eos_pr2(const SceneDev& s, float d, float& rho) {
    float p;
    if (LIN) {
        p = s.clinf * d;
    } else {
        const float e = d * s.inv_rho0f;
        const float u = 1.0f + e;
        float g = 1.0f;
        if constexpr (GAMMA > 0) {
#pragma unroll
            for (int k = 1; k < GAMMA; ++k) g = g * u + 1.0f;
        } else {
            for (int k = 1; k < s.gamma; ++k) g = g * u + 1.0f;
        }
        p = s.bf * (e * g) + s.p0f;
    }
    rho = s.rho0f + d;
    const float ir = prcp(rho);
    return p * ir * ir;
}

// Pair geometry from quantized positions, in units of the quantum qu:
// rv = q_i - q_j (the integer difference is exact; its float conversion
// rounds at most once), d2 = |rv|^2 in float.  In metres these are rv qu and
// d2 qu^2 exactly (qu is a power of two), and the pair passes that take them
// scale their length constants by powers of qu instead (QScale), so every
// product rounds as it would in metres.  Against the FP64 rule d^2 <= R^2 on
// the binary64 positions, |d2 - d^2 / qu^2| stays below qbandq near R (two
// half-quanta per axis, the float roundings, R2q's own rounding; derived
// where qbandq is set), so a candidate outside the band is decided by the
// float distance and one inside it -- or with a position beyond the
// quantized range -- by the FP64 rule itself.  Returns whether j is a
// neighbour; rv and d2 are the float geometry of the pair in quanta.
// (getj: returns j, called only inside the band)
template <int D, typename J>
__device__ __forceinline__ bool qgeom_f(const SceneDev& s, const double* __restrict__ xall, const double* xi,
                                        const int32_t* qi, const int32_t* qj, J&& getj, float* rv, float& d2) {
#pragma unroll
    for (int k = 0; k < D; ++k) rv[k] = (float)(int32_t)((uint32_t)qi[k] - (uint32_t)qj[k]);
    if (D == 2) rv[2] = 0.0f;
    d2 = D == 3 ? rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2] : rv[0] * rv[0] + rv[1] * rv[1];
    bool in = d2 <= s.R2q;
    // (bitwise: the common path computes everything and takes one branch,
    // around the rare fallback)
    if ((qi[0] == kQInvalid) | (qj[0] == kQInvalid) | (fabsf(d2 - s.R2q) <= s.qbandq)) {
        double xj[3], rvd[3];
        load_x<D>(xall, getj(), xj);
        const double d2d = dist2_exact<D>(xi, xj, rvd);
        in = d2d <= s.R2;
#pragma unroll
        for (int k = 0; k < D; ++k) rv[k] = (float)(rvd[k] * s.inv_qu);
        d2 = (float)(d2d * (s.inv_qu * s.inv_qu));
    }
    return in;
}

template <int D>
__device__ __forceinline__ bool qgeom(const SceneDev& s, const double* __restrict__ xall, const double* xi,
                                      const int32_t* qi, const int32_t* qj, int j, float* rv, float& d2) {
    return qgeom_f<D>(s, xall, xi, qi, qj, [j] { return j; }, rv, d2);
}

// Length constants of the pair arithmetic in the units of its geometry:
// metres (binary64 positions) or quanta (qgeom).  With quanta, 1 / h is
// multiplied by qu, 0.01 h^2 divided by qu^2, and factors of r / (r^2 + 0.01 h^2)
// by 1 / qu: powers of two, so the products equal the metre ones.
template <typename T>
struct QScale {
    T inv_h, eta2, len;  // 1 / h, 0.01 h^2, and the factor of one r / (r^2 + 0.01 h^2)
    __device__ __forceinline__ QScale(const SceneDev& s, bool quanta)
        : inv_h((T)(quanta ? s.inv_h * s.qu : s.inv_h)),
          eta2((T)(quanta ? 0.01 * (s.h * s.h) * (s.inv_qu * s.inv_qu) : 0.01 * (s.h * s.h))),
          len((T)(quanta ? s.inv_qu : 1.0)) {}
};

// The list walk of for_pairs_d with the candidates' geometry from their
// quantized positions (one 16-byte or 8-byte load instead of the binary64
// position): body(j, in, d2, rv, record, w) with in = the FP64 rule's
// decision and w the candidate's FluidSet::xq w bits.  The slots of a
// group past the particle's count are walked too, with in = false: their
// list entries are earlier candidates of the same list (or particle 0, the
// buffer's initial entry), so every value is finite and bodies that weight
// their terms by in (or gate on it) add nothing for them, without a branch
// per candidate.
// Gather groups as in for_pairs_d (SPLIT: the positions, then the records).
template <int D, int DEPTH, bool SPLIT = true, typename Load, typename Body>
__device__ __forceinline__ void for_pairs_q(const SceneDev& s, const FluidSet& S, const double* xi, int i,
                                            const int32_t* __restrict__ list, int cnt, Load&& load, Body&& body) {
    static_assert(DEPTH == 1 || DEPTH == 2 || DEPTH == 4, "DEPTH divides a 4-entry chunk");
    if (cnt <= 0) return;
    int32_t qi[4];
    load_q<D>(S.xq, i, qi);
    const size_t cap = (size_t)s.cap;
    const int nq = (cnt + 3) >> 2;
    if constexpr (D == 3 && DEPTH == 4 && SPLIT) {
        // the 3D continuity walk at depth 4 takes two chunks (8 candidates)
        // per group: 0.208 ms against 0.219 on the 1M-particle dam break (in
        // 2D, with shorter lists, the batched ensemble measured 10% slower).
        // The binary64 walks of the continuity and Shepard passes do the
        // same in 3D (for_pairs_d, WIDE): 0.362 against 0.375 ms and 0.290
        // against 0.306.
        int4 c0 = ld_chunk(list + list_at(cap, 0, i));
        int4 c1 = nq > 1 ? ld_chunk(list + list_at(cap, 4, i)) : c0;
        for (int q = 0; q < nq; q += 2) {
            int4 n0 = c0, n1 = c1;
            if (q + 2 < nq) n0 = ld_chunk(list + list_at(cap, 4 * (q + 2), i));
            if (q + 3 < nq) n1 = ld_chunk(list + list_at(cap, 4 * (q + 3), i));
            const int js[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            const int m = cnt - 4 * q;  // entries of these two chunks: min(8, m)
            int32_t qj[8][4];
            decltype(load(0)) rec[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) load_q<D>(S.xq, js[u], qj[u]);
#pragma unroll
            for (int u = 0; u < 8; ++u) rec[u] = load(after<D>(js[u], qj[u]));
            if constexpr (has_late<std::decay_t<Load>>::value) {
#pragma unroll
                for (int u = 0; u < 8; ++u) load.late(js[u], rec[u]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                // slots past the count: weighted out, not branched around
                float rv[3], d2;
                const bool in = qgeom<D>(s, S.x, xi, qi, qj[u], js[u], rv, d2) & (u < m);
                body(js[u], in, d2, rv, rec[u], qj[u][3]);
            }
            c0 = n0;
            c1 = n1;
        }
        return;
    }
    int4 cur = ld_chunk(list + list_at(cap, 0, i));
    for (int q = 0; q < nq; ++q) {
        int4 nxt = cur;
        if (q + 1 < nq) nxt = ld_chunk(list + list_at(cap, 4 * (q + 1), i));
        const int js[4] = {cur.x, cur.y, cur.z, cur.w};
        const int m = cnt - 4 * q;  // entries of this chunk: min(4, m)
#pragma unroll
        for (int g = 0; g < 4; g += DEPTH) {
            if (g < m) {
                int32_t qj[DEPTH][4];
                decltype(load(0)) rec[DEPTH];
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) load_q<D>(S.xq, js[g + u], qj[u]);
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) rec[u] = load(SPLIT ? after<D>(js[g + u], qj[u]) : js[g + u]);
                if constexpr (has_late<std::decay_t<Load>>::value) {
#pragma unroll
                    for (int u = 0; u < DEPTH; ++u) load.late(js[g + u], rec[u]);
                }
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) {
                    // slots past the count: weighted out, not branched around
                    float rv[3], d2;
                    const bool in = qgeom<D>(s, S.x, xi, qi, qj[u], js[g + u], rv, d2) & (g + u < m);
                    body(js[g + u], in, d2, rv, rec[u], qj[u][3]);
                }
            }
        }
        cur = nxt;
    }
}

// Velocity record of a neighbour (continuity pass).
template <typename T>
struct ARec {
    T v[3];
    T m, gap;
};

// DPP moves (GFX9 controls): by n lanes toward higher lanes inside each
// 16-lane row (kDppRowShr + n), lane 15 of each row to the next row, lane 31
// to rows 2 and 3; lanes outside the row mask, or whose source lies outside
// their row, get 0.  Every lane of the wavefront must be active.
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Segmented inclusive sum over the 64 lanes of a wavefront, every lane
// active: lane p gets the sum of lanes head..p, head <= p being the first
// lane of p's segment (segments are runs of consecutive lanes).  The steps
// of a Hillis-Steele scan as DPP moves -- shifts by 1, 2, 4 and 8 inside the
// rows, then the broadcasts of lanes 15 and 31 -- each adding the moved value
// only where it comes from the same segment.  The order of the additions
// depends only on the segment's lanes, and so does the result.
template <typename T>
__device__ __forceinline__ T wave_segmented_sum(T v, int head) {
    const int lane = (int)(threadIdx.x & 63);
    T y;
    y = dpp_move<kDppRowShr + 1, 0xf>(v);
    v += lane - 1 >= head ? y : T(0);
    y = dpp_move<kDppRowShr + 2, 0xf>(v);
    v += lane - 2 >= head ? y : T(0);
    y = dpp_move<kDppRowShr + 4, 0xf>(v);
    v += lane - 4 >= head ? y : T(0);
    y = dpp_move<kDppRowShr + 8, 0xf>(v);
    v += lane - 8 >= head ? y : T(0);
    y = dpp_move<kDppRowBcast15, 0xa>(v);  // rows 1 and 3: lanes 15 and 47
    v += head < (lane & ~15) ? y : T(0);
    y = dpp_move<kDppRowBcast31, 0xc>(v);  // rows 2 and 3: lane 31
    v += lane >= 32 && head < 32 ? y : T(0);
    return v;
}

// Inclusive scan over a wavefront (64 lanes, every lane active): one
// segment, in DPP moves (LDS-free; a shuffle per step waits on LDS).
__device__ __forceinline__ int wave_inclusive_scan(int x) { return wave_segmented_sum(x, 0); }

// Orders LDS accesses of one wavefront's lanes to each other (its LDS
// instructions complete in program order; this keeps the compiler's order).
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Brick walk (FLAG_BRICK).  Bricks are blocks of bdim cells; a workgroup
// takes one brick at a time, stages the quantized positions and A records of
// every particle of the brick's region -- its cells widened by the stencil
// on each side, i.e. every cell a neighbour of a brick particle can be in --
// in LDS with contiguous loads (the region's cell rows along x are particle
// ranges of the cell order), and walks its particles' candidate lists with
// the candidates' data read from LDS.  The list rebuild records, beside each
// candidate j, its LDS slot: the region's rows (fixed y, z) are staged in
// row order, row r from slot rb[r], so slot = j - rs[r] + rb[r].
// ---------------------------------------------------------------------------
__device__ __forceinline__ int brick_of_cell(const SceneDev& s, int c) {
    const int cx = c % s.nc[0];
    const int r = c / s.nc[0];
    const int cy = r % s.nc[1];
    const int cz = r / s.nc[1];
    return cx / s.bdim[0] + s.nbk[0] * (cy / s.bdim[1] + s.nbk[1] * (cz / s.bdim[2]));
}

// Cell rows of one brick and of its region, in LDS.  Region row r (r < nr)
// is cells [xa, xb] of row (y0 + r % ny, z0 + r / ny): particles
// [rs[r], rs[r] + rb[r + 1] - rb[r]), staged from slot rb[r].  Brick row r
// (r < no) is cells [bx0, bx1] of row (by0 + r % BY, bz0 + r / BY):
// particles [os[r], os[r] + ob[r + 1] - ob[r]).
struct BrickTables {
    int rs[kBrickRows];
    int rb[kBrickRows + 1];
    int os[kBrickOwn];
    int ob[kBrickOwn + 1];
    int y0, z0, ny, nr, no;
};

// Slot map of a brick region for the list rebuild: row(cy, cz) before the
// candidates of a cell row, slot(j) for each of them.
struct BrickMap {
    const BrickTables* t;
    int base;
    __device__ __forceinline__ void row(int cy, int cz) {
        const int r = (cy - t->y0) + t->ny * (cz - t->z0);
        base = t->rb[r] - t->rs[r];
    }
    __device__ __forceinline__ int slot(int j) const { return j + base; }
};

// Largest r in [0, n) with a[r] <= v (a non-decreasing, a[0] <= v).
__device__ __forceinline__ int row_search(const int* a, int n, int v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Fills the tables of brick b (all threads of the workgroup; ends with a
// barrier).  Rows outside the grid are empty.
__device__ __forceinline__ void brick_tables(const SceneDev& s, int b, BrickTables& t) {
    const int BX = s.bdim[0], BY = s.bdim[1], BZ = s.bdim[2];
    const int mx = s.stencil[0], my = s.stencil[1], mz = s.stencil[2];
    const int bx = b % s.nbk[0], b1 = b / s.nbk[0];
    const int x0 = bx * BX, y0 = (b1 % s.nbk[1]) * BY, z0 = (b1 / s.nbk[1]) * BZ;
    const int ny = BY + 2 * my, nr = ny * (BZ + 2 * mz), no = BY * BZ;
    const int32_t* cs = s.cell_start;
    const int tid = (int)threadIdx.x;
    if (tid < nr) {
        const int cy = y0 - my + tid % ny, cz = z0 - mz + tid / ny;
        int st = 0, len = 0;
        if (cy >= 0 && cy < s.nc[1] && cz >= 0 && cz < s.nc[2]) {
            const int xa = x0 - mx > 0 ? x0 - mx : 0;
            const int xb = x0 + BX - 1 + mx < s.nc[0] - 1 ? x0 + BX - 1 + mx : s.nc[0] - 1;
            st = cs[cell_linear(s, xa, cy, cz)];
            len = cs[cell_linear(s, xb, cy, cz) + 1] - st;
        }
        t.rs[tid] = st;
        t.rb[tid + 1] = len;
    } else if (tid >= kBrickRows && tid < kBrickRows + no) {
        const int r = tid - kBrickRows;
        const int cy = y0 + r % BY, cz = z0 + r / BY;
        int st = 0, len = 0;
        if (cy < s.nc[1] && cz < s.nc[2]) {
            const int xb = x0 + BX - 1 < s.nc[0] - 1 ? x0 + BX - 1 : s.nc[0] - 1;
            st = cs[cell_linear(s, x0, cy, cz)];
            len = cs[cell_linear(s, xb, cy, cz) + 1] - st;
        }
        t.os[r] = st;
        t.ob[r + 1] = len;
    }
    if (tid == 0) {
        t.y0 = y0 - my;
        t.z0 = z0 - mz;
        t.ny = ny;
        t.nr = nr;
        t.no = no;
    }
    __syncthreads();
    // running sums: wave 0 the region's rows (two per lane), wave 1 the brick's
    const int lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        const int a = 2 * lane < nr ? t.rb[2 * lane + 1] : 0;
        const int c = 2 * lane + 1 < nr ? t.rb[2 * lane + 2] : 0;
        const int ex = wave_inclusive_scan(a + c) - (a + c);
        if (2 * lane < nr) t.rb[2 * lane + 1] = ex + a;
        if (2 * lane + 1 < nr) t.rb[2 * lane + 2] = ex + a + c;
        if (lane == 0) t.rb[0] = 0;
    } else if (wave == 1) {
        const int a = lane < no ? t.ob[lane + 1] : 0;
        const int in = wave_inclusive_scan(a);
        if (lane < no) t.ob[lane + 1] = in;
        if (lane == 0) t.ob[0] = 0;
    }
    __syncthreads();
}

// Stages the region of the current tables (t.rb[t.nr] particles, at most
// kBrickCap) into lq / la; ends with a barrier.
__device__ __forceinline__ void brick_stage(const FluidSet& S, const BrickTables& t, int4* lq, float4* la) {
    const int total = t.rb[t.nr];
    constexpr int U = 4;
    for (int k0 = (int)threadIdx.x; k0 < total; k0 += U * kBrickThreads) {
        int4 q[U];
        float4 a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * kBrickThreads;
            const int kk = k < total ? k : total - 1;
            const int r = row_search(t.rb, t.nr, kk);
            const int j = t.rs[r] + (kk - t.rb[r]);
            q[u] = ldrec<int4>(S.xq, j);
            a[u] = ldrec<float4>(S.A, j);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * kBrickThreads;
            if (k < total) {
                lq[k] = q[u];
                la[k] = a[u];
            }
        }
    }
    __syncthreads();
}

// Calls f(i) for each particle of the current brick (each thread its own).
template <typename F>
__device__ __forceinline__ void for_brick_particles(const BrickTables& t, F&& f) {
    const int n = t.ob[t.no];
    for (int q = (int)threadIdx.x; q < n; q += kBrickThreads) {
        const int r = row_search(t.ob, t.no, q);
        f(t.os[r] + (q - t.ob[r]));
    }
}

// 4-entry chunk of LDS slots (FLAG_BRICK), loaded non-temporally.
typedef unsigned short esph_nus4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ ushort4 ld_slots(const uint16_t* p) {
    const esph_nus4 v = __builtin_nontemporal_load(reinterpret_cast<const esph_nus4*>(p));
    return make_ushort4(v.x, v.y, v.z, v.w);
}

// The fluid candidates staged in LDS (FLAG_BRICK): quantized positions with
// their w bits, and A records (velocity, mass), by slot.
struct BrickStage {
    const int4* q;
    const float4* a;
};

// The list walk of for_pairs_q over a staged brick region: each candidate's
// quantized position and A record come from its LDS slot (vslot) instead of
// device memory, and its index (vnbr) is read only inside the rounding band,
// where the FP64 rule needs its binary64 position.  Same candidates, order
// and arithmetic as for_pairs_q: body(in, d2, rv, a, w).
template <int DEPTH, typename Body>
__device__ __forceinline__ void for_pairs_brick(const SceneDev& s, const FluidSet& S, const double* xi, int i,
                                                int cnt, const BrickStage& st, Body&& body) {
    static_assert(DEPTH == 1 || DEPTH == 2 || DEPTH == 4, "DEPTH divides a 4-entry chunk");
    if (cnt <= 0) return;
    int32_t qi[4];
    load_q<3>(S.xq, i, qi);
    const size_t cap = (size_t)s.cap;
    const int nq = (cnt + 3) >> 2;
    ushort4 cur = ld_slots(s.vslot + list_at(cap, 0, i));
    for (int q = 0; q < nq; ++q) {
        ushort4 nxt = cur;
        if (q + 1 < nq) nxt = ld_slots(s.vslot + list_at(cap, 4 * (q + 1), i));
        const int ks[4] = {cur.x, cur.y, cur.z, cur.w};
        const int m = cnt - 4 * q;  // entries of this chunk: min(4, m)
#pragma unroll
        for (int g = 0; g < 4; g += DEPTH) {
            if (g < m) {
                int4 Q[DEPTH];
                float4 A[DEPTH];
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) {
                    Q[u] = st.q[ks[g + u]];
                    A[u] = st.a[ks[g + u]];
                }
#pragma unroll
                for (int u = 0; u < DEPTH; ++u) {
                    if (g + u < m) {
                        const int k = 4 * q + g + u;
                        const int32_t qj[4] = {Q[u].x, Q[u].y, Q[u].z, Q[u].w};
                        float rv[3], d2;
                        const bool in = qgeom_f<3>(s, S.x, xi, qi, qj, [&] { return s.vnbr[list_at(cap, k, i)]; },
                                                   rv, d2);
                        ARec<float> a;
                        a.v[0] = A[u].x;
                        a.v[1] = A[u].y;
                        a.v[2] = A[u].z;
                        a.m = A[u].w;
                        a.gap = 0.0f;
                        body(in, d2, rv, a, qj[3]);
                    }
                }
            }
        }
        cur = nxt;
    }
}

// The SCAN instances walk the step's overflow list with a small fixed grid:
// kOvfTiles tiles of kBlock entries per scene, striding over the list.
// Calls f(i) for the particle(s) of this thread: the tile slot (main
// kernels) or, in the SCAN instances, every kOvfTiles*kBlock-th entry of the
// step's overflow list.
template <bool SCAN, typename F>
__device__ __forceinline__ void for_work(const SceneDev& s, const SceneDyn& dyn, const Tile& tile, int first,
                                         F&& f) {
    const int q0 = (SCAN ? tile.begin : first) + (int)threadIdx.x;
    if (SCAN) {
        for (int q = q0; q < dyn.novf; q += kOvfTiles * kBlock) f(s.ovf[q]);
    } else if (q0 < dyn.n_live) {
        f(q0);
    }
}

__device__ __forceinline__ bool overflowed(const SceneDev& s, int i) {
    return s.nvnb[i] > s.kff || (s.nw > 0 && s.nnbw[i] > s.kfw);
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
    V a;
    if constexpr (sizeof(V) == 16)
        a = ldrec<V>(A, j);
    else
        a = static_cast<const V*>(A)[j];
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

// load_rec in two groups: the A record (velocity, mass) with the positions,
// then the B record (density, pressure, gap, temperature) and kappa once A
// has arrived.  Measured on the force walk of the 1M-particle 3D dam break:
// binary64 0.643 ms against 0.653 with the positions first and both records
// after them; float (quantized positions) 0.343 against 0.353 in three
// groups (positions, A, B) and 0.395 for the binary64 positions with one
// group.
template <int D, typename T, bool THERMAL>
struct RecLoaderAB {
    const FluidSet& S;
    __device__ __forceinline__ Rec<T> operator()(int j) const {
        Rec<T> r;
        load_A<D, T>(S.A, j, r.v, r.m, r.gap);
        return r;
    }
    __device__ __forceinline__ void late(int j, Rec<T>& r) const {
        using V = typename Vec4<T>::type;
        if (D == 3)
            asm volatile("" : "+v"(j) : "v"(r.v[0]), "v"(r.v[1]), "v"(r.v[2]), "v"(r.m));
        else
            asm volatile("" : "+v"(j) : "v"(r.v[0]), "v"(r.v[1]), "v"(r.m), "v"(r.gap));
        V b;
        if constexpr (sizeof(V) == 16)
            b = ldrec<V>(S.B, j);
        else
            b = static_cast<const V*>(S.B)[j];
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
};

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

__device__ __forceinline__ bool check_due(const SceneDev& s, const SceneDyn& d) {
    return (s.flags & FLAG_CHECK) && s.kcheck > 0 && (d.step + 1) % s.kcheck == 0;
}
__device__ __forceinline__ bool shepard_due(const SceneDev& s, const SceneDyn& d) {
    return s.kreinit > 0 && (d.step + 1) % s.kreinit == 0;
}

#define ESPH_DISPATCH_K(KVAR, BODY)                                                 \
    switch (KVAR) {                                                                 \
        case ESPH_CUBIC_SPLINE: { constexpr int K = ESPH_CUBIC_SPLINE; BODY; } break; \
        case ESPH_GAUSSIAN: { constexpr int K = ESPH_GAUSSIAN; BODY; } break;         \
        case ESPH_WENDLAND_C2: { constexpr int K = ESPH_WENDLAND_C2; BODY; } break;   \
        case ESPH_WENDLAND_C4: { constexpr int K = ESPH_WENDLAND_C4; BODY; } break;   \
        default: return hipErrorInvalidValue;                                       \
    }

// Gather depth of the pair passes (neighbours whose loads are in flight
// together): 1, 2 or 4; chosen per bucket by the tuner.
#define ESPH_DISPATCH_DEPTH(DVAR, BODY)                     \
    switch (DVAR) {                                         \
        case 1: { constexpr int DP = 1; BODY; } break;      \
        case 4: { constexpr int DP = 4; BODY; } break;      \
        default: { constexpr int DP = 2; BODY; } break;     \
    }

static inline int grid_for(int ntiles) { return ntiles > 0 ? ntiles : 1; }

// The pair passes' launch over the bucket's particle tiles (overflowed
// particles scan their cells inside the same launch).
template <typename L>
static hipError_t launch_both(const Bucket& b, hipStream_t st, L&& launch) {
    if (b.ntiles <= 0) return hipSuccess;
    launch(std::false_type{}, b.tiles, b.ntiles);
    return hipGetLastError();
}

// Workgroup size of the pair passes (threads), chosen per bucket by the tuner.
#define ESPH_DISPATCH_BS(BVAR, BODY)                         \
    switch (BVAR) {                                          \
        case 64: { constexpr int BSZ = 64; BODY; } break;    \
        case 128: { constexpr int BSZ = 128; BODY; } break;  \
        default: { constexpr int BSZ = kBlock; BODY; } break; \
    }
#define SINGLE_ARG(...) __VA_ARGS__

}  // namespace esph
