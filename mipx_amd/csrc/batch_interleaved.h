// batch_interleaved.h -- lane-per-instance iteration kernels for batches that
// share one constraint matrix (branch-and-bound node relaxations).
//
// Within a round the active instances of one precision are packed into dense
// groups of 64 (one instance per wave lane) and their iterates are stored
// interleaved: V[group][row][lane].  A wave then handles one column (or row)
// for 64 instances at once: the matrix entries of that column are uniform
// (scalar loads, L2-resident), and every gather of y[row] / x_r[col] is one
// contiguous 64-element buffer load whose scalar offset carries the element
// index.  Per-instance data (step, primal weight, Halpern weight, iteration
// limit of the round) sits in per-lane registers, set when a wave's work
// moves to another group.
//
// Round structure (driven by BatchEngine):
//   kSlotLists      compacts the active instances of each precision into slots
//   kInterleave     per-instance storage -> interleaved (x, x0, l, u, y, y0)
//   N x kILPrimal / kILDual
//   kDeinterleave   interleaved -> per-instance storage (x, y); x^ and y^ are
//                   written straight to per-instance storage on the last
//                   iteration of the round
// after which the per-instance evaluation / control kernel runs as usual.
#pragma once

#include <cstddef>
#include <type_traits>

#include "batch_kernels.h"
#include "pdlp_refine.h"

namespace mpx {

constexpr int kLanes = 64;

struct ILane {
    double step, pw;
    int32_t inner, lim;
};

// Interleaved iterate storage of one precision: V[group][element][lane].
template <typename T>
struct IlBufs {
    T *x = nullptr, *x0 = nullptr, *xr = nullptr, *l = nullptr, *u = nullptr, *y = nullptr,
      *y0 = nullptr;
};

// Slot lists: slots[p][s] = instance id for list p (0 = plain FP32, 1 =
// FP64, 2 = FP32 refinement), counts[p] = number of slots, lanes[p][s] =
// its scalars.  Only instances whose `owner` equals `owner` take part (all
// of them when owner < 0).  One block, in instance order (deterministic).
__global__ __launch_bounds__(1024) void kSlotLists(const BInst* __restrict__ insts, int K,
                                                   int owner, int* __restrict__ slots32,
                                                   int* __restrict__ slots64,
                                                   int* __restrict__ slotsR,
                                                   ILane* __restrict__ lanes32,
                                                   ILane* __restrict__ lanes64,
                                                   ILane* __restrict__ lanesR,
                                                   int* __restrict__ counts) {
    __shared__ int part[3][1024];
    __shared__ int carry[3];
    if (threadIdx.x == 0) carry[0] = carry[1] = carry[2] = 0;
    __syncthreads();
    for (int base = 0; base < K; base += 1024) {
        const int k = base + threadIdx.x;
        int pl[3] = {0, 0, 0};
        if (k < K) {
            const BCtrl& c = *insts[k].ctrl;
            const bool mine = owner < 0 || insts[k].owner == owner;
            const bool act = mine && c.phase != kPhaseDone && c.round_iters > 0;
            pl[0] = act && c.phase == kPhaseFp32;
            pl[2] = act && c.phase == kPhaseRefine;
            pl[1] = act && !iteratesFp32(c.phase);
        }
        for (int p = 0; p < 3; ++p) part[p][threadIdx.x] = pl[p];
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int a[3];
            for (int p = 0; p < 3; ++p) a[p] = threadIdx.x >= off ? part[p][threadIdx.x - off] : 0;
            __syncthreads();
            for (int p = 0; p < 3; ++p) part[p][threadIdx.x] += a[p];
            __syncthreads();
        }
        if (k < K && (pl[0] || pl[1] || pl[2])) {
            const BCtrl& c = *insts[k].ctrl;
            const ILane ln{c.step, c.pw, c.inner, c.round_iters};
            const int p = pl[0] ? 0 : (pl[1] ? 1 : 2);
            const int s = carry[p] + part[p][threadIdx.x] - 1;
            (p == 0 ? slots32 : p == 1 ? slots64 : slotsR)[s] = k;
            (p == 0 ? lanes32 : p == 1 ? lanes64 : lanesR)[s] = ln;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int p = 0; p < 3; ++p) carry[p] += part[p][1023];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int p = 0; p < 3; ++p) counts[p] = carry[p];
}

// The counts kSlotLists will find at the start of the next round (plain
// FP32, FP64 or polish, refinement), taken after this round's evaluation
// with the same tests, so that the host can leave the kernels of an empty
// set out of the next round.
__global__ __launch_bounds__(1024) void kNextCounts(const BInst* __restrict__ insts, int K,
                                                    int owner, int* __restrict__ out) {
    int c[3] = {0, 0, 0};
    for (int base = 0; base < K; base += 1024) {
        const int k = base + threadIdx.x;
        int pl[3] = {0, 0, 0};
        if (k < K) {
            const BCtrl& ct = *insts[k].ctrl;
            const bool mine = owner < 0 || insts[k].owner == owner;
            const bool act = mine && ct.phase != kPhaseDone && ct.round_iters > 0;
            pl[0] = act && ct.phase == kPhaseFp32;
            pl[2] = act && ct.phase == kPhaseRefine;
            pl[1] = act && !iteratesFp32(ct.phase);
        }
        for (int p = 0; p < 3; ++p) c[p] += __syncthreads_count(pl[p]);
    }
    if (threadIdx.x == 0)
        for (int p = 0; p < 3; ++p) out[p] = c[p];
}

// Work items of a group pass: (group, block of 4 columns or rows).  Fixed
// grid, count from device memory; groups are dealt XCD by XCD (group g's
// items run on XCD g % 8 when there are at least 8 groups) so that a group's
// interleaved y / x_r slice stays in one L2.
template <typename F>
__device__ inline void forGroupItems(int nslots, int per_group, F&& f) {
    const int ngroups = (nslots + kLanes - 1) / kLanes;
    const int total = ngroups * per_group;
    if (ngroups < 8) {
        for (int t = blockIdx.x; t < total; t += gridDim.x) f(t / per_group, t % per_group);
        return;
    }
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8, nslot = gridDim.x / 8;
    for (int g = xcd; g < ngroups; g += 8)
        for (int r = slot; r < per_group; r += nslot) f(g, r);
}

// Per-instance -> interleaved for `rows` elements of each slot's vector
// (selected by `which`: 0 x/x0/l/u, 1 y/y0).  A block transposes a tile of
// 64 instances x 64 elements through LDS: coalesced reads per instance,
// coalesced writes per element.
template <typename T>
__device__ inline const BVecs<T>& vecsOf(const BInst& in);
template <>
__device__ inline const BVecs<float>& vecsOf<float>(const BInst& in) { return in.f; }
template <>
__device__ inline const BVecs<double>& vecsOf<double>(const BInst& in) { return in.d; }

template <typename T>
__device__ inline bool sameBits(T a, T b) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(unsigned long long, a) == __builtin_bit_cast(unsigned long long, b);
    else
        return __builtin_bit_cast(unsigned, a) == __builtin_bit_cast(unsigned, b);
}

// Byte offset, within an instance's BInst, of the pointer to source vector
// `a` of the vector set `which` (0 x, x0, l, u; 1 y, y0; 2 the refinement's
// c'; 3 its rl', ru', q; sets 2 and 3 are FP32 only).
template <typename T>
__device__ inline size_t ilSourceOffset(int which, int a) {
    using V = BVecs<T>;
    const size_t vb = std::is_same_v<T, float> ? offsetof(BInst, f) : offsetof(BInst, d);
    if (which == 0)
        return vb + (a == 0 ? offsetof(V, x) : a == 1 ? offsetof(V, x0) : a == 2 ? offsetof(V, l)
                                                                               : offsetof(V, u));
    if (which == 1) return vb + (a == 0 ? offsetof(V, y) : offsetof(V, y0));
    if (which == 2) return offsetof(BInst, rc);
    return a == 0 ? offsetof(BInst, rrl) : a == 1 ? offsetof(BInst, rru) : offsetof(BInst, rq);
}

template <typename T>
__device__ inline const T* ilSource(const BInst* in, size_t off) {
    return *reinterpret_cast<const T* const*>(reinterpret_cast<const char*>(in) + off);
}

// Rows of a 64 x 64 transpose tile that one thread moves (kBlock / 64 rows
// per pass).
constexpr int kTileRows = kLanes / (kBlock / kLanes);

// The instances of the tile rows a thread moves, resolved once per tile:
// the row is the same for a whole wave, so its slot and instance come from
// scalar loads; all of a thread's row loads are then issued together.
__device__ inline void tileRowInsts(const BInst* __restrict__ insts, const int* __restrict__ slots,
                                    int nslots, int g, const BInst* (&row)[kTileRows]) {
    const int r0 = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kLanes);
#pragma unroll
    for (int k = 0; k < kTileRows; ++k) {
        const int s = g * kLanes + r0 + k * (kBlock / kLanes);
        row[k] = s < nslots ? insts + __builtin_amdgcn_readfirstlane(slots[s]) : nullptr;
    }
}

// For which = 0 with `flags` given, also the bound flags of the group's
// 64-column tile: bit e - e0 set when some taking-part lane's l or u at
// column e differs (bitwise) from the owner's BL[e] / BU[e].
template <typename T>
__global__ __launch_bounds__(kBlock) void kInterleave(
    const BInst* __restrict__ insts, const int* __restrict__ slots, const int* __restrict__ count,
    int len, int which, T* __restrict__ A, T* __restrict__ B, T* __restrict__ C,
    T* __restrict__ D, const T* __restrict__ BL, const T* __restrict__ BU,
    unsigned long long* __restrict__ flags) {
    __shared__ T tile[kLanes][kLanes + 1];
    __shared__ unsigned long long own;
    const int nslots = *count;
    const int per_group = (len + kLanes - 1) / kLanes;
    const int nsrc = which == 0 ? 4 : which == 1 ? 2 : which == 2 ? 1 : 3;
    const bool mark = which == 0 && flags != nullptr;
    const int lane = static_cast<int>(threadIdx.x % kLanes);
    const int r0 = static_cast<int>(threadIdx.x / kLanes);
    forGroupItems(nslots, per_group, [&](int g, int tb) {
        const int e0 = tb * kLanes;
        // Cleared before the first barrier below; read after the last one.
        if (mark && threadIdx.x == 0) own = 0ull;
        const BInst* row[kTileRows];
        tileRowInsts(insts, slots, nslots, g, row);
        for (int a = 0; a < nsrc; ++a) {
            // load: row r = slot within the group, column = element offset
            const int e = e0 + lane;
            const size_t off = ilSourceOffset<T>(which, a);
            T v[kTileRows];
#pragma unroll
            for (int k = 0; k < kTileRows; ++k)
                v[k] = row[k] != nullptr && e < len ? ilSource<T>(row[k], off)[e] : T(0);
#pragma unroll
            for (int k = 0; k < kTileRows; ++k) tile[r0 + k * (kBlock / kLanes)][lane] = v[k];
            __syncthreads();
            T* dst = a == 0 ? A : a == 1 ? B : a == 2 ? C : D;
#pragma unroll
            for (int k = 0; k < kTileRows; ++k) {
                const int r = r0 + k * (kBlock / kLanes);
                const int er = e0 + r;
                if (er >= len) continue;
                const T w = tile[lane][r];
                dst[(static_cast<size_t>(g) * len + er) * kLanes + lane] = w;
                if (mark && a >= 2) {
                    const T b = (a == 2 ? BL : BU)[er];
                    const bool diff = g * kLanes + lane < nslots && !sameBits(w, b);
                    if (__ballot(diff) != 0ull && lane == 0) atomicOr(&own, 1ull << r);
                }
            }
            __syncthreads();
        }
        if (mark && threadIdx.x == 0) flags[static_cast<size_t>(g) * per_group + tb] = own;
    });
}

// Interleaved -> per-instance x (which = 0, len = n) or y (which = 1, len = m).
template <typename T>
__global__ __launch_bounds__(kBlock) void kDeinterleave(const BInst* __restrict__ insts,
                                                         const int* __restrict__ slots,
                                                         const int* __restrict__ count, int len,
                                                         int which, const T* __restrict__ A) {
    __shared__ T tile[kLanes][kLanes + 1];
    const int nslots = *count;
    const int per_group = (len + kLanes - 1) / kLanes;
    const int lane = static_cast<int>(threadIdx.x % kLanes);
    const int r0 = static_cast<int>(threadIdx.x / kLanes);
    forGroupItems(nslots, per_group, [&](int g, int tb) {
        const int e0 = tb * kLanes;
        const BInst* row[kTileRows];
        tileRowInsts(insts, slots, nslots, g, row);
#pragma unroll
        for (int k = 0; k < kTileRows; ++k) {
            const int e = e0 + r0 + k * (kBlock / kLanes);
            tile[lane][r0 + k * (kBlock / kLanes)] =
                e < len ? A[(static_cast<size_t>(g) * len + e) * kLanes + lane] : T(0);
        }
        __syncthreads();
        const int e = e0 + lane;
#pragma unroll
        for (int k = 0; k < kTileRows; ++k) {
            const int r = r0 + k * (kBlock / kLanes);
            if (row[k] != nullptr && e < len) {
                const BVecs<T>& vv = vecsOf<T>(*row[k]);
                (which == 0 ? vv.x : vv.y)[e] = tile[r][lane];
            }
        }
        __syncthreads();
    });
}

// Streamed per-lane vectors use non-temporal accesses so that they do not
// evict the gathered y / x_r slice of the group from L2.
template <typename T>
__device__ inline T ldNT(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ inline void stNT(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}

// Buffer-resource view of one group's interleaved vector (len elements x 64
// lanes of T).  Element i of lane l sits at byte i * 64 * sizeof(T) +
// l * sizeof(T): the element offset goes into the load's scalar offset and the
// lane offset is a constant vector register, so a gather costs one scalar
// shift and no vector address arithmetic.  Offsets stay below 2^31 bytes
// (ilFits); 0x00020000 is the gfx9 descriptor word for raw dword access.
constexpr int kIlMaxLen = 0x7fffffff / (kLanes * 8);  // 4,194,303
__host__ __device__ constexpr bool ilFits(int n, int m) { return n <= kIlMaxLen && m <= kIlMaxLen; }

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t groupRsrc(const T* base, int len) {
    const int bytes = len * kLanes * static_cast<int>(sizeof(T));
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), 0, bytes, 0x00020000);
}

template <typename T>
__device__ inline T gatherLane(__amdgpu_buffer_rsrc_t r, int voff, int i) {
    const int soff = i * kLanes * static_cast<int>(sizeof(T));
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Gathered dot product sum_q val[q] * V[idx[q]] of the lane's column of the
// group view `r` (lane byte offset `voff`), G entries in flight; the
// additions stay in storage order.  The last 1..G-1 entries go in one round
// trip: indices past the end repeat the last entry and are not added.  With
// G = 8 a C2 row (16 entries on average) takes two round trips and a column
// (6 entries) one: these kernels wait on memory three quarters of the time
// at the full 8 waves per SIMD, so entries in flight per wave set the pace.
template <typename T, int G = 8>
__device__ inline T laneDot(const int* __restrict__ idx, const T* __restrict__ val, int qb, int qe,
                            __amdgpu_buffer_rsrc_t r, int voff) {
    T acc = T(0);
    int q = qb;
    for (; q + G <= qe; q += G) {
        int ix[G];
        T a[G], v[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            ix[u] = idx[q + u];
            a[u] = val[q + u];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) v[u] = gatherLane<T>(r, voff, ix[u]);
#pragma unroll
        for (int u = 0; u < G; ++u) acc += a[u] * v[u];
    }
    if (q < qe) {
        int ix[G - 1];
        T a[G - 1], v[G - 1];
#pragma unroll
        for (int u = 0; u < G - 1; ++u) {
            const int qc = min(q + u, qe - 1);
            ix[u] = idx[qc];
            a[u] = val[qc];
        }
#pragma unroll
        for (int u = 0; u < G - 1; ++u) v[u] = gatherLane<T>(r, voff, ix[u]);
#pragma unroll
        for (int u = 0; u < G - 1; ++u)
            if (q + u < qe) acc += a[u] * v[u];
    }
    return acc;
}

// Gathered dot products acc[c] = sum_q val[q] * V[idx[q]] over the entries
// [qb[c], qe[c]) of CW columns (rows) at once: entries go in chunks of four
// per column, so up to 4 * CW gathers are in flight and a short tail costs
// one round trip, not one per entry.  Loads are unconditional (an index past
// a column's end repeats its last entry; an empty column reads entry 0) and
// each column's additions stay in storage order (an entry past the end is
// skipped, not added as zero).  Callers guarantee nnz > 0 when any column has
// entries.
template <typename T, int CW>
__device__ inline void laneDotN(const int* __restrict__ idx, const T* __restrict__ val,
                                const int (&qb)[CW], const int (&qe)[CW],
                                __amdgpu_buffer_rsrc_t r, int voff, T (&acc)[CW]) {
    int most = 0;
#pragma unroll
    for (int c = 0; c < CW; ++c) most = max(most, qe[c] - qb[c]);
    for (int t = 0; t < most; t += 4) {
        int ix[CW][4];
        T a[CW][4], v[CW][4];
#pragma unroll
        for (int c = 0; c < CW; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = qb[c] + t + u;
                const int qc = q < qe[c] ? q : qe[c] > qb[c] ? qe[c] - 1 : 0;
                ix[c][u] = idx[qc];
                a[c][u] = val[qc];
            }
#pragma unroll
        for (int c = 0; c < CW; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u) v[c][u] = gatherLane<T>(r, voff, ix[c][u]);
#pragma unroll
        for (int c = 0; c < CW; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (qb[c] + t + u < qe[c]) acc[c] += a[c][u] * v[c][u];
    }
}

// A lane's per-iteration scalars in the iteration precision: the round's
// last iteration, the Halpern weight k/(k+1) of iteration `kiter` (formed in
// FP64, then rounded), step size and primal weight.
template <typename T>
__device__ inline void laneScalars(const ILane& ln, int kiter, int& lim, T& lam, T& step, T& pw) {
    lim = ln.lim;
    const int kk = ln.inner + kiter;
    lam = static_cast<T>(static_cast<double>(kk) / static_cast<double>(kk + 1));
    step = static_cast<T>(ln.step);
    pw = static_cast<T>(ln.pw);
}

// CW = columns (rows) per wave in the iteration kernels: their epilogue
// operands and gathers are issued together so that several memory round trips
// are in flight per wave; each column still sums its entries in storage
// order, so every CW gives bit-identical results and the choice is left to
// the online autotuner.

// Primal half step: each wave handles kCW consecutive columns for the 64
// instances of its group (4 waves per block).  Shared data: CSC structure and
// values, c, tau_base.
// REFINE: the refinement's correction phase, whose objective c' is per lane
// (interleaved IC) instead of the matrix's shared c.
template <typename T, int kCW, bool REFINE = false>
__global__ __launch_bounds__(kBlock) void kILPrimal(
    const BInst* __restrict__ insts, const int* __restrict__ slots, const int* __restrict__ count,
    const ILane* __restrict__ lanes, int n, int m, const int* __restrict__ cs,
    const int* __restrict__ ri, const T* __restrict__ valt, const T* __restrict__ c,
    const T* __restrict__ tau_base, const T* __restrict__ IY, T* __restrict__ IX,
    const T* __restrict__ IX0, T* __restrict__ IXR, const T* __restrict__ IL,
    const T* __restrict__ IU, const T* __restrict__ BL, const T* __restrict__ BU,
    const unsigned long long* __restrict__ bflags, int kiter, const T* __restrict__ IC = nullptr) {
    const int nslots = *count;
    const int per_group = (n + 4 * kCW - 1) / (4 * kCW);
    const int ntiles = (n + kLanes - 1) / kLanes;
    // A lane's scalars are the same for every column of its group: loaded,
    // and the Halpern weight k/(k+1) formed, when the group changes (work
    // items come group by group).
    int curg = -1, lim = 0;
    T lam = T(0), step = T(1), pw = T(1);
    T* xh = nullptr;  // x^ of a lane whose round ends at this iteration
    forGroupItems(nslots, per_group, [&](int g, int jb) {
        // Wave-uniform column: in a scalar register, so the column's matrix
        // entries, c and tau come through the scalar cache.
        const int j0 = __builtin_amdgcn_readfirstlane(
            (jb * 4 + static_cast<int>(threadIdx.x) / kLanes) * kCW);
        const int lane = static_cast<int>(threadIdx.x) % kLanes;
        if (j0 >= n) return;
        const int s = g * kLanes + lane;
        if (g != curg) {
            curg = g;
            laneScalars(s < nslots ? lanes[s] : ILane{1.0, 1.0, 0, 0}, kiter, lim, lam, step, pw);
            // Loaded here, not beside the column's update: a conditional load
            // there would make every column wait for all of its loads.
            xh = s < nslots && kiter == lim ? vecsOf<T>(insts[slots[s]]).xh : nullptr;
        }
        if (kiter > lim) return;  // lane idle this iteration
        const __amdgpu_buffer_rsrc_t ry = groupRsrc(IY + static_cast<size_t>(g) * m * kLanes, m);
        const int voff = lane * static_cast<int>(sizeof(T));
        const size_t base = (static_cast<size_t>(g) * n + j0) * kLanes + lane;
        T xv[kCW], x0v[kCW], lv[kCW], uv[kCW], acc[kCW], cv[kCW];
        int qb[kCW], qe[kCW];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const bool in = j0 + cc < n;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            xv[cc] = in ? ldNT(IX + o) : T(0);
            x0v[cc] = in ? ldNT(IX0 + o) : T(0);
            if constexpr (REFINE) cv[cc] = in ? ldNT(IC + o) : T(0);
        }
        // FP64: the kCW columns lie in one 64-column tile (kCW divides 64); a
        // column whose bounds match the owner's in every lane reads the
        // owner's value (all lanes load one cached address): primal 365 ->
        // 346 us on C2.  FP32 streams its own copies: with the flags its
        // primal measured slower (180 -> 189 us, 193 us with each group's
        // flags staged in LDS instead of a scalar load per column).
        unsigned long long fw = ~0ull;
        if constexpr (sizeof(T) == 8) fw = bflags[static_cast<size_t>(g) * ntiles + (j0 >> 6)];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const bool in = j0 + cc < n;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            if constexpr (sizeof(T) == 8) {
                const bool own = (fw >> ((j0 + cc) & 63)) & 1ull;
                lv[cc] = in ? *(own ? IL + o : BL + j0 + cc) : T(0);
                uv[cc] = in ? *(own ? IU + o : BU + j0 + cc) : T(0);
            } else {
                lv[cc] = in ? ldNT(IL + o) : T(0);
                uv[cc] = in ? ldNT(IU + o) : T(0);
            }
            qb[cc] = in ? cs[j0 + cc] : 0;
            qe[cc] = in ? cs[j0 + cc + 1] : 0;
            acc[cc] = T(0);
        }
        if constexpr (kCW == 1) acc[0] = laneDot(ri, valt, qb[0], qe[0], ry, voff);
        else laneDotN<T, kCW>(ri, valt, qb, qe, ry, voff, acc);
        const bool last = kiter == lim;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const int j = j0 + cc;
            if (j >= n) break;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            const T tau = step * tau_base[j] / pw;
            T cj;
            if constexpr (REFINE) cj = cv[cc];
            else cj = c[j];
            const T t = xv[cc] - tau * (cj + acc[cc]);
            const T h = clampRef(t, lv[cc], uv[cc]);
            const T r = T(2) * h - xv[cc];
            IXR[o] = r;  // gathered by the dual step: keep it cacheable
            stNT(IX + o, lam * r + (T(1) - lam) * x0v[cc]);
            if (last) xh[j] = h;
        }
    });
}

// REFINE: per-lane row data of the refinement (interleaved IRL, IRU, IQ)
// and the shifted prox of pdlp_refine.h.
template <typename T, int kCW, bool REFINE = false>
__global__ __launch_bounds__(kBlock) void kILDual(
    const BInst* __restrict__ insts, const int* __restrict__ slots, const int* __restrict__ count,
    const ILane* __restrict__ lanes, int n, int m, const int* __restrict__ rs,
    const int* __restrict__ ci, const T* __restrict__ val, const T* __restrict__ sigma_base,
    const T* __restrict__ rl, const T* __restrict__ ru, const T* __restrict__ IXR,
    T* __restrict__ IY, const T* __restrict__ IY0, int kiter,
    const T* __restrict__ IRL = nullptr, const T* __restrict__ IRU = nullptr,
    const T* __restrict__ IQ = nullptr) {
    const int nslots = *count;
    const int per_group = (m + 4 * kCW - 1) / (4 * kCW);
    int curg = -1, lim = 0;
    T lam = T(0), step = T(1), pw = T(1);
    T* yh = nullptr;  // y^ of a lane whose round ends at this iteration
    forGroupItems(nslots, per_group, [&](int g, int ib) {
        const int i0 = __builtin_amdgcn_readfirstlane(
            (ib * 4 + static_cast<int>(threadIdx.x) / kLanes) * kCW);
        const int lane = static_cast<int>(threadIdx.x) % kLanes;
        if (i0 >= m) return;
        const int s = g * kLanes + lane;
        if (g != curg) {
            curg = g;
            laneScalars(s < nslots ? lanes[s] : ILane{1.0, 1.0, 0, 0}, kiter, lim, lam, step, pw);
            yh = s < nslots && kiter == lim ? vecsOf<T>(insts[slots[s]]).yh : nullptr;
        }
        if (kiter > lim) return;
        const __amdgpu_buffer_rsrc_t rx = groupRsrc(IXR + static_cast<size_t>(g) * n * kLanes, n);
        const int voff = lane * static_cast<int>(sizeof(T));
        const size_t base = (static_cast<size_t>(g) * m + i0) * kLanes + lane;
        T yv[kCW], y0v[kCW], acc[kCW], rlv[kCW], ruv[kCW], qv[kCW];
        int qb[kCW], qe[kCW];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const bool in = i0 + cc < m;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            yv[cc] = in ? IY[o] : T(0);  // y is gathered by the next primal step
            y0v[cc] = in ? ldNT(IY0 + o) : T(0);
            if constexpr (REFINE) {
                rlv[cc] = in ? ldNT(IRL + o) : T(0);
                ruv[cc] = in ? ldNT(IRU + o) : T(0);
                qv[cc] = in ? ldNT(IQ + o) : T(0);
            }
            qb[cc] = in ? rs[i0 + cc] : 0;
            qe[cc] = in ? rs[i0 + cc + 1] : 0;
            acc[cc] = T(0);
        }
        if constexpr (kCW == 1) acc[0] = laneDot(ci, val, qb[0], qe[0], rx, voff);
        else laneDotN<T, kCW>(ci, val, qb, qe, rx, voff, acc);
        const bool last = kiter == lim;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const int i = i0 + cc;
            if (i >= m) break;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            const T sig = step * sigma_base[i] * pw;
            const T v = yv[cc] + sig * acc[cc];
            T h;
            if constexpr (REFINE) h = dualProxShifted(v, sig, rlv[cc], ruv[cc], qv[cc]);
            else h = dualProx(v, sig, rl[i], ru[i]);
            const T r = T(2) * h - yv[cc];
            IY[o] = lam * r + (T(1) - lam) * y0v[cc];
            if (last) yh[i] = h;
        }
    });
}

}  // namespace mpx
