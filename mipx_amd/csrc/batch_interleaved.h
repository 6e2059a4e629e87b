// batch_interleaved.h -- lane-per-instance iteration kernels for batches that
// share one constraint matrix (branch-and-bound node relaxations).
//
// Within a round the active instances of one precision are packed into dense
// groups of 64 (one instance per wave lane) and their iterates are stored
// interleaved: V[group][row][lane].  A wave then handles one column (or row)
// for 64 instances at once: the matrix entries of that column are uniform
// (scalar loads, L2-resident), and every gather of y[row] / x_r[col] is one
// contiguous 64-element load.  Per-instance data (primal weight, inner count,
// iteration limit of the round) sits in per-lane registers.
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

#include "batch_kernels.h"

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

// Slot lists: slots[p][s] = instance id for precision p (0 = FP32, 1 = FP64),
// counts[p] = number of slots, lanes[p][s] = its scalars.  Only instances
// whose `owner` equals `owner` take part (all of them when owner < 0).  One
// block, in instance order (deterministic).
__global__ __launch_bounds__(1024) void kSlotLists(const BInst* __restrict__ insts, int K,
                                                   int owner, int* __restrict__ slots32,
                                                   int* __restrict__ slots64,
                                                   ILane* __restrict__ lanes32,
                                                   ILane* __restrict__ lanes64,
                                                   int* __restrict__ counts) {
    __shared__ int part[2][1024];
    __shared__ int carry[2];
    if (threadIdx.x == 0) carry[0] = carry[1] = 0;
    __syncthreads();
    for (int base = 0; base < K; base += 1024) {
        const int k = base + threadIdx.x;
        int p32 = 0, p64 = 0;
        if (k < K) {
            const BCtrl& c = *insts[k].ctrl;
            const bool mine = owner < 0 || insts[k].owner == owner;
            const bool act = mine && c.phase != kPhaseDone && c.round_iters > 0;
            p32 = act && c.phase == kPhaseFp32;
            p64 = act && c.phase != kPhaseFp32;
        }
        part[0][threadIdx.x] = p32;
        part[1][threadIdx.x] = p64;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int a0 = threadIdx.x >= off ? part[0][threadIdx.x - off] : 0;
            const int a1 = threadIdx.x >= off ? part[1][threadIdx.x - off] : 0;
            __syncthreads();
            part[0][threadIdx.x] += a0;
            part[1][threadIdx.x] += a1;
            __syncthreads();
        }
        if (k < K && (p32 || p64)) {
            const BCtrl& c = *insts[k].ctrl;
            const ILane ln{c.step, c.pw, c.inner, c.round_iters};
            if (p32) {
                const int s = carry[0] + part[0][threadIdx.x] - 1;
                slots32[s] = k;
                lanes32[s] = ln;
            } else {
                const int s = carry[1] + part[1][threadIdx.x] - 1;
                slots64[s] = k;
                lanes64[s] = ln;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            carry[0] += part[0][1023];
            carry[1] += part[1][1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[0] = carry[0];
        counts[1] = carry[1];
    }
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
__global__ __launch_bounds__(kBlock) void kInterleave(const BInst* __restrict__ insts,
                                                       const int* __restrict__ slots,
                                                       const int* __restrict__ count, int len,
                                                       int which, T* __restrict__ A,
                                                       T* __restrict__ B, T* __restrict__ C,
                                                       T* __restrict__ D) {
    __shared__ T tile[kLanes][kLanes + 1];
    const int nslots = *count;
    const int per_group = (len + kLanes - 1) / kLanes;
    const int nsrc = which == 0 ? 4 : 2;
    forGroupItems(nslots, per_group, [&](int g, int tb) {
        const int e0 = tb * kLanes;
        for (int a = 0; a < nsrc; ++a) {
            // load: row r = slot within the group, column = element offset
            for (int r = threadIdx.x / kLanes; r < kLanes; r += kBlock / kLanes) {
                const int s = g * kLanes + r;
                const int e = e0 + static_cast<int>(threadIdx.x % kLanes);
                T v = T(0);
                if (s < nslots && e < len) {
                    const BVecs<T>& vv = vecsOf<T>(insts[slots[s]]);
                    const T* src = which == 0 ? (a == 0 ? vv.x : a == 1 ? vv.x0 : a == 2 ? vv.l : vv.u)
                                              : (a == 0 ? vv.y : vv.y0);
                    v = src[e];
                }
                tile[r][threadIdx.x % kLanes] = v;
            }
            __syncthreads();
            T* dst = a == 0 ? A : a == 1 ? B : a == 2 ? C : D;
            for (int r = threadIdx.x / kLanes; r < kLanes; r += kBlock / kLanes) {
                const int e = e0 + r;
                const int lane = static_cast<int>(threadIdx.x % kLanes);
                if (e < len) dst[(static_cast<size_t>(g) * len + e) * kLanes + lane] = tile[lane][r];
            }
            __syncthreads();
        }
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
    forGroupItems(nslots, per_group, [&](int g, int tb) {
        const int e0 = tb * kLanes;
        for (int r = threadIdx.x / kLanes; r < kLanes; r += kBlock / kLanes) {
            const int e = e0 + r;
            const int lane = static_cast<int>(threadIdx.x % kLanes);
            tile[lane][r] = e < len ? A[(static_cast<size_t>(g) * len + e) * kLanes + lane] : T(0);
        }
        __syncthreads();
        for (int r = threadIdx.x / kLanes; r < kLanes; r += kBlock / kLanes) {
            const int s = g * kLanes + r;
            const int e = e0 + static_cast<int>(threadIdx.x % kLanes);
            if (s < nslots && e < len) {
                const BVecs<T>& vv = vecsOf<T>(insts[slots[s]]);
                (which == 0 ? vv.x : vv.y)[e] = tile[r][threadIdx.x % kLanes];
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

// Gathered dot product sum_q val[q] * V[idx[q] * 64] with four entries in
// flight; the additions stay in storage order.
template <typename T>
__device__ inline T laneDot(const int* __restrict__ idx, const T* __restrict__ val, int qb, int qe,
                            const T* __restrict__ V) {
    T acc = T(0);
    int q = qb;
    for (; q + 4 <= qe; q += 4) {
        const int i0 = idx[q], i1 = idx[q + 1], i2 = idx[q + 2], i3 = idx[q + 3];
        const T a0 = val[q], a1 = val[q + 1], a2 = val[q + 2], a3 = val[q + 3];
        const T v0 = V[static_cast<size_t>(i0) * kLanes], v1 = V[static_cast<size_t>(i1) * kLanes];
        const T v2 = V[static_cast<size_t>(i2) * kLanes], v3 = V[static_cast<size_t>(i3) * kLanes];
        acc += a0 * v0;
        acc += a1 * v1;
        acc += a2 * v2;
        acc += a3 * v3;
    }
    for (; q < qe; ++q) acc += val[q] * V[static_cast<size_t>(idx[q]) * kLanes];
    return acc;
}

// Gathered dot products acc[c] = sum_q val[q] * V[idx[q] * 64] over the
// entries [qb[c], qe[c]) of CW columns (rows) at once: entries go in chunks
// of four per column, so up to 4 * CW gathers are in flight and a short tail
// costs one round trip, not one per entry.  Each column's additions stay in
// storage order (an entry past the end is skipped, not added as zero).
// (Measured slower than laneDot for one column per wave; used for CW > 1.)
template <typename T, int CW>
__device__ inline void laneDotN(const int* __restrict__ idx, const T* __restrict__ val,
                                const int (&qb)[CW], const int (&qe)[CW],
                                const T* __restrict__ V, T (&acc)[CW]) {
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
                const bool ok = q < qe[c];
                ix[c][u] = ok ? idx[q] : 0;
                a[c][u] = ok ? val[q] : T(0);
            }
#pragma unroll
        for (int c = 0; c < CW; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u) v[c][u] = V[static_cast<size_t>(ix[c][u]) * kLanes];
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
template <typename T, int kCW>
__global__ __launch_bounds__(kBlock) void kILPrimal(
    const BInst* __restrict__ insts, const int* __restrict__ slots, const int* __restrict__ count,
    const ILane* __restrict__ lanes, int n, int m, const int* __restrict__ cs,
    const int* __restrict__ ri, const T* __restrict__ valt, const T* __restrict__ c,
    const T* __restrict__ tau_base, const T* __restrict__ IY, T* __restrict__ IX,
    const T* __restrict__ IX0, T* __restrict__ IXR, const T* __restrict__ IL,
    const T* __restrict__ IU, int kiter) {
    const int nslots = *count;
    const int per_group = (n + 4 * kCW - 1) / (4 * kCW);
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
        const T* Y = IY + static_cast<size_t>(g) * m * kLanes + lane;
        const size_t base = (static_cast<size_t>(g) * n + j0) * kLanes + lane;
        T xv[kCW], x0v[kCW], lv[kCW], uv[kCW], acc[kCW];
        int qb[kCW], qe[kCW];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const bool in = j0 + cc < n;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            xv[cc] = in ? ldNT(IX + o) : T(0);
            x0v[cc] = in ? ldNT(IX0 + o) : T(0);
            lv[cc] = in ? ldNT(IL + o) : T(0);
            uv[cc] = in ? ldNT(IU + o) : T(0);
            qb[cc] = in ? cs[j0 + cc] : 0;
            qe[cc] = in ? cs[j0 + cc + 1] : 0;
            acc[cc] = T(0);
        }
        if constexpr (kCW == 1) acc[0] = laneDot(ri, valt, qb[0], qe[0], Y);
        else laneDotN<T, kCW>(ri, valt, qb, qe, Y, acc);
        const bool last = kiter == lim;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const int j = j0 + cc;
            if (j >= n) break;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            const T tau = step * tau_base[j] / pw;
            const T t = xv[cc] - tau * (c[j] + acc[cc]);
            const T h = clampRef(t, lv[cc], uv[cc]);
            const T r = T(2) * h - xv[cc];
            IXR[o] = r;  // gathered by the dual step: keep it cacheable
            stNT(IX + o, lam * r + (T(1) - lam) * x0v[cc]);
            if (last) xh[j] = h;
        }
    });
}

template <typename T, int kCW>
__global__ __launch_bounds__(kBlock) void kILDual(
    const BInst* __restrict__ insts, const int* __restrict__ slots, const int* __restrict__ count,
    const ILane* __restrict__ lanes, int n, int m, const int* __restrict__ rs,
    const int* __restrict__ ci, const T* __restrict__ val, const T* __restrict__ sigma_base,
    const T* __restrict__ rl, const T* __restrict__ ru, const T* __restrict__ IXR,
    T* __restrict__ IY, const T* __restrict__ IY0, int kiter) {
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
        const T* X = IXR + static_cast<size_t>(g) * n * kLanes + lane;
        const size_t base = (static_cast<size_t>(g) * m + i0) * kLanes + lane;
        T yv[kCW], y0v[kCW], acc[kCW];
        int qb[kCW], qe[kCW];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const bool in = i0 + cc < m;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            yv[cc] = in ? IY[o] : T(0);  // y is gathered by the next primal step
            y0v[cc] = in ? ldNT(IY0 + o) : T(0);
            qb[cc] = in ? rs[i0 + cc] : 0;
            qe[cc] = in ? rs[i0 + cc + 1] : 0;
            acc[cc] = T(0);
        }
        if constexpr (kCW == 1) acc[0] = laneDot(ci, val, qb[0], qe[0], X);
        else laneDotN<T, kCW>(ci, val, qb, qe, X, acc);
        const bool last = kiter == lim;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            const int i = i0 + cc;
            if (i >= m) break;
            const size_t o = base + static_cast<size_t>(cc) * kLanes;
            const T sig = step * sigma_base[i] * pw;
            const T v = yv[cc] + sig * acc[cc];
            const T h = dualProx(v, sig, rl[i], ru[i]);
            const T r = T(2) * h - yv[cc];
            IY[o] = lam * r + (T(1) - lam) * y0v[cc];
            if (last) yh[i] = h;
        }
    });
}

}  // namespace mpx
