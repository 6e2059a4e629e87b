// pdlp_kernels.h -- hand-written CDNA4 kernels of the single-LP PDLP path.
//
// Layout in HBM (one LP): A in CSR (row_starts[m+1], col_indices[nnz],
// values[nnz]) and A' in CSR = A in CSC (col_starts[n+1], row_indices[nnz],
// values_t[nnz], row indices increasing inside each column, like the
// reference's buildTransposeCSR, src/lp/pdlp.cpp:671-707).  Vectors are
// contiguous arrays.  A row (or column) is handled by an aligned group of G
// lanes of one wave64: the G lanes stream the row's entries with unit stride
// (coalesced), keep a private partial sum and finish with a butterfly.  G = 1
// sums each row sequentially in storage order, which is the reference CPU's
// order (sparse_matrix.cpp:181-208) and makes the result bit-identical.
//
// Per PDHG iteration two kernels run, each an SpMV with the Halpern update
// fused into its epilogue (reference: cuSPARSE SpMV + primalHalpernStepKernel
// / dualHalpernStepKernel, pdlp_kernels.cu:46-157):
//   primalStep:  x^ = clamp(x - tau (c + A'y), l, u); x_r = 2x^ - x;
//                x = lam x_r + (1 - lam) x0                   (one group per column)
//   dualStep:    v = y + sigma A x_r; y^ = prox_rowtype(v); y_r = 2y^ - y;
//                y = lam y_r + (1 - lam) y0                   (one group per row)
// with tau = step tau_base / w, sigma = step sigma_base w and
// lam = (inner + k) / (inner + k + 1) (src/lp/pdlp.cpp:986-1071).  x^ and y^
// are only stored on the last iteration of a block, where the evaluation
// reads them; y_r is never stored (the reference never reads it back).
#pragma once

#include "device_common.h"
#include "pdlp_refine.h"

namespace mpx {

// Scalars the iteration kernels read through a pointer, so that one captured
// hipGraph survives restarts (new step / primal weight / inner count).
struct IterScalars {
    double step;
    double primal_weight;
    int32_t inner;
    int32_t pad;
};

template <int G, typename T, typename TX>
__device__ inline T groupDot(const int* __restrict__ starts, const int* __restrict__ idx,
                             const T* __restrict__ val, const TX* __restrict__ x, int row,
                             int lane) {
    const int b = starts[row], e = starts[row + 1];
    T acc = T(0);
    for (int k = b + lane; k < e; k += G) acc += val[k] * static_cast<T>(x[idx[k]]);
    return groupSum<G>(acc);
}

// ---- SpMV y = A x --------------------------------------------------------
template <int G, typename T>
__global__ __launch_bounds__(kBlock) void kSpmv(int rows, const int* __restrict__ starts,
                                                const int* __restrict__ idx,
                                                const T* __restrict__ val,
                                                const T* __restrict__ x, T* __restrict__ y) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int row = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    if (row >= rows) return;
    const T s = groupDot<G>(starts, idx, val, x, row, lane);
    if (lane == 0) y[row] = s;
}

// Group dot product over the entries [b, e) (G lanes, butterfly finish).
template <int G, typename T, typename TX>
__device__ inline T groupDotBE(int b, int e, const int* __restrict__ idx, const T* __restrict__ val,
                               const TX* __restrict__ x, int lane) {
    T acc = T(0);
    for (int k = b + lane; k < e; k += G) acc += val[k] * static_cast<T>(x[idx[k]]);
    return groupSum<G>(acc);
}

// Columns (rows) with more than `skip_above` entries are left to the
// one-wave-per-column kernels below, so a power-law tail of long columns
// does not hold a whole wave of short ones.
template <typename T, bool STORE_HAT>
__device__ inline void primalEpilogue(int j, T aty, T* __restrict__ x, const T* __restrict__ x0,
                                      T* __restrict__ xr, T* __restrict__ xh,
                                      const T* __restrict__ c, const T* __restrict__ tau_base,
                                      const T* __restrict__ lo, const T* __restrict__ hi,
                                      const IterScalars* __restrict__ sc, int k) {
    const int kk = sc->inner + k;
    const T lam = static_cast<T>(static_cast<double>(kk) / static_cast<double>(kk + 1));
    const T step = static_cast<T>(sc->step), pw = static_cast<T>(sc->primal_weight);
    const T tau = step * tau_base[j] / pw;
    const T xj = x[j];
    const T t = xj - tau * (c[j] + aty);
    const T h = clampRef(t, lo[j], hi[j]);
    const T r = T(2) * h - xj;
    if (STORE_HAT) xh[j] = h;
    xr[j] = r;
    x[j] = lam * r + (T(1) - lam) * x0[j];
}

// ---- primal half step (CSC SpMV + projection + reflection + Halpern) -------
template <int G, typename T, bool STORE_HAT>
__global__ __launch_bounds__(kBlock) void kPrimalStep(
    int n, const int* __restrict__ cstarts, const int* __restrict__ ridx,
    const T* __restrict__ valt, const T* __restrict__ y, T* __restrict__ x,
    const T* __restrict__ x0, T* __restrict__ xr, T* __restrict__ xh,
    const T* __restrict__ c, const T* __restrict__ tau_base, const T* __restrict__ lo,
    const T* __restrict__ hi, const IterScalars* __restrict__ sc, int k, int skip_above) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int j = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    if (j >= n) return;
    const int b = cstarts[j], e = cstarts[j + 1];
    if (e - b > skip_above) return;
    const T aty = groupDotBE<G>(b, e, ridx, valt, y, lane);
    if (lane != 0) return;
    primalEpilogue<T, STORE_HAT>(j, aty, x, x0, xr, xh, c, tau_base, lo, hi, sc, k);
}

// One wave per listed (long) column.
template <typename T, bool STORE_HAT>
__global__ __launch_bounds__(kBlock) void kPrimalStepLong(
    int count, const int* __restrict__ list, const int* __restrict__ cstarts,
    const int* __restrict__ ridx, const T* __restrict__ valt, const T* __restrict__ y,
    T* __restrict__ x, const T* __restrict__ x0, T* __restrict__ xr, T* __restrict__ xh,
    const T* __restrict__ c, const T* __restrict__ tau_base, const T* __restrict__ lo,
    const T* __restrict__ hi, const IterScalars* __restrict__ sc, int k) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(gid / kWave));
    const int lane = static_cast<int>(gid % kWave);
    if (w >= count) return;
    const int j = list[w];
    const T aty = groupDotBE<kWave>(cstarts[j], cstarts[j + 1], ridx, valt, y, lane);
    if (lane != 0) return;
    primalEpilogue<T, STORE_HAT>(j, aty, x, x0, xr, xh, c, tau_base, lo, hi, sc, k);
}

// Row-type prox of the dual step (src/lp/pdlp.cpp:1043-1067): equality,
// upper-only, lower-only, ranged; the row type follows from the bounds.
template <typename T>
__device__ inline T dualProx(T v, T sig, T rl, T ru) {
    if (rl == ru) return v - sig * rl;
    if (!isfinite(rl)) return maxRef(T(0), v - sig * ru);
    if (!isfinite(ru)) return minRef(T(0), v - sig * rl);
    if (v >= sig * ru) return v - sig * ru;
    if (v <= sig * rl) return v - sig * rl;
    return T(0);
}

// q (may be null): the refinement's base multipliers; with q the prox is
// the shifted one of pdlp_refine.h.
template <typename T, bool STORE_HAT>
__device__ inline void dualEpilogue(int i, T axr, T* __restrict__ y, const T* __restrict__ y0,
                                    T* __restrict__ yh, const T* __restrict__ sigma_base,
                                    const T* __restrict__ rl, const T* __restrict__ ru,
                                    const T* __restrict__ q, const IterScalars* __restrict__ sc,
                                    int k) {
    const int kk = sc->inner + k;
    const T lam = static_cast<T>(static_cast<double>(kk) / static_cast<double>(kk + 1));
    const T step = static_cast<T>(sc->step), pw = static_cast<T>(sc->primal_weight);
    const T sig = step * sigma_base[i] * pw;
    const T yi = y[i];
    const T v = yi + sig * axr;
    const T h = q ? dualProxShifted(v, sig, rl[i], ru[i], q[i]) : dualProx(v, sig, rl[i], ru[i]);
    const T r = T(2) * h - yi;
    if (STORE_HAT) yh[i] = h;
    y[i] = lam * r + (T(1) - lam) * y0[i];
}

// ---- dual half step (CSR SpMV of x_r + prox + reflection + Halpern) --------
template <int G, typename T, bool STORE_HAT>
__global__ __launch_bounds__(kBlock) void kDualStep(
    int m, const int* __restrict__ rstarts, const int* __restrict__ cidx,
    const T* __restrict__ val, const T* __restrict__ xr, T* __restrict__ y,
    const T* __restrict__ y0, T* __restrict__ yh, const T* __restrict__ sigma_base,
    const T* __restrict__ rl, const T* __restrict__ ru, const T* __restrict__ q,
    const IterScalars* __restrict__ sc, int k, int skip_above) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    if (i >= m) return;
    const int b = rstarts[i], e = rstarts[i + 1];
    if (e - b > skip_above) return;
    const T axr = groupDotBE<G>(b, e, cidx, val, xr, lane);
    if (lane != 0) return;
    dualEpilogue<T, STORE_HAT>(i, axr, y, y0, yh, sigma_base, rl, ru, q, sc, k);
}

template <typename T, bool STORE_HAT>
__global__ __launch_bounds__(kBlock) void kDualStepLong(
    int count, const int* __restrict__ list, const int* __restrict__ rstarts,
    const int* __restrict__ cidx, const T* __restrict__ val, const T* __restrict__ xr,
    T* __restrict__ y, const T* __restrict__ y0, T* __restrict__ yh,
    const T* __restrict__ sigma_base, const T* __restrict__ rl, const T* __restrict__ ru,
    const T* __restrict__ q, const IterScalars* __restrict__ sc, int k) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(gid / kWave));
    const int lane = static_cast<int>(gid % kWave);
    if (w >= count) return;
    const int i = list[w];
    const T axr = groupDotBE<kWave>(rstarts[i], rstarts[i + 1], cidx, val, xr, lane);
    if (lane != 0) return;
    dualEpilogue<T, STORE_HAT>(i, axr, y, y0, yh, sigma_base, rl, ru, q, sc, k);
}

// Row i's dot product so far after chunk `ch`, over the chunk's entries
// [b, e), for the group's lane 0 (part[i]: the sum after the earlier
// chunks).  With one lane (G = 1) the sum continues part[i] entry by entry,
// so the chunked product is the one-launch sum in storage order; with G
// lanes the chunk's own group sum is added to it, ((p_0 + p_1) + ...) +
// p_{C-1}.
template <int G, typename T, typename TX>
__device__ inline T chunkDot(int b, int e, const int* __restrict__ idx, const T* __restrict__ val,
                             const TX* __restrict__ x, int lane, int ch, const T* __restrict__ part,
                             int i) {
    if constexpr (G == 1) {
        T acc = ch > 0 ? part[i] : T(0);
        for (int k = b; k < e; ++k) acc += val[k] * static_cast<T>(x[idx[k]]);
        return acc;
    } else {
        const T acc = groupDotBE<G>(b, e, idx, val, x, lane);
        return ch > 0 ? part[i] + acc : acc;
    }
}

// Dual half step with A x_r in column chunks: chunk c of C holds, row by
// row, the entries whose columns lie in chunk c (rows in order, the entries
// of a row in column order; row starts rs_c).  A chunk's slice of x_r fits
// in L2, so the gathers of one launch hit L2 instead of going to memory.
// Chunks 0..C-2 leave their rows' sums so far in `part` (chunkDot); the last
// finishes the row.
template <int G, typename T, bool STORE_HAT>
__global__ __launch_bounds__(kBlock) void kDualChunk(
    int m, const int* __restrict__ rs_c, const int* __restrict__ ci_c, const T* __restrict__ val_c,
    const T* __restrict__ xr, T* __restrict__ part, int c, int C, T* __restrict__ y,
    const T* __restrict__ y0, T* __restrict__ yh, const T* __restrict__ sigma_base,
    const T* __restrict__ rl, const T* __restrict__ ru, const T* __restrict__ q,
    const IterScalars* __restrict__ sc, int k) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    if (i >= m) return;
    const T acc = chunkDot<G>(rs_c[i], rs_c[i + 1], ci_c, val_c, xr, lane, c, part, i);
    if (lane != 0) return;
    if (c < C - 1) part[i] = acc;
    else dualEpilogue<T, STORE_HAT>(i, acc, y, y0, yh, sigma_base, rl, ru, q, sc, k);
}

// ---- evaluation (src/lp/pdlp.cpp:1076-1202) --------------------------------
// Column pass: A'y^ and A'(y^ - y0) per column, then the per-column terms of
// |dx|^2, <A'dy, dx>, the dual residual (sum of squares and max, original
// units), c'x^ and the column part of the dual objective.  One partial per
// block per field, reduced later in a fixed order (no atomics).
enum ColField { kDx2 = 0, kInter, kDsq, kDmax, kPobj, kDcol, kNumColFields };
enum RowField { kPsq = 0, kPmax, kDy2, kDrow, kNumRowFields };

template <int G, typename TI>
__global__ __launch_bounds__(kBlock) void kEvalCol(
    int n, const int* __restrict__ cstarts, const int* __restrict__ ridx,
    const double* __restrict__ valt, const TI* __restrict__ yh, const TI* __restrict__ y0,
    const TI* __restrict__ xh, const TI* __restrict__ x0, const double* __restrict__ c,
    const double* __restrict__ lo, const double* __restrict__ hi,
    const double* __restrict__ col_scale, double os, double* __restrict__ partials,
    int nblocks) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int j = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    double v[kNumColFields] = {0, 0, 0, 0, 0, 0};
    if (j < n) {
        const int b = cstarts[j], e = cstarts[j + 1];
        double aty = 0.0, atdy = 0.0;
        for (int q = b + lane; q < e; q += G) {
            const int i = ridx[q];
            const double a = valt[q];
            const double yi = static_cast<double>(yh[i]);
            aty += a * yi;
            atdy += a * (yi - static_cast<double>(y0[i]));
        }
        aty = groupSum<G>(aty);
        atdy = groupSum<G>(atdy);
        if (lane == 0) {
            const double xj = static_cast<double>(xh[j]);
            const double dx = xj - static_cast<double>(x0[j]);
            v[kDx2] = dx * dx;
            v[kInter] = atdy * dx;
            const double g = c[j] + aty;
            const double lb = lo[j], ub = hi[j];
            // Activity is tested against the bounds the iteration projected
            // onto (the iterate type's rounding of the FP64 bounds).
            const double lbi = static_cast<double>(static_cast<TI>(lb));
            const double ubi = static_cast<double>(static_cast<TI>(ub));
            double dr;
            if (xj <= lbi + 1e-12 && isfinite(lb)) dr = minRef(g, 0.0);
            else if (xj >= ubi - 1e-12 && isfinite(ub)) dr = maxRef(g, 0.0);
            else dr = g;
            const double d = dr / (os * col_scale[j]);
            v[kDsq] = d * d;
            v[kDmax] = fabs(d);
            v[kPobj] = c[j] * xj;
            double gp = g;
            if (!isfinite(lb)) gp = minRef(gp, 0.0);
            if (!isfinite(ub)) gp = maxRef(gp, 0.0);
            if (gp > 0.0 && isfinite(lb)) v[kDcol] = gp * lb;
            else if (gp < 0.0 && isfinite(ub)) v[kDcol] = gp * ub;
        }
    }
    const bool isMax[kNumColFields] = {false, false, false, true, false, false};
    double out[kNumColFields];
    blockReduce<kNumColFields>(v, isMax, out);
    if (threadIdx.x == 0)
        for (int f = 0; f < kNumColFields; ++f) partials[f * nblocks + blockIdx.x] = out[f];
}

// Row pass: A x^ per row, the scaled primal violation (original units), the
// dual movement |dy|^2 and the row part of the dual objective.
template <int G, typename TI>
__global__ __launch_bounds__(kBlock) void kEvalRow(
    int m, const int* __restrict__ rstarts, const int* __restrict__ cidx,
    const double* __restrict__ val, const TI* __restrict__ xh, const TI* __restrict__ yh,
    const TI* __restrict__ y0, const double* __restrict__ rl, const double* __restrict__ ru,
    const double* __restrict__ row_scale, double cscale, double* __restrict__ partials,
    int nblocks) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    double v[kNumRowFields] = {0, 0, 0, 0};
    if (i < m) {
        const double ax = groupDot<G>(rstarts, cidx, val, xh, i, lane);
        if (lane == 0) {
            const double lo = rl[i], hi = ru[i];
            const double vs = ax - clampRef(ax, lo, hi);
            const double p = vs / (cscale * row_scale[i]);
            v[kPsq] = p * p;
            v[kPmax] = fabs(p);
            const double yi = static_cast<double>(yh[i]);
            const double dy = yi - static_cast<double>(y0[i]);
            v[kDy2] = dy * dy;
            double yp = yi;
            if (!isfinite(hi)) yp = minRef(yp, 0.0);
            if (!isfinite(lo)) yp = maxRef(yp, 0.0);
            if (yp > 0.0 && isfinite(hi)) v[kDrow] = -(yp * hi);
            else if (yp < 0.0 && isfinite(lo)) v[kDrow] = -(yp * lo);
        }
    }
    const bool isMax[kNumRowFields] = {false, true, false, false};
    double out[kNumRowFields];
    blockReduce<kNumRowFields>(v, isMax, out);
    if (threadIdx.x == 0)
        for (int f = 0; f < kNumRowFields; ++f) partials[f * nblocks + blockIdx.x] = out[f];
}

// Final reduction of `nf` fields of `nblocks` partials each, in a fixed order:
// thread t sums partials t, t+256, ... sequentially, then a block reduction.
// maxMask bit f marks a max field.  One block.
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void kFinalize(const double* __restrict__ partials,
                                                    int nblocks, int nf, unsigned maxMask,
                                                    double* __restrict__ out) {
    __shared__ double red[kBlock];
    for (int f = 0; f < nf; ++f) {
        const bool mx = f < 32 && ((maxMask >> f) & 1u);
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) {
            const double p = partials[f * nblocks + b];
            a = mx ? maxRef(a, p) : a + p;
        }
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                const double o = red[threadIdx.x + s];
                red[threadIdx.x] = mx ? maxRef(red[threadIdx.x], o) : red[threadIdx.x] + o;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[f] = red[0];
        __syncthreads();
    }
}

// ---- elementwise helpers ---------------------------------------------------
template <typename TD, typename TS>
__global__ void kConvert(long long count, const TS* __restrict__ src, TD* __restrict__ dst) {
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < count) dst[t] = static_cast<TD>(src[t]);
}

// x0 = x = x^ (restart, reference restartCopyKernelImpl pdlp_kernels.cu:377-405).
template <typename T>
__global__ void kRestartCopy(int count, const T* __restrict__ hat, T* __restrict__ cur,
                             T* __restrict__ anchor) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) {
        const T v = hat[t];
        cur[t] = v;
        anchor[t] = v;
    }
}

}  // namespace mpx
