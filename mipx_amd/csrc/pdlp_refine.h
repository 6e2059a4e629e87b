// pdlp_refine.h -- FP64 iterative refinement of mixed-precision PDLP.
//
// After the plain FP32 phase, the iterate is split into an FP64 base point
// (x_b, y_b) and an FP32 correction (dx, dy): x = x_b + dx, y = y_b + dy.
// The FP32 kernels then run the same Halpern PDHG on the correction, whose
// data are formed in FP64 and rounded once:
//   c'  = c + A' y_b            (the base point's reduced costs)
//   l'  = l - x_b,  u' = u - x_b
//   rl' = rl - A x_b, ru' = ru - A x_b,   q = y_b
// so that every FP32 quantity is the size of the remaining correction, not
// of the solution: near the optimum the FP32 rounding error shrinks with it.
// The primal step keeps its form, dx^ = clamp(dx - tau (c' + A'dy), l', u').
// The dual prox of y = y_b + dy (row type from the bounds, as in
// dualProx) becomes, with w = dy + sigma A dx_r,
//   equality:    dy^ = w - sigma rl'
//   upper-only:  dy^ = max(-q, w - sigma ru')
//   lower-only:  dy^ = min(-q, w - sigma rl')
//   ranged:      w - sigma ru' if that is >= -q, else w - sigma rl' if that is
//                <= -q, else -q
// which is exactly dualProx(y_b + w + sigma A x_b) - y_b; q only enters the
// comparisons and the "row inactive" value, never a sum with dy.
// At each restart the base moves to the evaluated point (re-centring, one
// FP64 pass with both SpMVs) and the correction restarts from zero.
#pragma once

#include "device_common.h"

namespace mpx {

template <typename T>
__device__ inline T dualProxShifted(T w, T sig, T rl, T ru, T q) {
    if (rl == ru) return w - sig * rl;
    const T th = w - sig * ru, tl = w - sig * rl;
    if (!isfinite(rl)) return maxRef(-q, th);
    if (!isfinite(ru)) return minRef(-q, tl);
    if (th >= -q) return th;
    if (tl <= -q) return tl;
    return -q;
}

// Column pass of a re-centring: c' = c + A'y_b (sequential per column, CSC),
// l' = l - x_b, u' = u - x_b, rounded to T; the correction iterates zero.
template <typename T>
__global__ __launch_bounds__(kBlock) void kRefineCols(
    int n, const int* __restrict__ cs, const int* __restrict__ ri, const double* __restrict__ valt,
    const double* __restrict__ c, const double* __restrict__ lo, const double* __restrict__ hi,
    const double* __restrict__ xb, const double* __restrict__ yb, T* __restrict__ cr,
    T* __restrict__ lr, T* __restrict__ ur, T* __restrict__ dx, T* __restrict__ dx0,
    T* __restrict__ dxh) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    double g = 0.0;
    for (int k = cs[j]; k < cs[j + 1]; ++k) g += valt[k] * yb[ri[k]];
    cr[j] = static_cast<T>(c[j] + g);
    lr[j] = static_cast<T>(lo[j] - xb[j]);
    ur[j] = static_cast<T>(hi[j] - xb[j]);
    dx[j] = T(0);
    dx0[j] = T(0);
    dxh[j] = T(0);
}

// Row pass: rl' = rl - A x_b, ru' = ru - A x_b, q = y_b; correction zero.
template <typename T>
__global__ __launch_bounds__(kBlock) void kRefineRows(
    int m, const int* __restrict__ rs, const int* __restrict__ ci, const double* __restrict__ val,
    const double* __restrict__ rl, const double* __restrict__ ru, const double* __restrict__ xb,
    const double* __restrict__ yb, T* __restrict__ rlr, T* __restrict__ rur, T* __restrict__ q,
    T* __restrict__ dy, T* __restrict__ dy0, T* __restrict__ dyh) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    double r = 0.0;
    for (int k = rs[i]; k < rs[i + 1]; ++k) r += val[k] * xb[ci[k]];
    rlr[i] = static_cast<T>(rl[i] - r);
    rur[i] = static_cast<T>(ru[i] - r);
    q[i] = static_cast<T>(yb[i]);
    dy[i] = T(0);
    dy0[i] = T(0);
    dyh[i] = T(0);
}

// The evaluated point in FP64: x^ = x_b + dx^, snapped to the FP64 bound
// when dx^ sits exactly on the rounded one (the projection put it there);
// y^ = y_b + dy^, and exactly 0 when the prox found the row inactive
// (dy^ == -q) on a row that is not an equality; the anchors likewise.
template <typename T>
__global__ __launch_bounds__(kBlock) void kRefineMaterialize(
    int n, int m, const double* __restrict__ xb, const double* __restrict__ yb,
    const T* __restrict__ dxh, const T* __restrict__ dx0, const T* __restrict__ dyh,
    const T* __restrict__ dy0, const T* __restrict__ lr, const T* __restrict__ ur,
    const double* __restrict__ lo, const double* __restrict__ hi, const T* __restrict__ q,
    const T* __restrict__ rlr, const T* __restrict__ rur, double* __restrict__ xh,
    double* __restrict__ x0, double* __restrict__ yh, double* __restrict__ y0) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) {
        const T d = dxh[t];
        xh[t] = d == lr[t] ? lo[t] : (d == ur[t] ? hi[t] : xb[t] + static_cast<double>(d));
        x0[t] = xb[t] + static_cast<double>(dx0[t]);
    }
    if (t < m) {
        const T d = dyh[t];
        const bool inactive = rlr[t] != rur[t] && d == -q[t];
        yh[t] = inactive ? 0.0 : yb[t] + static_cast<double>(d);
        y0[t] = yb[t] + static_cast<double>(dy0[t]);
    }
}

}  // namespace mpx
