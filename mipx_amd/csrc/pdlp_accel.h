// pdlp_accel.h -- device kernels of the two optional accelerations of the
// single-LP solver: the adaptive step of the GPU host loop (reference
// src/lp/pdlp.cpp:2271-2307) and type-I Anderson acceleration at restarts
// (pdlp.cpp:293-437, applied at 2418-2466).  Both are off by default.
//
// Every reduction is a fixed-order two-level sum (per-block blockReduce, then
// kFinalize over the partials), so a run is reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "pdlp_kernels.h"

namespace mpx {

// ---- adaptive step ----------------------------------------------------------
// Row pass: a = A x^ per row (FP64 matrix), partial sums of
// sigma_i (a_i - prev_i)^2, then prev := a.
template <int G, typename TI>
__global__ __launch_bounds__(kBlock) void kAdaptRows(int m, const int* __restrict__ rstarts,
                                                     const int* __restrict__ cidx,
                                                     const double* __restrict__ val,
                                                     const TI* __restrict__ xh,
                                                     const double* __restrict__ sig,
                                                     double* __restrict__ prev_ax,
                                                     double* __restrict__ partials) {
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    double v[1] = {0.0};
    if (i < m) {
        const double ax = groupDot<G>(rstarts, cidx, val, xh, i, lane);
        if (lane == 0) {
            const double d = ax - prev_ax[i];
            v[0] = sig[i] * d * d;
            prev_ax[i] = ax;
        }
    }
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) partials[blockIdx.x] = out[0];
}

// Column pass: partial sums of (x^_j - prev_j)^2 / tau_j, then prev := x^.
template <typename TI>
__global__ __launch_bounds__(kBlock) void kAdaptCols(int n, const TI* __restrict__ xh,
                                                     const double* __restrict__ tau,
                                                     double* __restrict__ prev_x,
                                                     double* __restrict__ partials) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0.0};
    if (j < n) {
        const double xj = static_cast<double>(xh[j]);
        const double d = xj - prev_x[j];
        v[0] = d * d / tau[j];
        prev_x[j] = xj;
    }
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) partials[blockIdx.x] = out[0];
}

// ---- Anderson acceleration --------------------------------------------------
// History: `w` slots of z = [x^; y^] and r = z - [x0; y0], each of dim = n + m
// doubles, slot s at z + s * dim.  The d newest slots in age order are
// slot(i) = (first + i) % w, i = 0..d-1.

// Stores the restart candidate and its residual into slot `s`.
template <typename T>
__global__ void kAaPush(int n, int m, const T* __restrict__ xh, const T* __restrict__ yh,
                        const T* __restrict__ x0, const T* __restrict__ y0, double* __restrict__ z,
                        double* __restrict__ r) {
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < n) {
        const double h = static_cast<double>(xh[t]);
        z[t] = h;
        r[t] = h - static_cast<double>(x0[t]);
    } else if (t < static_cast<long long>(n) + m) {
        const long long i = t - n;
        const double h = static_cast<double>(yh[i]);
        z[t] = h;
        r[t] = h - static_cast<double>(y0[i]);
    }
}

__device__ inline void aaPair(int p, int d, int& i, int& j) {
    i = 0;
    while (p >= d - i) {
        p -= d - i;
        ++i;
    }
    j = i + p;
}

// Gram entries r_i . r_j for the d(d+1)/2 pairs i <= j (row-major pair order,
// blockIdx.y); each block sums a fixed set of elements (grid stride over
// gridDim.x blocks) into partials[pair * gridDim.x + block].
__global__ __launch_bounds__(kBlock) void kAaGram(long long dim, int d, int first, int w,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ partials) {
    int i, j;
    aaPair(static_cast<int>(blockIdx.y), d, i, j);
    const double* __restrict__ a = r + static_cast<long long>((first + i) % w) * dim;
    const double* __restrict__ b = r + static_cast<long long>((first + j) % w) * dim;
    double v[1] = {0.0};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; k < dim; k += stride)
        v[0] += a[k] * b[k];
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) partials[static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x] = out[0];
}

// |sum_i alpha_i r_i|^2: per element the weighted sum in age order, then a
// fixed-order reduction.
__global__ __launch_bounds__(kBlock) void kAaResid(long long dim, int d, int first, int w,
                                                   const double* __restrict__ r,
                                                   const double* __restrict__ alpha,
                                                   double* __restrict__ partials) {
    double v[1] = {0.0};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; k < dim; k += stride) {
        double s = 0.0;
        for (int i = 0; i < d; ++i) s += alpha[i] * r[static_cast<long long>((first + i) % w) * dim + k];
        v[0] += s * s;
    }
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) partials[blockIdx.x] = out[0];
}

// Accepted extrapolation: z = sum_i alpha_i z_i; x0 = x = clamp(z_x, l, u)
// (scaled column bounds), y0 = y = z_y.
template <typename T>
__global__ void kAaApply(int n, int m, int d, int first, int w, const double* __restrict__ z,
                         const double* __restrict__ alpha, const double* __restrict__ l,
                         const double* __restrict__ u, T* __restrict__ x, T* __restrict__ x0,
                         T* __restrict__ y, T* __restrict__ y0) {
    const long long dim = static_cast<long long>(n) + m;
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= dim) return;
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += alpha[i] * z[static_cast<long long>((first + i) % w) * dim + t];
    if (t < n) {
        const T v = static_cast<T>(clampRef(s, l[t], u[t]));
        x[t] = v;
        x0[t] = v;
    } else {
        const T v = static_cast<T>(s);
        y[t - n] = v;
        y0[t - n] = v;
    }
}

}  // namespace mpx
