// pdlp_setup.h -- on-device problem scaling and power iteration.
//
// Device restatement of the reference host code buildScaledProblem
// (src/lp/pdlp.cpp:469-669) and estimateSpectralNorm (pdlp.cpp:748-795).
// Every per-row / per-column quantity is computed by one thread in the
// reference's own order (row maxima, squared norms summed in storage order),
// so the scaled matrix, row/column scales and Pock-Chambolle bases are
// bit-identical to the reference's.  Only the two global sums of the
// bound/objective rescaling and the power-iteration dot products are
// reduced in a different (fixed, deterministic) order.
#pragma once

#include "device_common.h"

namespace mpx {

// Ruiz row pass: row max |a|, skip if <= 1e-12, scale by 1/sqrt(max).
__global__ void kRuizRows(int m, const int* __restrict__ rs, double* __restrict__ val,
                          double* __restrict__ row_scale) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const int b = rs[i], e = rs[i + 1];
    double rmax = 0.0;
    for (int k = b; k < e; ++k) rmax = maxRef(rmax, fabs(val[k]));
    if (rmax <= 1e-12) return;
    const double f = 1.0 / sqrt(rmax);
    row_scale[i] *= f;
    for (int k = b; k < e; ++k) val[k] *= f;
}

// Ruiz column pass, part 1: column max |a| (through the CSC permutation) and
// the column scale update.
__global__ void kRuizColMax(int n, const int* __restrict__ cs, const int* __restrict__ perm,
                            const double* __restrict__ val, double* __restrict__ cmax,
                            double* __restrict__ col_scale) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    double mx = 0.0;
    for (int q = cs[j]; q < cs[j + 1]; ++q) mx = maxRef(mx, fabs(val[perm[q]]));
    cmax[j] = mx;
    if (mx > 1e-12) col_scale[j] *= 1.0 / sqrt(mx);
}

// Ruiz column pass, part 2: a_k *= 1/sqrt(cmax_j) where cmax_j > 1e-12.
__global__ void kRuizColApply(long long nnz, const int* __restrict__ ci,
                              const double* __restrict__ cmax, double* __restrict__ val) {
    const long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= nnz) return;
    const double mx = cmax[ci[k]];
    if (mx > 1e-12) val[k] *= 1.0 / sqrt(mx);
}

// Objective and bounds into the Ruiz-scaled space (pdlp.cpp:476-479,560-573).
__global__ void kScaleColumns(int n, double sgn, const double* __restrict__ obj,
                              const double* __restrict__ col_scale, const double* __restrict__ l,
                              const double* __restrict__ u, double* __restrict__ c,
                              double* __restrict__ ls, double* __restrict__ us) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double cj = sgn * obj[j];
    c[j] = cj * col_scale[j];
    const double lj = l[j], uj = u[j];
    ls[j] = isfinite(lj) ? lj / col_scale[j] : lj;
    us[j] = isfinite(uj) ? uj / col_scale[j] : uj;
}

__global__ void kScaleRows(int m, const double* __restrict__ row_scale,
                           const double* __restrict__ rl, const double* __restrict__ ru,
                           double* __restrict__ rls, double* __restrict__ rus) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const double a = rl[i], b = ru[i];
    rls[i] = isfinite(a) ? a * row_scale[i] : a;
    rus[i] = isfinite(b) ? b * row_scale[i] : b;
}

// CSC values from CSR values through the transpose permutation.
__global__ void kGather(long long nnz, const int* __restrict__ perm, const double* __restrict__ src,
                        double* __restrict__ dst) {
    const long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (k < nnz) dst[k] = src[perm[k]];
}

// Pock-Chambolle bases: 1 / max(sqrt(sum a^2), 1e-8), summed in storage order
// (rows: CSR order; columns: increasing row order = CSC order).
__global__ void kPockChambolle(int count, const int* __restrict__ starts,
                               const double* __restrict__ val, double* __restrict__ base) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    double s = 0.0;
    for (int k = starts[i]; k < starts[i + 1]; ++k) s += val[k] * val[k];
    base[i] = 1.0 / maxRef(sqrt(s), 1e-8);
}

// Partial sums for the bound/objective rescaling norms: sum c^2 over columns
// and sum of finite rl^2 + ru^2 over rows (per-row order rl then ru).
__global__ __launch_bounds__(kBlock) void kRescaleNorms(int n, int m, const double* __restrict__ c,
                                                        const double* __restrict__ rl,
                                                        const double* __restrict__ ru,
                                                        double* __restrict__ partials,
                                                        int nblocks) {
    double v[2] = {0.0, 0.0};
    for (long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; t < n;
         t += static_cast<long long>(nblocks) * kBlock)
        v[0] += c[t] * c[t];
    for (long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; t < m;
         t += static_cast<long long>(nblocks) * kBlock) {
        double s = 0.0;
        if (isfinite(rl[t])) s += rl[t] * rl[t];
        if (isfinite(ru[t])) s += ru[t] * ru[t];
        v[1] += s;
    }
    const bool isMax[2] = {false, false};
    double out[2];
    blockReduce<2>(v, isMax, out);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = out[0];
        partials[nblocks + blockIdx.x] = out[1];
    }
}

// Applies os to c and cs to the row and column bounds (pdlp.cpp:619-640).
__global__ void kApplyRescale(int n, int m, const double* __restrict__ scales,
                              double* __restrict__ c, double* __restrict__ l,
                              double* __restrict__ u, double* __restrict__ rl,
                              double* __restrict__ ru) {
    const double os = scales[0], cs = scales[1];
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < n) {
        c[t] *= os;
        if (isfinite(l[t])) l[t] *= cs;
        if (isfinite(u[t])) u[t] *= cs;
    }
    if (t < m) {
        if (isfinite(rl[t])) rl[t] *= cs;
        if (isfinite(ru[t])) ru[t] *= cs;
    }
}

// os = 1/(sqrt(sum c^2) + 1), cs = 1/(sqrt(sum b^2) + 1) from the finalized sums.
__global__ void kRescaleFactors(const double* __restrict__ sums, double* __restrict__ scales) {
    scales[0] = 1.0 / (sqrt(sums[0]) + 1.0);
    scales[1] = 1.0 / (sqrt(sums[1]) + 1.0);
}

// ---- power iteration on A'A ------------------------------------------------
// State (device): x (normalized), xn = A'A x, sigma^2, residual, done flag.
struct PowerState {
    double sigma_sq;
    double resid;
    double xnorm;
    int32_t done;
    int32_t iters;
};

// x <- x / ||x|| for the random start (norm from a finalized sum).
__global__ void kScaleBy(int n, const double* __restrict__ sumsq, double* __restrict__ x) {
    double nrm = sqrt(sumsq[0]);
    if (nrm < 1e-15) nrm = 1.0;
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < n) x[j] /= nrm;
}

template <int G>
__global__ __launch_bounds__(kBlock) void kPowerSpmvA(int m, const int* __restrict__ rs,
                                                      const int* __restrict__ ci,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y,
                                                      const PowerState* __restrict__ ps) {
    if (ps->done) return;
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    if (i >= m) return;
    double acc = 0.0;
    for (int k = rs[i] + lane; k < rs[i + 1]; k += G) acc += val[k] * x[ci[k]];
    acc = groupSum<G>(acc);
    if (lane == 0) y[i] = acc;
}

// xn = A' y and per-block partials of <x, xn>.
template <int G>
__global__ __launch_bounds__(kBlock) void kPowerSpmvAt(
    int n, const int* __restrict__ cs, const int* __restrict__ ri, const double* __restrict__ valt,
    const double* __restrict__ y, const double* __restrict__ x, double* __restrict__ xn,
    double* __restrict__ partials, int nblocks, const PowerState* __restrict__ ps) {
    if (ps->done) return;
    const long long gid = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int j = static_cast<int>(gid / G), lane = static_cast<int>(gid % G);
    double v[1] = {0.0};
    if (j < n) {
        double acc = 0.0;
        for (int k = cs[j] + lane; k < cs[j + 1]; k += G) acc += valt[k] * y[ri[k]];
        acc = groupSum<G>(acc);
        if (lane == 0) {
            xn[j] = acc;
            v[0] = x[j] * acc;
        }
    }
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) partials[blockIdx.x] = out[0];
}

// Residual and norm partials: sum (xn - s2 x)^2 and sum xn^2.
__global__ __launch_bounds__(kBlock) void kPowerResid(int n, const double* __restrict__ x,
                                                      const double* __restrict__ xn,
                                                      const double* __restrict__ s2p,
                                                      double* __restrict__ partials, int nblocks,
                                                      const PowerState* __restrict__ ps) {
    if (ps->done) return;
    const double s2 = s2p[0];
    double v[2] = {0.0, 0.0};
    for (long long j = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; j < n;
         j += static_cast<long long>(nblocks) * kBlock) {
        const double d = xn[j] - s2 * x[j];
        v[0] += d * d;
        v[1] += xn[j] * xn[j];
    }
    const bool isMax[2] = {false, false};
    double out[2];
    blockReduce<2>(v, isMax, out);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = out[0];
        partials[nblocks + blockIdx.x] = out[1];
    }
}

// Step bookkeeping (one thread): sigma^2, residual, xnorm, stop test.  The
// reference breaks before normalizing when ||xn|| < 1e-15 and otherwise
// normalizes, then breaks when resid < tol |sigma^2|.
__global__ void kPowerUpdate(const double* __restrict__ s2p, const double* __restrict__ rn,
                             double tol, int max_iter, PowerState* __restrict__ ps) {
    if (ps->done) return;
    ps->sigma_sq = s2p[0];
    ps->resid = sqrt(rn[0]);
    ps->xnorm = sqrt(rn[1]);
    ps->iters += 1;
}

__global__ void kPowerNormalize(int n, const double* __restrict__ xn, double* __restrict__ x,
                                const PowerState* __restrict__ ps) {
    if (ps->done || ps->xnorm < 1e-15) return;
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < n) x[j] = xn[j] / ps->xnorm;
}

__global__ void kPowerCheck(double tol, int max_iter, PowerState* __restrict__ ps) {
    if (ps->done) return;
    if (ps->xnorm < 1e-15 || ps->resid < tol * fabs(ps->sigma_sq) || ps->iters >= max_iter)
        ps->done = 1;
}

// Warm start into the scaled space (pdlp.cpp:1812-1827):
// x = x_orig / col_scale * cs, y = y_orig * os / (sgn * row_scale).
__global__ void kWarmStart(int n, int m, double sgn, const double* __restrict__ wx,
                           const double* __restrict__ wy, const double* __restrict__ col_scale,
                           const double* __restrict__ row_scale, const double* __restrict__ scales,
                           double* __restrict__ x, double* __restrict__ x0, double* __restrict__ y,
                           double* __restrict__ y0) {
    const double os = scales[0], cs = scales[1];
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < n) {
        const double v = wx[t] / col_scale[t] * cs;
        x[t] = v;
        x0[t] = v;
    }
    if (t < m) {
        const double v = wy[t] * os / (sgn * row_scale[t]);
        y[t] = v;
        y0[t] = v;
    }
}

// Recovery into original space (pdlp.cpp:1677-1697): x clamped to the
// original bounds, y = sgn y^ row_scale / os.
template <typename TI>
__global__ void kRecover(int n, int m, double sgn, const TI* __restrict__ xh,
                         const TI* __restrict__ yh, const double* __restrict__ col_scale,
                         const double* __restrict__ row_scale, const double* __restrict__ scales,
                         const double* __restrict__ l, const double* __restrict__ u, int clamp,
                         double* __restrict__ xo, double* __restrict__ yo) {
    const double os = scales[0], cs = scales[1];
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < n) {
        const double inv_cs = 1.0 / cs;
        double xj = static_cast<double>(xh[t]) * col_scale[t] * inv_cs;
        if (clamp) {
            if (isfinite(l[t])) xj = maxRef(xj, l[t]);
            if (isfinite(u[t])) xj = minRef(xj, u[t]);
        }
        xo[t] = xj;
    }
    if (t < m) {
        const double inv_os = 1.0 / os;
        yo[t] = sgn * static_cast<double>(yh[t]) * row_scale[t] * inv_os;
    }
}

}  // namespace mpx
