// pdlp_kkt.hip -- relative KKT residuals of an original-space (x, y) on the
// device: the termination quantities of the reference's evaluation block in
// original units (src/lp/pdlp.cpp:1105-1202 for the CPU loop, 2196-2269 for
// the GPU loop), computed from the LP as given rather than from a solver's
// scaled state, so a caller can certify any returned solution where it lies.
//
// Same formulas as the long-double certificate of the test oracle, in FP64:
//   rows:    v_i = (A x)_i - clamp((A x)_i, rl_i, ru_i); ybar_i = sgn y_i with
//            the sign rules of rows without an upper / lower bound (a wrong
//            sign counts as dual-sign violation and is dropped)
//   columns: g_j = sgn (c_j + (A'y)_j); the part of g_j that no finite bound
//            can absorb is the dual residual; bounds contribute to the dual
//            objective with the sign of g_j
//   rel_primal = ||v||_2 / (1 + ||b||_2), rel_dual = ||dres||_2 / (1 + ||c||_2),
//   rel_gap = |pobj - dobj| / (1 + |pobj| + |dobj|)   (original-problem norms,
//   pdlp.cpp:53-81)
// One thread per row (column) sums its entries sequentially in storage
// order; each block leaves five partial fields in a scratch buffer, and one
// block combines the partials in block order: the result is deterministic.
//
// The partial sums live in a stream-ordered allocation of the call.

#include "../../include/mipx_pdlp.h"
#include "device_common.h"

namespace mpx {
namespace {

constexpr int kKktFields = 5;
// Row fields: sum of squared violations, max violation, ||b||^2 terms,
// dual-objective row part, max dual-sign violation.
enum { kRPsq = 0, kRPmax = 1, kRBsq = 2, kRDrow = 3, kRSign = 4 };
// Column fields: ||c||^2, squared dual residual, dual-objective column part,
// max bound violation, primal objective.
enum { kCCsq = 0, kCDsq = 1, kCDcol = 2, kCBviol = 3, kCPobj = 4 };

struct KktArgs {
    int m, n, sense_max;
    double offset;
    const int *rs, *ci, *cs, *ri;
    const double *val, *valt, *obj, *lo, *hi, *rl, *ru, *x, *y;
};

__global__ __launch_bounds__(kBlock) void kKktRows(KktArgs a, double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double f[kKktFields] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < a.m) {
        double ax = 0.0;
        for (int k = a.rs[i]; k < a.rs[i + 1]; ++k) ax += a.val[k] * a.x[a.ci[k]];
        const double lo = a.rl[i], hi = a.ru[i];
        double v = 0.0;
        if (ax < lo) v = ax - lo;
        else if (ax > hi) v = ax - hi;
        f[kRPsq] = v * v;
        f[kRPmax] = fabs(v);
        const bool hasLo = isfinite(lo), hasHi = isfinite(hi);
        if (hasHi) f[kRBsq] += hi * hi;
        if (hasLo && lo != hi) f[kRBsq] += lo * lo;
        double yb = (a.sense_max ? -1.0 : 1.0) * a.y[i];
        if (!hasHi && yb > 0.0) {
            f[kRSign] = yb;
            yb = 0.0;
        }
        if (!hasLo && yb < 0.0) {
            f[kRSign] = -yb;
            yb = 0.0;
        }
        if (yb > 0.0 && hasHi) f[kRDrow] = -yb * hi;
        else if (yb < 0.0 && hasLo) f[kRDrow] = -yb * lo;
    }
    const bool mx[kKktFields] = {false, true, false, false, true};
    double out[kKktFields];
    blockReduce<kKktFields>(f, mx, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < kKktFields; ++q) part[blockIdx.x * kKktFields + q] = out[q];
}

__global__ __launch_bounds__(kBlock) void kKktCols(KktArgs a, double* __restrict__ part) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    double f[kKktFields] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (j < a.n) {
        double aty = 0.0;
        for (int k = a.cs[j]; k < a.cs[j + 1]; ++k) aty += a.valt[k] * a.y[a.ri[k]];
        const double cj = a.obj[j], lb = a.lo[j], ub = a.hi[j], xj = a.x[j];
        const bool hasL = isfinite(lb), hasU = isfinite(ub);
        f[kCCsq] = cj * cj;
        if (hasL && xj < lb) f[kCBviol] = lb - xj;
        if (hasU && xj > ub) f[kCBviol] = maxRef(f[kCBviol], xj - ub);
        f[kCPobj] = cj * xj;
        const double g = (a.sense_max ? -1.0 : 1.0) * (cj + aty);
        double gf = g;
        if (!hasL) gf = minRef(gf, 0.0);
        if (!hasU) gf = maxRef(gf, 0.0);
        const double dres = g - gf;
        f[kCDsq] = dres * dres;
        if (gf > 0.0 && hasL) f[kCDcol] = gf * lb;
        else if (gf < 0.0 && hasU) f[kCDcol] = gf * ub;
    }
    const bool mx[kKktFields] = {false, false, false, true, false};
    double out[kKktFields];
    blockReduce<kKktFields>(f, mx, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < kKktFields; ++q) part[blockIdx.x * kKktFields + q] = out[q];
}

// One block: thread t combines the partials of blocks t, t + 256, ... in
// order, then the block reduction; thread 0 forms the eight outputs.
__global__ __launch_bounds__(kBlock) void kKktFinish(KktArgs a, const double* __restrict__ rpart,
                                                     int rblocks, const double* __restrict__ cpart,
                                                     int cblocks, double* __restrict__ out) {
    double f[2 * kKktFields];
    const bool mx[2 * kKktFields] = {false, true, false, false, true,
                                     false, false, false, true, false};
    for (int q = 0; q < 2 * kKktFields; ++q) f[q] = 0.0;
    for (int b = threadIdx.x; b < rblocks; b += kBlock)
        for (int q = 0; q < kKktFields; ++q) {
            const double v = rpart[b * kKktFields + q];
            f[q] = mx[q] ? maxRef(f[q], v) : f[q] + v;
        }
    for (int b = threadIdx.x; b < cblocks; b += kBlock)
        for (int q = 0; q < kKktFields; ++q) {
            const double v = cpart[b * kKktFields + q];
            f[kKktFields + q] = mx[kKktFields + q] ? maxRef(f[kKktFields + q], v) : f[kKktFields + q] + v;
        }
    double s[2 * kKktFields];
    blockReduce<2 * kKktFields>(f, mx, s);
    if (threadIdx.x == 0) {
        const double* r = s;
        const double* c = s + kKktFields;
        const double bnorm = sqrt(r[kRBsq]), cnorm = sqrt(c[kCCsq]);
        const double sgn = a.sense_max ? -1.0 : 1.0;
        const double pob = a.offset + c[kCPobj];
        const double dob = a.offset + sgn * (c[kCDcol] + r[kRDrow]);
        out[0] = sqrt(r[kRPsq]) / (1.0 + bnorm);
        out[1] = sqrt(c[kCDsq]) / (1.0 + cnorm);
        out[2] = fabs(pob - dob) / (1.0 + fabs(pob) + fabs(dob));
        out[3] = pob;
        out[4] = dob;
        out[5] = c[kCBviol];
        out[6] = r[kRPmax];
        out[7] = r[kRSign];
    }
}

// Scratch doubles one call needs: five partial fields per row block and per
// column block.
size_t kktScratchDoubles(int m, int n) {
    return static_cast<size_t>(ceilDiv(m, kBlock) + ceilDiv(n, kBlock)) * kKktFields;
}

int kktArgs(const mpx_lp* lp, const int32_t* cs, const int32_t* ri, const double* valt,
            const double* x, const double* y, double* out, KktArgs& a) {
    if (!lp || !cs || !x || !out || lp->num_rows < 0 || lp->num_cols <= 0 || lp->nnz < 0 ||
        !lp->row_starts || !lp->obj || !lp->col_lower || !lp->col_upper ||
        (lp->num_rows > 0 && (!lp->row_lower || !lp->row_upper || !y)) ||
        (lp->nnz > 0 && (!lp->col_indices || !lp->values || !ri || !valt))) {
        setError("mpx_kkt_check_device: invalid arguments");
        return MPX_ERR_INVALID;
    }
    a = KktArgs{lp->num_rows, lp->num_cols, lp->sense ? 1 : 0, lp->obj_offset,
                lp->row_starts, lp->col_indices, cs, ri, lp->values, valt, lp->obj,
                lp->col_lower, lp->col_upper, lp->row_lower, lp->row_upper, x, y};
    return MPX_OK;
}

void launchKkt(const KktArgs& a, double* scratch, double* out, hipStream_t st) {
    const int rb = ceilDiv(a.m, kBlock), cb = ceilDiv(a.n, kBlock);
    double* rpart = scratch;
    double* cpart = scratch + static_cast<size_t>(rb) * kKktFields;
    if (rb > 0) kKktRows<<<rb, kBlock, 0, st>>>(a, rpart);
    kKktCols<<<cb, kBlock, 0, st>>>(a, cpart);
    kKktFinish<<<1, kBlock, 0, st>>>(a, rpart, rb, cpart, cb, out);
    MPX_HIP(hipGetLastError());
}

}  // namespace
}  // namespace mpx

extern "C" int mpx_kkt_check_device(const mpx_lp* lp, const int32_t* col_starts,
                                    const int32_t* row_indices, const double* col_values,
                                    const double* x, const double* y, double* out,
                                    void* stream) {
    using namespace mpx;
    KktArgs a;
    if (const int rc = kktArgs(lp, col_starts, row_indices, col_values, x, y, out, a)) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    try {
        double* scratch = nullptr;
        MPX_HIP(hipMallocAsync(reinterpret_cast<void**>(&scratch),
                               kktScratchDoubles(a.m, a.n) * sizeof(double), st));
        launchKkt(a, scratch, out, st);
        MPX_HIP(hipFreeAsync(scratch, st));
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

