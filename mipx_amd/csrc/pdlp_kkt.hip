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
// Two entry points: mpx_kkt_check_device takes its scratch from a
// stream-ordered allocation, mpx_kkt_check_device_noalloc from a fixed
// device array (for callers that cannot allocate).  Every partial is
// written before it is read in the same call, so the fixed array needs no
// clearing; two calls that overlap on one device race on it.
#include <mutex>
#include <string>

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

// ---- allocation-free variant ---------------------------------------------
// Fixed scratch of the allocation-free entry point: 1 MiB of partial sums per
// device, held until the process ends; enough for LPs whose row and column
// blocks of 256 number at most kKktNoAllocDoubles / 5 = 26,214 together
// (about 6.7 million rows plus columns).  Calls on one device share it.
constexpr size_t kKktNoAllocDoubles = size_t(1) << 17;
__device__ __align__(256) double scratch[kKktNoAllocDoubles];

// The device address of `scratch` on each device, looked up once per device
// with that device current (the name itself is only a host-side stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* noAllocScratch() {
    int dev = 0;
    MPX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        setError("mpx_kkt_check_device_noalloc: device ordinal out of range");
        throw DeviceError();
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        MPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
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

extern "C" int mpx_kkt_check_device_noalloc(const mpx_lp* lp, const int32_t* col_starts,
                                            const int32_t* row_indices, const double* col_values,
                                            const double* x, const double* y, double* out,
                                            void* stream) {
    using namespace mpx;
    KktArgs a;
    if (const int rc = kktArgs(lp, col_starts, row_indices, col_values, x, y, out, a)) return rc;
    if (kktScratchDoubles(a.m, a.n) > kKktNoAllocDoubles) {
        setError("mpx_kkt_check_device_noalloc: the LP needs " +
                 std::to_string(kktScratchDoubles(a.m, a.n)) +
                 " doubles of partial sums, more than the fixed scratch of " +
                 std::to_string(kKktNoAllocDoubles) +
                 " (at most 26,214 row plus column blocks of 256); use mpx_kkt_check_device");
        return MPX_ERR_INVALID;
    }
    try {
        launchKkt(a, noAllocScratch(), out, static_cast<hipStream_t>(stream));
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}
