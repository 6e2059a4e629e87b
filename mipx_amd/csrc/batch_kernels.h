// batch_kernels.h -- batched PDLP: per-instance state, evaluation sums and
// certificates.
//
// Many LPs (branch-and-bound node relaxations sharing one A, or independent
// LPs of uneven size) are iterated together: the iteration kernels
// (packed tiles in pdlp_batch.hip, lane-per-instance groups in
// batch_interleaved.h) run N iterations per round, then every instance is
// evaluated in FP64 in size-based parts and its restart / PI primal-weight /
// termination / certificate / polish decisions (the reference's host logic,
// src/lp/pdlp.cpp:2228-2469) run on the device in one workgroup per
// instance.  Instances never share mutable state, so an instance's result
// does not depend on what else is in the launch.
//
// Precision: iterations in T (float for MPX_PREC_MIXED, with float
// accumulation), evaluation always in FP64 against the FP64 scaled matrix.
// State stays resident in HBM between rounds (hipGraph-captured rounds).
#pragma once

#include "device_common.h"
#include "pdlp_kernels.h"

namespace mpx {

// (lanes per column, lanes per row) presets the autotuner chooses from.
constexpr int kNumPresets = 8;
__host__ __device__ constexpr int presetGC(int p) {
    return p == 0 ? 1 : p <= 2 ? 2 : p <= 4 ? 4 : p <= 6 ? 8 : 16;
}
__host__ __device__ constexpr int presetGR(int p) {
    return p == 0 ? 1 : p == 1 ? 2 : p <= 3 ? 4 : p <= 5 ? 8 : 16;
}

// Per-instance control state (FP64 scalars of the reference's host loop).
struct BCtrl {
    double step, pw;
    double fpe0, last_fpe;
    double pw_err_sum, pw_last_err, best_pw, best_score, precond_pw;
    double cnorm, bnorm;       // original-problem termination norms
    double os, cs, sgn, offset;
    double best_kkt;
    // last evaluation
    double rel_p, rel_d, rel_p_inf, rel_d_inf, gap, pobj, dobj;
    int32_t inner, total, pcount, phase, status, restarts, fp32_iters, stall;
    int32_t power_iters;
    int32_t round_iters, skip_eval;
    int32_t refined, refine_iters;  // 1 once the instance entered kPhaseRefine; its iterations
    int32_t pending;  // BPending: vector work of this round's decision, done in parts
    int32_t pad;
    unsigned long long t_finish;  // device wall clock when the instance finished
};

// Vector work that a control decision leaves to the parted kernels after it
// (kControlCopy, then kControlRecenter): restart copies, phase changes and
// the refinement's re-centring, each column / row slice by its own part.
enum BPending {
    kPendNone = 0,
    kPendPolish = 1,        // x0 = x = x^, y0 = y = y^ (FP64)
    kPendToFp64 = 2,        // FP64 x^ = x = x0 = FP32 x^ (and y)
    kPendToRefine = 3,      // FP64 base = FP32 x^, y^; then re-centre
    kPendRefineToFp64 = 4,  // x = x0 = x^, y = y0 = y^ (FP64)
    kPendRecenter = 5,      // FP64 base = x^, y^; then re-centre
    kPendRestart = 6        // x = x0 = x^, y = y0 = y^ in the phase's precision
};

// kPhaseRefine: FP32 iterations on the correction around an FP64 base point
// (iterative refinement, pdlp_refine.h), evaluated in FP64.
enum BPhase { kPhaseFp32 = 0, kPhaseFp64 = 1, kPhasePolish = 2, kPhaseDone = 3, kPhaseRefine = 4 };

// Phases whose iterations run in the FP32 kernels.
__host__ __device__ inline bool iteratesFp32(int phase) {
    return phase == kPhaseFp32 || phase == kPhaseRefine;
}

struct BOptions {
    int32_t max_iter, N, polish_max_iter, norm_inf, update_pw, enable_polish, precond_refresh,
        do_pc;
    double primal_tol, dual_tol, opt_tol, polish_p_tol, polish_d_tol;
    double refine_p_tol, refine_d_tol, refine_g_tol;  // FP64 phase of mixed precision
    int32_t mixed, refine;  // refine: the FP32 phase hands over to kPhaseRefine
    double handover;  // FP32 -> refinement (or FP64) at this KKT ratio
    double pid_kp, pid_ki, pid_kd, pid_i_smooth;
    double rs_suff, rs_nec, rs_art, precond_ratio;
};

template <typename T>
struct BVecs {
    T *x, *x0, *xr, *xh, *y, *y0, *yh;
    const T *val, *valt, *c, *tau, *l, *u, *sig, *rl, *ru;
    const T* q;  // refinement row shift (FP32 correction phase), else null
};

struct BInst {
    int32_t m, n, nnz, sense_max;
    int32_t preset, owner;
    int32_t il;  // iterated by the lane-per-instance kernels of its shared matrix
    const int *rs, *ci, *cs, *ri;
    // original bounds / objective (certificates)
    const double *lo_o, *hi_o, *rl_o, *ru_o, *obj_o;
    // scaled fp64 problem
    const double *val, *valt, *c, *l, *u, *rl, *ru, *row_scale, *col_scale, *sig, *tau;
    BVecs<float> f;
    BVecs<double> d;
    BCtrl* ctrl;
    // Correction data of the refinement phase (mixed precision): c', rl',
    // ru', q; l' and u' overwrite the instance's FP32 bounds.  At the first
    // re-centring f.c / f.rl / f.ru / f.q are pointed at them.
    float *rc, *rrl, *rru, *rq;
};

// A distinct constraint matrix of a batch (an "owner"): its original data,
// the scaled problem shared by its instances, FP32 copies and the
// power-iteration scratch.
struct BOwner {
    int32_t m, n, nnz, do_ruiz, ruiz_iters, do_pc, do_rescale, norm_inf, sense_max, pad;
    const int *rs, *ci, *cs, *ri, *perm;
    const double *val_o, *obj_o, *lo_o, *hi_o, *rl_o, *ru_o;
    double *val, *valt, *c, *l, *u, *rl, *ru, *row_scale, *col_scale, *cmax, *sig, *tau;
    float *valf, *valtf, *cf, *tauf, *sigf, *rlf, *ruf;
    double* scal;  // [os, cs, cnorm, bnorm, sigma_sq, power_iters]
    // power iteration: start vector (kept, so setup can be re-run) and
    // scratch (n, m, n)
    double *px0, *px, *py, *pxn;
    double offset;
    int32_t sv_max_iter;
    double sv_tol;
};

// ---- block helpers -----------------------------------------------------------

template <int G, typename T, typename TX>
__device__ inline T bDot(const int* __restrict__ starts, const int* __restrict__ idx,
                         const T* __restrict__ val, const TX* x, int row, int lane) {
    const int b = starts[row], e = starts[row + 1];
    T acc = T(0);
    for (int k = b + lane; k < e; k += G) acc += val[k] * static_cast<T>(x[idx[k]]);
    return groupSum<G>(acc);
}

// Certificate terms (src/lp/pdlp.cpp:119-284) gathered by the evaluation on
// the rounds whose decision may check the certificates, in original units
// before the rays are normalised: sums of squares for the norms, the sums of
// the objective terms, and the largest value of the quantities that the
// checks compare against tol times the norm (negated where the check is
// a lower one).  Column terms use A'y^, row terms A x^, both already formed
// by the evaluation.
enum CertColField {
    kCxsq = 0,  // |x_o|^2
    kCfBad,     // Farkas: max of -(A'y_o)_j (finite lower bound), (A'y_o)_j (finite upper)
    kCcol,      // column part of the Farkas sum
    kCrBad,     // primal ray: max of -x_o (finite lower bound), x_o (finite upper)
    kCobj,      // c_o'x_o
    kNumCertCol
};
enum CertRowField {
    kRysq = 0,  // |y_o|^2
    kRfBad,     // Farkas: max of -y_o (only an upper bound), y_o (only a lower bound)
    kRrow,      // row part of the Farkas sum
    kRrBad,     // primal ray: max of |A x_o| (both bounds), A x_o (upper only), -A x_o (lower only)
    kNumCertRow
};
constexpr int kNumEvalCore = static_cast<int>(kNumColFields) + static_cast<int>(kNumRowFields);
constexpr int kCertCol0 = kNumEvalCore, kCertRow0 = kNumEvalCore + kNumCertCol;
constexpr int kNumEval = kNumEvalCore + kNumCertCol + kNumCertRow;

// Whether field f of an evaluation part combines by maximum (else by sum).
__host__ __device__ constexpr bool evalFieldIsMax(int f) {
    return f == static_cast<int>(kDmax) || f == static_cast<int>(kNumColFields) + kPmax ||
           f == kCertCol0 + kCfBad || f == kCertCol0 + kCrBad || f == kCertRow0 + kRfBad ||
           f == kCertRow0 + kRrBad;
}

// The certificates are checked at the evaluation that ends a round with the
// iteration total a multiple of `every` (ten evaluation blocks), outside
// the polish phase.
__device__ inline bool certDue(const BCtrl& c, int every) {
    const int after = c.total + c.round_iters;
    return c.phase != kPhasePolish && every > 0 && after >= every && after % every == 0;
}
// Columns (rows) per lane group and pass in the batched evaluation.
constexpr int kEvalCW = 4;
// An instance's evaluation is split into parts of about kEvalChunk matrix
// entries plus vector elements (at most kEvalMaxParts), each a workgroup;
// the part count depends only on the instance's size, so its sums do not
// depend on the rest of the batch.
constexpr int kEvalMaxParts = 64;
constexpr long long kEvalChunk = 1 << 17;

__device__ inline int evalParts(const BInst& in) {
    const long long work = static_cast<long long>(in.nnz) + in.n + in.m;
    const long long p = (work + kEvalChunk - 1) / kEvalChunk;
    return static_cast<int>(p < 1 ? 1 : (p > kEvalMaxParts ? kEvalMaxParts : p));
}

struct BEval {
    double fpe, rel_p, rel_d, rel_p_inf, rel_d_inf, gap, pobj, dobj, dx2, dy2;
};

// Elements of the LDS stage of a staged evaluation (kStage): y^ and y0 for
// the column pass (2m), then x^ for the row pass (n).
constexpr int kEvalStageElems = 8192;

// The one LDS stage of a kernel's staged evaluations (a single array for all
// lane-group presets).
template <typename T>
__device__ inline T* evalStage() {
    __shared__ T stage[kEvalStageElems];
    return stage;
}

// FP64 evaluation sums of (xh, yh) against (x0, y0) over the columns
// [jb, je) and rows [ib, ie), same formulas as the single path
// (src/lp/pdlp.cpp:1076-1202): the kNumColFields column fields then the
// kNumRowFields row fields, stored to dst by thread 0.
//
// kStage (whole evaluations with 2m and n at most kEvalStageElems): the
// gathered vectors are first copied into LDS, y^ and y0 for the column
// pass, then x^ for the row pass, and gathered from there.  With thousands
// of instances in flight their gathered vectors do not stay in L2; the
// values, and so the fields, are the same.
//
// cert: also the certificate terms (CertColField, CertRowField) after the
// kNumEvalCore fields; otherwise those fields are left unwritten.
template <int GC, int GR, typename T, bool kStage = false>
__device__ inline void bEvalFields(const BInst& in, const BVecs<T>& v, const BCtrl& c,
                                   double* dst, int jb, int je, int ib, int ie, bool cert) {
    const int tid = threadIdx.x;
    const double ninf = -INFINITY;
    double cc[kNumCertCol] = {0.0, ninf, 0.0, ninf, 0.0};
    double crr[kNumCertRow] = {0.0, ninf, 0.0, ninf};
    T* stage = kStage ? evalStage<T>() : nullptr;
    if constexpr (kStage) {
        __syncthreads();
        for (int e = tid; e < in.m; e += kBlock) {
            stage[e] = v.yh[e];
            stage[in.m + e] = v.y0[e];
        }
        __syncthreads();
    }
    // Each lane group takes kEvalCW columns (rows) per pass, PC (PR) apart,
    // with their loads issued together: a thread still visits its columns in
    // the order grp, grp + PC, grp + 2 PC, ... and sums each column's entries
    // per lane in storage order, so the fields are the same bits as with one
    // column per pass, with kEvalCW dependent load chains in flight.
    constexpr int PC = kBlock / GC, PR = kBlock / GR;
    const int lc = tid % GC, gc = tid / GC;
    double cf[kNumColFields] = {0, 0, 0, 0, 0, 0};
    // Column starts of the coming pass, loaded while the current pass's
    // entries are in flight (the pass then needs one round trip less);
    // not in the FP32 staged kernel, where the extra registers would cost a
    // wave per SIMD.
    constexpr bool kAhead = !(kStage && sizeof(T) == 4);
    int nb[kEvalCW], ne[kEvalCW];
    auto colStarts = [&](int b) {
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            const int jj = b + gc + w * PC;
            nb[w] = jj < je ? in.cs[jj] : 0;
            ne[w] = jj < je ? in.cs[jj + 1] : 0;
        }
    };
    if constexpr (kAhead) colStarts(jb);
    for (int base = jb; base < je; base += kEvalCW * PC) {
        if constexpr (!kAhead) colStarts(base);
        int j[kEvalCW], qb[kEvalCW], qe[kEvalCW];
        double aty[kEvalCW], atdy[kEvalCW], xj[kEvalCW], x0j[kEvalCW];
        // The column's own operands, loaded with its first loads (lane 0 of
        // the group uses them) instead of after its sums.
        double cj[kEvalCW], lbj[kEvalCW], ubj[kEvalCW], csj[kEvalCW];
        int most = 0;
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            j[w] = base + gc + w * PC;
            const bool ok = j[w] < je;
            qb[w] = ok ? nb[w] + lc : 0;
            qe[w] = ok ? ne[w] : 0;
            xj[w] = ok ? static_cast<double>(v.xh[j[w]]) : 0.0;
            x0j[w] = ok ? static_cast<double>(v.x0[j[w]]) : 0.0;
            const bool own = ok && lc == 0;
            cj[w] = own ? in.c[j[w]] : 0.0;
            lbj[w] = own ? in.l[j[w]] : 0.0;
            ubj[w] = own ? in.u[j[w]] : 0.0;
            csj[w] = own ? in.col_scale[j[w]] : 1.0;
            const int cnt = qe[w] > qb[w] ? (qe[w] - qb[w] + GC - 1) / GC : 0;
            most = cnt > most ? cnt : most;
            aty[w] = atdy[w] = 0.0;
        }
        if constexpr (kAhead) colStarts(base + kEvalCW * PC);
        for (int t = 0; t < most; ++t) {
#pragma unroll
            for (int w = 0; w < kEvalCW; ++w) {
                const int q = qb[w] + t * GC;
                if (q < qe[w]) {
                    const int i = in.ri[q];
                    const double a = in.valt[q];
                    const double yi = static_cast<double>(kStage ? stage[i] : v.yh[i]);
                    const double y0i = static_cast<double>(kStage ? stage[in.m + i] : v.y0[i]);
                    aty[w] += a * yi;
                    atdy[w] += a * (yi - y0i);
                }
            }
        }
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            aty[w] = groupSum<GC>(aty[w]);
            atdy[w] = groupSum<GC>(atdy[w]);
            if (lc == 0 && j[w] < je) {
                const double dx = xj[w] - x0j[w];
                cf[kDx2] += dx * dx;
                cf[kInter] += atdy[w] * dx;
                const double g = cj[w] + aty[w];
                const double lb = lbj[w], ub = ubj[w];
                // Activity against the bounds the iteration projected onto.
                const double lbi = static_cast<double>(static_cast<T>(lb));
                const double ubi = static_cast<double>(static_cast<T>(ub));
                double dr;
                if (xj[w] <= lbi + 1e-12 && isfinite(lb)) dr = minRef(g, 0.0);
                else if (xj[w] >= ubi - 1e-12 && isfinite(ub)) dr = maxRef(g, 0.0);
                else dr = g;
                const double d = dr / (c.os * csj[w]);
                cf[kDsq] += d * d;
                cf[kDmax] = maxRef(cf[kDmax], fabs(d));
                cf[kPobj] += cj[w] * xj[w];
                double gp = g;
                if (!isfinite(lb)) gp = minRef(gp, 0.0);
                if (!isfinite(ub)) gp = maxRef(gp, 0.0);
                if (gp > 0.0 && isfinite(lb)) cf[kDcol] += gp * lb;
                else if (gp < 0.0 && isfinite(ub)) cf[kDcol] += gp * ub;
                if (cert) {
                    // Farkas column terms on y_o = sgn R y^ / os: (A'y_o)_j =
                    // sgn (A~'y^)_j / (C_j os); primal-ray terms on x_o =
                    // C x^ / cs (pdlp.cpp:119-284).
                    const double lo = in.lo_o[j[w]], hi = in.hi_o[j[w]];
                    const double atr = c.sgn * aty[w] / csj[w] / c.os;
                    const double xo = xj[w] * csj[w] * (1.0 / c.cs);
                    cc[kCxsq] += xo * xo;
                    if (isfinite(lo)) {
                        cc[kCfBad] = maxRef(cc[kCfBad], -atr);
                        cc[kCrBad] = maxRef(cc[kCrBad], -xo);
                    }
                    if (isfinite(hi)) {
                        cc[kCfBad] = maxRef(cc[kCfBad], atr);
                        cc[kCrBad] = maxRef(cc[kCrBad], xo);
                    }
                    if (atr > 0.0 && isfinite(lo)) cc[kCcol] += atr * lo;
                    else if (atr < 0.0 && isfinite(hi)) cc[kCcol] += atr * hi;
                    cc[kCobj] += in.obj_o[j[w]] * xo;
                }
            }
        }
    }
    // Block sums of the column fields now (each field reduces on its own,
    // so reducing them before the row pass changes no bits).
    {
        bool mx[kNumColFields];
        double out[kNumColFields];
        for (int f = 0; f < kNumColFields; ++f) mx[f] = f == kDmax;
        blockReduce<kNumColFields>(cf, mx, out);
        if (tid == 0)
            for (int f = 0; f < kNumColFields; ++f) dst[f] = out[f];
    }
    if (cert) {
        bool mx[kNumCertCol];
        double out[kNumCertCol];
        for (int f = 0; f < kNumCertCol; ++f) mx[f] = evalFieldIsMax(kCertCol0 + f);
        blockReduce<kNumCertCol>(cc, mx, out);
        if (tid == 0)
            for (int f = 0; f < kNumCertCol; ++f) dst[kCertCol0 + f] = out[f];
    }
    if constexpr (kStage) {
        __syncthreads();
        for (int e = tid; e < in.n; e += kBlock) stage[e] = v.xh[e];
        __syncthreads();
    }
    const int lr = tid % GR, gr = tid / GR;
    double rf[kNumRowFields] = {0, 0, 0, 0};
    auto rowStarts = [&](int b) {
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            const int ii = b + gr + w * PR;
            nb[w] = ii < ie ? in.rs[ii] : 0;
            ne[w] = ii < ie ? in.rs[ii + 1] : 0;
        }
    };
    if constexpr (kAhead) rowStarts(ib);
    for (int base = ib; base < ie; base += kEvalCW * PR) {
        if constexpr (!kAhead) rowStarts(base);
        int i[kEvalCW], qb[kEvalCW], qe[kEvalCW];
        double ax[kEvalCW], yv[kEvalCW], y0v[kEvalCW], loi[kEvalCW], hii[kEvalCW], rsi[kEvalCW];
        int most = 0;
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            i[w] = base + gr + w * PR;
            const bool ok = i[w] < ie;
            qb[w] = ok ? nb[w] + lr : 0;
            qe[w] = ok ? ne[w] : 0;
            yv[w] = ok ? static_cast<double>(v.yh[i[w]]) : 0.0;
            y0v[w] = ok ? static_cast<double>(v.y0[i[w]]) : 0.0;
            const bool own = ok && lr == 0;
            loi[w] = own ? in.rl[i[w]] : 0.0;
            hii[w] = own ? in.ru[i[w]] : 0.0;
            rsi[w] = own ? in.row_scale[i[w]] : 1.0;
            const int cnt = qe[w] > qb[w] ? (qe[w] - qb[w] + GR - 1) / GR : 0;
            most = cnt > most ? cnt : most;
            ax[w] = 0.0;
        }
        if constexpr (kAhead) rowStarts(base + kEvalCW * PR);
        for (int t = 0; t < most; ++t) {
#pragma unroll
            for (int w = 0; w < kEvalCW; ++w) {
                const int q = qb[w] + t * GR;
                if (q < qe[w]) {
                    const int col = in.ci[q];
                    ax[w] += in.val[q] * static_cast<double>(kStage ? stage[col] : v.xh[col]);
                }
            }
        }
#pragma unroll
        for (int w = 0; w < kEvalCW; ++w) {
            ax[w] = groupSum<GR>(ax[w]);
            if (lr == 0 && i[w] < ie) {
                const double lo = loi[w], hi = hii[w];
                const double vs = ax[w] - clampRef(ax[w], lo, hi);
                const double p = vs / (c.cs * rsi[w]);
                rf[kPsq] += p * p;
                rf[kPmax] = maxRef(rf[kPmax], fabs(p));
                const double dy = yv[w] - y0v[w];
                rf[kDy2] += dy * dy;
                double yp = yv[w];
                if (!isfinite(hi)) yp = minRef(yp, 0.0);
                if (!isfinite(lo)) yp = maxRef(yp, 0.0);
                if (yp > 0.0 && isfinite(hi)) rf[kDrow] -= yp * hi;
                else if (yp < 0.0 && isfinite(lo)) rf[kDrow] -= yp * lo;
                if (cert) {
                    const double rlo = in.rl_o[i[w]], rhi = in.ru_o[i[w]];
                    const bool fl = isfinite(rlo), fh = isfinite(rhi);
                    const double yo = c.sgn * yv[w] * rsi[w] * (1.0 / c.os);
                    crr[kRysq] += yo * yo;
                    if (fh && !fl) crr[kRfBad] = maxRef(crr[kRfBad], -yo);
                    if (fl && !fh) crr[kRfBad] = maxRef(crr[kRfBad], yo);
                    if (rlo == rhi) crr[kRrow] += yo * rlo;
                    else if (yo > 0.0 && fh) crr[kRrow] += yo * rhi;
                    else if (yo < 0.0 && fl) crr[kRrow] += yo * rlo;
                    const double ar = ax[w] / rsi[w] / c.cs;
                    if (fl && fh) crr[kRrBad] = maxRef(crr[kRrBad], fabs(ar));
                    else if (fh) crr[kRrBad] = maxRef(crr[kRrBad], ar);
                    else if (fl) crr[kRrBad] = maxRef(crr[kRrBad], -ar);
                }
            }
        }
    }
    {
        bool mx[kNumRowFields];
        double out[kNumRowFields];
        for (int f = 0; f < kNumRowFields; ++f) mx[f] = f == kPmax;
        blockReduce<kNumRowFields>(rf, mx, out);
        if (tid == 0)
            for (int f = 0; f < kNumRowFields; ++f) dst[kNumColFields + f] = out[f];
    }
    if (cert) {
        bool mx[kNumCertRow];
        double out[kNumCertRow];
        for (int f = 0; f < kNumCertRow; ++f) mx[f] = evalFieldIsMax(kCertRow0 + f);
        blockReduce<kNumCertRow>(crr, mx, out);
        if (tid == 0)
            for (int f = 0; f < kNumCertRow; ++f) dst[kCertRow0 + f] = out[f];
    }
}

// The certificate verdict from an instance's combined certificate terms
// (the checks of src/lp/pdlp.cpp:119-284 with the normalisation by the ray
// norm moved onto the tolerance): 1 = infeasible (Farkas ray), 2 =
// unbounded (primal ray), 0 = neither.
__device__ inline int certVerdict(const double* f, int sense_max) {
    const double* cc = f + kCertCol0;
    const double* cr = f + kCertRow0;
    const double tol = 1e-6;
    const double ny = sqrt(cr[kRysq]);
    if (ny >= 1e-15) {
        const bool bad = cr[kRfBad] > tol * ny || cc[kCfBad] > tol * ny;
        if (!bad && cr[kRrow] / ny - cc[kCcol] / ny < -tol) return 1;
    }
    const double nx = sqrt(cc[kCxsq]);
    if (nx < 1e-15) return 0;
    const bool bad = cc[kCrBad] > tol * nx || cr[kRrBad] > tol * nx;
    const double sg = sense_max ? -1.0 : 1.0;
    if (!bad && sg * (cc[kCobj] / nx) < -tol) return 2;
    return 0;
}

// The termination quantities from the evaluation sums f (column fields, then
// row fields) and the instance's scalars.
__device__ inline BEval bEvalFinish(const double* f, const BCtrl& c) {
    const double* mc = f;
    const double* mr = f + kNumColFields;
    BEval e;
    e.dx2 = mc[kDx2];
    e.dy2 = mr[kDy2];
    const double movement = e.dx2 * c.pw + e.dy2 / c.pw;
    const double interaction = mc[kInter] * (2.0 * c.step);
    e.fpe = sqrt(maxRef(movement + interaction, 0.0));
    e.rel_p = sqrt(mr[kPsq]) / (1.0 + c.bnorm);
    e.rel_p_inf = mr[kPmax] / (1.0 + c.bnorm);
    e.rel_d = sqrt(mc[kDsq]) / (1.0 + c.cnorm);
    e.rel_d_inf = mc[kDmax] / (1.0 + c.cnorm);
    const double unscale = 1.0 / (c.os * c.cs);
    e.pobj = c.offset + c.sgn * mc[kPobj] * unscale;
    e.dobj = c.offset + c.sgn * (mc[kDcol] + mr[kDrow]) * unscale;
    e.gap = fabs(e.pobj - e.dobj) / (1.0 + fabs(e.pobj) + fabs(e.dobj));
    return e;
}

// Re-centring of an instance's refinement (one workgroup): the FP64 base
// (x_b in d.x, y_b in d.y) is already in place; forms c' = c + A'y_b,
// l' = l - x_b, u' = u - x_b, rl' = rl - A x_b, ru' = ru - A x_b and q = y_b
// in FP64 (sequential sums per column / row) rounded to FP32, points the
// FP32 iteration vectors at them and zeroes the correction iterates.
__device__ inline void refineRecenter(const BInst& in) {
    BInst& w = const_cast<BInst&>(in);
    const int tid = threadIdx.x;
    float* lf = const_cast<float*>(in.f.l);
    float* uf = const_cast<float*>(in.f.u);
    for (int j = tid; j < in.n; j += kBlock) {
        double g = 0.0;
        for (int k = in.cs[j]; k < in.cs[j + 1]; ++k) g += in.valt[k] * in.d.y[in.ri[k]];
        const double xb = in.d.x[j];
        in.rc[j] = static_cast<float>(in.c[j] + g);
        lf[j] = static_cast<float>(in.l[j] - xb);
        uf[j] = static_cast<float>(in.u[j] - xb);
        in.f.x[j] = in.f.x0[j] = in.f.xh[j] = 0.0f;
    }
    for (int i = tid; i < in.m; i += kBlock) {
        double r = 0.0;
        for (int k = in.rs[i]; k < in.rs[i + 1]; ++k) r += in.val[k] * in.d.x[in.ci[k]];
        in.rrl[i] = static_cast<float>(in.rl[i] - r);
        in.rru[i] = static_cast<float>(in.ru[i] - r);
        in.rq[i] = static_cast<float>(in.d.y[i]);
        in.f.y[i] = in.f.y0[i] = in.f.yh[i] = 0.0f;
    }
    if (tid == 0) {
        w.f.c = in.rc;
        w.f.rl = in.rrl;
        w.f.ru = in.rru;
        w.f.q = in.rq;
    }
    __syncthreads();
}

// Column slice [jb, je) and row slice [ib, ie) of a re-centring whose FP64
// base (x_b in d.x, y_b in d.y) is in place: the same per-element sums and
// roundings as refineRecenter (one thread per column / row, entries in
// storage order), so a parted re-centring is bit-identical to it.  The FP32
// pointer switch (f.c, f.rl, f.ru, f.q) is the control kernel's.
__device__ inline void refineRecenterSlice(const BInst& in, int jb, int je, int ib, int ie) {
    float* lf = const_cast<float*>(in.f.l);
    float* uf = const_cast<float*>(in.f.u);
    // A thread takes its columns (rows) kEvalCW at a time, their entry
    // chains in flight together; each sum still runs over its own entries
    // in storage order.
    constexpr int W = kEvalCW;
    const int tid = static_cast<int>(threadIdx.x);
    for (int j0 = jb + tid; j0 < je; j0 += W * kBlock) {
        int cb[W], ce[W];
        double g[W];
        int most = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int j = j0 + w * kBlock;
            cb[w] = j < je ? in.cs[j] : 0;
            ce[w] = j < je ? in.cs[j + 1] : 0;
            most = ce[w] - cb[w] > most ? ce[w] - cb[w] : most;
            g[w] = 0.0;
        }
        for (int t = 0; t < most; ++t) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const int k = cb[w] + t;
                if (k < ce[w]) g[w] += in.valt[k] * in.d.y[in.ri[k]];
            }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int j = j0 + w * kBlock;
            if (j >= je) break;
            const double xb = in.d.x[j];
            in.rc[j] = static_cast<float>(in.c[j] + g[w]);
            lf[j] = static_cast<float>(in.l[j] - xb);
            uf[j] = static_cast<float>(in.u[j] - xb);
            in.f.x[j] = in.f.x0[j] = in.f.xh[j] = 0.0f;
        }
    }
    for (int i0 = ib + tid; i0 < ie; i0 += W * kBlock) {
        int cb[W], ce[W];
        double r[W];
        int most = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int i = i0 + w * kBlock;
            cb[w] = i < ie ? in.rs[i] : 0;
            ce[w] = i < ie ? in.rs[i + 1] : 0;
            most = ce[w] - cb[w] > most ? ce[w] - cb[w] : most;
            r[w] = 0.0;
        }
        for (int t = 0; t < most; ++t) {
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const int k = cb[w] + t;
                if (k < ce[w]) r[w] += in.val[k] * in.d.x[in.ci[k]];
            }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int i = i0 + w * kBlock;
            if (i >= ie) break;
            in.rrl[i] = static_cast<float>(in.rl[i] - r[w]);
            in.rru[i] = static_cast<float>(in.ru[i] - r[w]);
            in.rq[i] = static_cast<float>(in.d.y[i]);
            in.f.y[i] = in.f.y0[i] = in.f.yh[i] = 0.0f;
        }
    }
}

// Copies of a pending decision over one column / row slice (element-wise,
// so parts never read what another part writes).
template <typename T>
__device__ inline void pendingCopySlice(const BInst& in, int code, int jb, int je, int ib,
                                        int ie) {
    const int t0 = static_cast<int>(threadIdx.x);
    switch (code) {
        case kPendPolish:
        case kPendRefineToFp64:
            for (int j = jb + t0; j < je; j += kBlock) in.d.x0[j] = in.d.x[j] = in.d.xh[j];
            for (int i = ib + t0; i < ie; i += kBlock) in.d.y0[i] = in.d.y[i] = in.d.yh[i];
            break;
        case kPendToFp64:
            for (int j = jb + t0; j < je; j += kBlock) {
                const double h = static_cast<double>(in.f.xh[j]);
                in.d.xh[j] = h;
                in.d.x[j] = h;
                in.d.x0[j] = h;
            }
            for (int i = ib + t0; i < ie; i += kBlock) {
                const double h = static_cast<double>(in.f.yh[i]);
                in.d.yh[i] = h;
                in.d.y[i] = h;
                in.d.y0[i] = h;
            }
            break;
        case kPendToRefine:
            for (int j = jb + t0; j < je; j += kBlock) in.d.x[j] = static_cast<double>(in.f.xh[j]);
            for (int i = ib + t0; i < ie; i += kBlock) in.d.y[i] = static_cast<double>(in.f.yh[i]);
            break;
        case kPendRecenter:
            for (int j = jb + t0; j < je; j += kBlock) in.d.x[j] = in.d.xh[j];
            for (int i = ib + t0; i < ie; i += kBlock) in.d.y[i] = in.d.yh[i];
            break;
        case kPendRestart: {
            const BVecs<T>* vp;
            if constexpr (sizeof(T) == 4) vp = &in.f;
            else vp = &in.d;
            const BVecs<T>& v = *vp;
            for (int j = jb + t0; j < je; j += kBlock) {
                const T h = v.xh[j];
                v.x[j] = h;
                v.x0[j] = h;
            }
            for (int i = ib + t0; i < ie; i += kBlock) {
                const T h = v.yh[i];
                v.y[i] = h;
                v.y0[i] = h;
            }
            break;
        }
        default:
            break;
    }
}

// The FP64 evaluation point of a refining instance over one column / row
// slice, element-wise: base plus correction; a column whose correction sits
// on its shifted bound takes the FP64 bound itself, and a ranged row whose
// correction cancels the base multiplier takes zero.
__device__ inline void refineMaterializeSlice(const BInst& in, int jb, int je, int ib, int ie) {
    for (int j = jb + static_cast<int>(threadIdx.x); j < je; j += kBlock) {
        const float d = in.f.xh[j];
        const double xb = in.d.x[j];
        in.d.xh[j] = d == in.f.l[j] ? in.l[j] : (d == in.f.u[j] ? in.u[j] : xb + static_cast<double>(d));
        in.d.x0[j] = xb + static_cast<double>(in.f.x0[j]);
    }
    for (int i = ib + static_cast<int>(threadIdx.x); i < ie; i += kBlock) {
        const float d = in.f.yh[i];
        const bool inactive = in.rrl[i] != in.rru[i] && d == -in.rq[i];
        in.d.yh[i] = inactive ? 0.0 : in.d.y[i] + static_cast<double>(d);
        in.d.y0[i] = in.d.y[i] + static_cast<double>(in.f.y0[i]);
    }
}

}  // namespace mpx
