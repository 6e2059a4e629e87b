// pdlp_batch.hip -- batched PDLP (many LPs per launch) and the
// continuous-batching serving loop.
//
// A round is N = termination_eval_frequency iterations of every running
// instance followed by its evaluation and control, captured once in a
// hipGraph and replayed:
//   * instances that share one constraint matrix (branch-and-bound node
//     relaxations) iterate in the lane-per-instance kernels of
//     batch_interleaved.h: 64 instances per group, one lane each, so a wave
//     reads the shared matrix once for 64 instances;
//   * independent LPs iterate in packed tiles (instance, block of columns or
//     rows) over a fixed grid, each with its own lane-group preset;
//   * kEvalScan / kEvalParts / kEvalControl evaluate every instance in
//     size-based parts and run the reference host loop's termination,
//     certificate, restart, primal-weight and polish decisions per instance
//     (src/lp/pdlp.cpp:2228-2676) on the device; kControlCopy /
//     kControlRecenter then do the decisions' vector work (restart copies,
//     phase changes, the refinement's re-centring) in the same parts.
// The host reads one count of unfinished instances per round.
//
// Setup is on the device too: the scaling (Ruiz, Pock-Chambolle,
// bound/objective rescaling; bit-identical per element to the reference's
// buildScaledProblem, pdlp.cpp:469-669) and the power iteration
// (pdlp.cpp:748-795) run as tile passes over a cohort of matrices
// (batch_setup.h).  Instances that share A share all of it and only scale
// their own column bounds.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mipx_pdlp.h"
#include "batch_interleaved.h"
#include "batch_kernels.h"
#include "batch_setup.h"
#include "packed_presets.h"
#include "device_common.h"

namespace mpx {


// Heuristic preset from mean entries per column / row (about two entries per
// lane), for LP shapes the autotuned table (packed_presets.h) has no entry for.
int heuristicPreset(long long nnz, int m, int n) {
    const double ac = n ? static_cast<double>(nnz) / n : 1.0;
    const double ar = m ? static_cast<double>(nnz) / m : 1.0;
    int best = 0;
    double bestScore = 1e30;
    for (int p = 0; p < kNumPresets; ++p) {
        const double sc = std::abs(std::log2(2.0 * presetGC(p) / ac)) +
                          std::abs(std::log2(2.0 * presetGR(p) / ar));
        if (sc < bestScore) {
            bestScore = sc;
            best = p;
        }
    }
    return best;
}

// ---- control logic (thread 0; the reference's host loop) -------------------
__device__ inline double kktRatio(const BEval& e, const BOptions& o) {
    const double tp = o.norm_inf ? e.rel_p_inf : e.rel_p;
    const double td = o.norm_inf ? e.rel_d_inf : e.rel_d;
    return maxRef(maxRef(tp / o.primal_tol, td / o.dual_tol), e.gap / o.opt_tol);
}

__device__ inline void updatePrimalWeight(BCtrl& c, const BEval& e, const BOptions& o,
                                          bool polish) {
    const double pd = sqrt(e.dx2), dd = sqrt(e.dy2);
    const double ratio = (e.rel_p > 1e-15) ? e.rel_d / e.rel_p : 0.0;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = log(dd) - log(pd) - log(c.pw);
        c.pw_err_sum = o.pid_i_smooth * c.pw_err_sum + err;
        const double derr = err - c.pw_last_err;
        c.pw *= exp(o.pid_kp * err + o.pid_ki * c.pw_err_sum + o.pid_kd * derr);
        c.pw = clampRef(c.pw, 1e-4, 1e8);
        c.pw_last_err = err;
        if (!polish) {
            const double score =
                fabs(log10(maxRef(e.rel_d, 1e-15)) - log10(maxRef(e.rel_p, 1e-15)));
            if (score < c.best_score) {
                c.best_score = score;
                c.best_pw = c.pw;
            }
        }
    } else if (!polish) {
        c.pw = c.best_pw;
        c.pw_err_sum = 0.0;
        c.pw_last_err = 0.0;
    }
}

// Decision codes written by thread 0 for the workgroup.
enum BAction { kActNone = 0, kActRestart = 1, kActToFp64 = 2, kActConverged = 3, kActFinish = 4,
               kActCert = 5, kActToRefine = 6, kActRefineToFp64 = 7 };

// Mixed precision: FP32 hands over to FP64 once the FP64-judged KKT ratio
// (max of residual / tolerance) is within this factor, or after this many
// evaluations without halving it.
constexpr double kFp32ExitRatio = 3.0;
constexpr int kFp32Stall = 20;
// A refining instance that goes this many evaluations without improving its
// KKT ratio by 1% continues in FP64 (a precision stall of the FP32 matrix;
// the seeded families reach 1e-8 in the correction phase without one).
constexpr int kRefineStall = 10 * kFp32Stall;

template <typename T>
__device__ inline const BVecs<T>& pickVecs(const BInst& in);
template <>
__device__ inline const BVecs<float>& pickVecs<float>(const BInst& in) { return in.f; }
template <>
__device__ inline const BVecs<double>& pickVecs<double>(const BInst& in) { return in.d; }

__global__ void kZeroInt(int* p) { *p = 0; }

__global__ void kStampStart(unsigned long long* t) { *t = wall_clock64(); }

// Completion stamp of an instance (device wall clock), first time it is done.
__device__ inline void stampDone(BCtrl& c) {
    if (c.phase == kPhaseDone && c.t_finish == 0ull) c.t_finish = wall_clock64();
}

// ---- packed iteration kernels ----------------------------------------------
// All active instances of one precision form one flat list of tiles; a tile
// is (instance, first column) covering kBlock/G columns (or rows) of that
// instance.  A fixed grid walks the list (persistent), so one captured launch
// serves any number of active instances.  XCD-aware dealing: workgroup b runs
// on XCD b % 8; tiles come in chunks of (grid / 8) consecutive tiles and
// chunk q is dealt to the workgroups of XCD q % 8, so the tiles of one
// instance share one XCD's L2 for the gathered y / x_r slice.  Lists shorter
// than one chunk per XCD are dealt round-robin over the whole grid.
struct Tile {
    int32_t inst;
    int32_t first;
};

template <typename F>
__device__ inline void forTiles(int cnt, F&& f) {
    const int nslot = static_cast<int>(gridDim.x) / kNumXcd;
    if (cnt < kNumXcd * nslot) {
        for (int t = blockIdx.x; t < cnt; t += gridDim.x) f(t);
        return;
    }
    const int xcd = blockIdx.x % kNumXcd, slot = blockIdx.x / kNumXcd;
    for (int ch = xcd; ch * nslot < cnt; ch += kNumXcd) {
        const int t = ch * nslot + slot;
        if (t < cnt) f(t);
    }
}

// Per-round record of an active packed instance in the precision it runs
// this round: every pointer and scalar a tile needs, written once per round
// by kTileFill, so that a tile costs one indirection instead of a chain of
// dependent loads.  step and pw are rounded to T exactly as the iteration
// uses them.
template <typename T>
struct PRec {
    const int *cs, *ri, *rs, *ci;
    const T *valt, *val, *c, *tau, *l, *u, *x0, *sig, *rl, *ru, *y0;
    T *x, *xr, *xh, *y, *yh;
    const T* q;  // refinement row shift, or null
    T step, pw;
    int32_t inner, lim, n, m;
};

// Columns (rows) per lane group in a packed tile: their loads are issued
// together; each column still sums its entries per lane in storage order, so
// results do not depend on this constant.
constexpr int kPackCW = 4;

template <int G, typename T>
__device__ inline void primalTile(const PRec<T>& rec, int first, int kiter) {
    if (kiter > rec.lim) return;
    struct {
        decltype(asGlobal(rec.cs)) cs;
        decltype(asGlobal(rec.ri)) ri;
        decltype(asGlobal(rec.valt)) valt, c, tau, l, u, x0;
        decltype(asGlobal(rec.x)) y, x, xr, xh;
        T step, pw;
        int32_t inner, lim, n;
    } r{asGlobal(rec.cs), asGlobal(rec.ri), asGlobal(rec.valt), asGlobal(rec.c), asGlobal(rec.tau),
        asGlobal(rec.l), asGlobal(rec.u), asGlobal(rec.x0), asGlobal(rec.y), asGlobal(rec.x),
        asGlobal(rec.xr), asGlobal(rec.xh), rec.step, rec.pw, rec.inner, rec.lim, rec.n};
    constexpr int per = kBlock / G;
    const int lane = static_cast<int>(threadIdx.x) % G, grp = static_cast<int>(threadIdx.x) / G;
    int j[kPackCW], b[kPackCW], cnt[kPackCW];
    T acc[kPackCW];
    int most = 0;
#pragma unroll
    for (int q = 0; q < kPackCW; ++q) {
        j[q] = first + grp + q * per;
        const bool ok = j[q] < r.n;
        const int bb = ok ? r.cs[j[q]] : 0, ee = ok ? r.cs[j[q] + 1] : 0;
        b[q] = bb + lane;
        cnt[q] = ee > b[q] ? (ee - b[q] + G - 1) / G : 0;
        most = cnt[q] > most ? cnt[q] : most;
        acc[q] = T(0);
    }
    for (int s = 0; s < most; ++s) {
#pragma unroll
        for (int q = 0; q < kPackCW; ++q)
            if (s < cnt[q]) {
                const int k = b[q] + s * G;
                acc[q] += r.valt[k] * r.y[r.ri[k]];
            }
    }
    const int kk = r.inner + kiter;
    const T lam = static_cast<T>(static_cast<double>(kk) / static_cast<double>(kk + 1));
#pragma unroll
    for (int q = 0; q < kPackCW; ++q) {
        const T aty = groupSum<G>(acc[q]);
        if (lane != 0 || j[q] >= r.n) continue;
        const int jj = j[q];
        const T tau = r.step * r.tau[jj] / r.pw;
        const T xj = r.x[jj];
        const T tt = xj - tau * (r.c[jj] + aty);
        const T h = clampRef(tt, r.l[jj], r.u[jj]);
        const T rr = T(2) * h - xj;
        if (kiter == r.lim) r.xh[jj] = h;
        r.xr[jj] = rr;
        r.x[jj] = lam * rr + (T(1) - lam) * r.x0[jj];
    }
}

template <int G, typename T>
__device__ inline void dualTile(const PRec<T>& rec, int first, int kiter) {
    if (kiter > rec.lim) return;
    struct {
        decltype(asGlobal(rec.rs)) rs;
        decltype(asGlobal(rec.ci)) ci;
        decltype(asGlobal(rec.val)) val, sig, rl, ru, y0, q;
        decltype(asGlobal(rec.xr)) xr, y, yh;
        T step, pw;
        int32_t inner, lim, m;
    } r{asGlobal(rec.rs), asGlobal(rec.ci), asGlobal(rec.val), asGlobal(rec.sig), asGlobal(rec.rl),
        asGlobal(rec.ru), asGlobal(rec.y0), asGlobal(rec.q), asGlobal(rec.xr), asGlobal(rec.y),
        asGlobal(rec.yh), rec.step, rec.pw, rec.inner, rec.lim, rec.m};
    constexpr int per = kBlock / G;
    const int lane = static_cast<int>(threadIdx.x) % G, grp = static_cast<int>(threadIdx.x) / G;
    int i[kPackCW], b[kPackCW], cnt[kPackCW];
    T acc[kPackCW];
    int most = 0;
#pragma unroll
    for (int q = 0; q < kPackCW; ++q) {
        i[q] = first + grp + q * per;
        const bool ok = i[q] < r.m;
        const int bb = ok ? r.rs[i[q]] : 0, ee = ok ? r.rs[i[q] + 1] : 0;
        b[q] = bb + lane;
        cnt[q] = ee > b[q] ? (ee - b[q] + G - 1) / G : 0;
        most = cnt[q] > most ? cnt[q] : most;
        acc[q] = T(0);
    }
    for (int s = 0; s < most; ++s) {
#pragma unroll
        for (int q = 0; q < kPackCW; ++q)
            if (s < cnt[q]) {
                const int k = b[q] + s * G;
                acc[q] += r.val[k] * r.xr[r.ci[k]];
            }
    }
    const int kk = r.inner + kiter;
    const T lam = static_cast<T>(static_cast<double>(kk) / static_cast<double>(kk + 1));
#pragma unroll
    for (int q = 0; q < kPackCW; ++q) {
        const T axr = groupSum<G>(acc[q]);
        if (lane != 0 || i[q] >= r.m) continue;
        const int ii = i[q];
        const T sig = r.step * r.sig[ii] * r.pw;
        const T yi = r.y[ii];
        const T vv = yi + sig * axr;
        const T h = r.q ? dualProxShifted(vv, sig, r.rl[ii], r.ru[ii], r.q[ii])
                        : dualProx(vv, sig, r.rl[ii], r.ru[ii]);
        const T rr = T(2) * h - yi;
        if (kiter == r.lim) r.yh[ii] = h;
        r.y[ii] = lam * rr + (T(1) - lam) * r.y0[ii];
    }
}

// The lane-group size of a tile is its instance's own preset, so an
// instance's arithmetic does not depend on the rest of the batch.
template <typename T>
__global__ __launch_bounds__(kBlock) void kPackedPrimal(const BInst* __restrict__ insts,
                                                         const PRec<T>* __restrict__ recs,
                                                         const Tile* __restrict__ tiles,
                                                         const int* __restrict__ total, int kiter) {
    forTiles(*total, [&](int t) {
        const Tile tl = tiles[t];
        const PRec<T>& r = recs[tl.inst];
        switch (presetGC(insts[tl.inst].preset)) {
            case 1: primalTile<1, T>(r, tl.first, kiter); break;
            case 2: primalTile<2, T>(r, tl.first, kiter); break;
            case 4: primalTile<4, T>(r, tl.first, kiter); break;
            case 8: primalTile<8, T>(r, tl.first, kiter); break;
            default: primalTile<16, T>(r, tl.first, kiter); break;
        }
    });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kPackedDual(const BInst* __restrict__ insts,
                                                       const PRec<T>* __restrict__ recs,
                                                       const Tile* __restrict__ tiles,
                                                       const int* __restrict__ total, int kiter) {
    forTiles(*total, [&](int t) {
        const Tile tl = tiles[t];
        const PRec<T>& r = recs[tl.inst];
        switch (presetGR(insts[tl.inst].preset)) {
            case 1: dualTile<1, T>(r, tl.first, kiter); break;
            case 2: dualTile<2, T>(r, tl.first, kiter); break;
            case 4: dualTile<4, T>(r, tl.first, kiter); break;
            case 8: dualTile<8, T>(r, tl.first, kiter); break;
            default: dualTile<16, T>(r, tl.first, kiter); break;
        }
    });
}

// Columns / rows covered by one packed tile.
__host__ __device__ constexpr int tileCols(int preset) { return kBlock / presetGC(preset) * kPackCW; }
__host__ __device__ constexpr int tileRows(int preset) { return kBlock / presetGR(preset) * kPackCW; }

// Per-instance tile counts for the four lists (FP32 columns, FP32 rows,
// FP64 columns, FP64 rows) of the coming round.
__global__ void kTileCounts(const BInst* __restrict__ insts, int K, int* __restrict__ cnt) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= K) return;
    const BInst& in = insts[k];
    const int cpt = tileCols(in.preset), rpt = tileRows(in.preset);
    const BCtrl& c = *in.ctrl;
    const bool act = !in.il && c.phase != kPhaseDone && c.round_iters > 0;
    const bool p32 = act && iteratesFp32(c.phase), p64 = act && !iteratesFp32(c.phase);
    const int ct = (in.n + cpt - 1) / cpt, rt = (in.m + rpt - 1) / rpt;
    cnt[0 * K + k] = p32 ? ct : 0;
    cnt[1 * K + k] = p32 ? rt : 0;
    cnt[2 * K + k] = p64 ? ct : 0;
    cnt[3 * K + k] = p64 ? rt : 0;
}

// In-place exclusive scan of the four count arrays (one block), totals[4];
// totals[4] = number of unfinished instances.
__global__ __launch_bounds__(1024) void kScan4(int* __restrict__ cnt, int K,
                                               const BInst* __restrict__ insts,
                                               int* __restrict__ totals) {
    __shared__ int part[1024];
    __shared__ int carry;
    for (int a = 0; a < 4; ++a) {
        int* arr = cnt + static_cast<size_t>(a) * K;
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int base = 0; base < K; base += 1024) {
            const int k = base + threadIdx.x;
            const int v = k < K ? arr[k] : 0;
            part[threadIdx.x] = v;
            __syncthreads();
            for (int off = 1; off < 1024; off <<= 1) {
                const int add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
                __syncthreads();
                part[threadIdx.x] += add;
                __syncthreads();
            }
            if (k < K) arr[k] = carry + part[threadIdx.x] - v;
            __syncthreads();
            if (threadIdx.x == 0) carry += part[1023];
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[a] = carry;
        __syncthreads();
    }
    // Unfinished instances (order-independent integer count).
    int act = 0;
    for (int k = threadIdx.x; k < K; k += 1024) act += insts[k].ctrl->phase != kPhaseDone;
    part[threadIdx.x] = act;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[4] = part[0];
}

template <typename T>
__device__ inline void fillRec(const BInst& in, const BCtrl& c, PRec<T>& r) {
    const BVecs<T>& v = pickVecs<T>(in);
    r.cs = in.cs;
    r.ri = in.ri;
    r.rs = in.rs;
    r.ci = in.ci;
    r.valt = v.valt;
    r.val = v.val;
    r.c = v.c;
    r.tau = v.tau;
    r.l = v.l;
    r.u = v.u;
    r.x0 = v.x0;
    r.sig = v.sig;
    r.rl = v.rl;
    r.ru = v.ru;
    r.y0 = v.y0;
    r.x = v.x;
    r.xr = v.xr;
    r.xh = v.xh;
    r.y = v.y;
    r.yh = v.yh;
    r.q = v.q;
    r.step = static_cast<T>(c.step);
    r.pw = static_cast<T>(c.pw);
    r.inner = c.inner;
    r.lim = c.round_iters;
    r.n = in.n;
    r.m = in.m;
}

// Tile lists and per-round records of the active packed instances.
__global__ void kTileFill(const BInst* __restrict__ insts, int K, const int* __restrict__ off,
                          Tile* __restrict__ c32, Tile* __restrict__ r32, Tile* __restrict__ c64,
                          Tile* __restrict__ r64, PRec<float>* __restrict__ rec32,
                          PRec<double>* __restrict__ rec64) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= K) return;
    const BInst& in = insts[k];
    const int cpt = tileCols(in.preset), rpt = tileRows(in.preset);
    const BCtrl& c = *in.ctrl;
    const bool act = !in.il && c.phase != kPhaseDone && c.round_iters > 0;
    if (!act) return;
    const bool p32 = iteratesFp32(c.phase);
    if (p32) fillRec<float>(in, c, rec32[k]);
    else fillRec<double>(in, c, rec64[k]);
    Tile* cl = p32 ? c32 : c64;
    Tile* rl = p32 ? r32 : r64;
    const int oc = off[(p32 ? 0 : 2) * K + k], orr = off[(p32 ? 1 : 3) * K + k];
    const int ct = (in.n + cpt - 1) / cpt, rt = (in.m + rpt - 1) / rpt;
    for (int t = 0; t < ct; ++t) cl[oc + t] = Tile{k, t * cpt};
    for (int t = 0; t < rt; ++t) rl[orr + t] = Tile{k, t * rpt};
}

// ---- per-instance evaluation and control (one workgroup per instance) --------
// Runs once per round after the round's iterations: the FP64 evaluation, the
// termination / certificate / restart / primal-weight decisions of the
// reference's loop (src/lp/pdlp.cpp:2228-2469) and the polish phase switch
// (2477-2676); then sets the instance's iteration count for the next round.
template <typename T>
__device__ void evalControl(const BInst& in, const BOptions& o, const double* parts) {
    BCtrl& c = *in.ctrl;
    const int tid = threadIdx.x;
    constexpr bool kFp32 = sizeof(T) == 4;
    __shared__ int s_act, s_phase, s_ran;
    if (tid == 0) {
        s_phase = c.phase;
        s_ran = c.round_iters;
    }
    __syncthreads();
    if (s_phase == kPhaseDone || kFp32 != (s_phase == kPhaseFp32) || s_ran <= 0) return;
    if (!kFp32) {
        // An instance handed over from FP32 earlier in this round has not
        // iterated in FP64 yet.
        __shared__ int s_skip;
        if (tid == 0) {
            s_skip = c.skip_eval;
            c.skip_eval = 0;
        }
        __syncthreads();
        if (s_skip) return;
    }
    const bool polish = s_phase == kPhasePolish;
    const bool refining = s_phase == kPhaseRefine;
    const int ran = s_ran;
    if (tid == 0) c.pending = kPendNone;
    // The sums of the instance's evaluation parts (kEvalParts, same round),
    // combined in part order: the workgroup stages all parts in LDS with one
    // coalesced pass, then one thread per field adds (or takes the maximum
    // of) that field over the parts in order.  A large LP has up to
    // kEvalMaxParts parts; one thread loading them field by field was a
    // chain of P * kNumEval dependent global loads.
    // The certificate terms are there (and combined) only on the rounds
    // whose decision may check the certificates.
    __shared__ double fs[kNumEval];
    __shared__ double staged[kEvalMaxParts * kNumEval];
    const int P = evalParts(in);
    const int nf = certDue(c, 10 * o.N) ? kNumEval : kNumEvalCore;
    for (int t = tid; t < P * kNumEval; t += kBlock)
        if (t % kNumEval < nf) staged[t] = parts[t];
    __syncthreads();
    if (tid < nf) {
        const int f = tid;
        const bool mx = evalFieldIsMax(f);
        double a = staged[f];
        for (int p = 1; p < P; ++p) {
            const double x = staged[p * kNumEval + f];
            a = mx ? maxRef(a, x) : a + x;
        }
        fs[f] = a;
    }
    __syncthreads();
    const BEval e = bEvalFinish(fs, c);
    __syncthreads();
    if (tid == 0) {
        c.inner += ran;
        c.total += ran;
        if (polish) c.pcount += ran;
        if (kFp32 || refining) c.fp32_iters += ran;
        if (refining) c.refine_iters += ran;
        if (c.fpe0 < 0.0) c.fpe0 = e.fpe;
        c.rel_p = e.rel_p;
        c.rel_d = e.rel_d;
        c.rel_p_inf = e.rel_p_inf;
        c.rel_d_inf = e.rel_d_inf;
        c.gap = e.gap;
        c.pobj = e.pobj;
        c.dobj = e.dobj;
        const double tp = o.norm_inf ? e.rel_p_inf : e.rel_p;
        const double td = o.norm_inf ? e.rel_d_inf : e.rel_d;
        int act = -1;
        if (polish) {
            if (tp <= o.polish_p_tol && td <= o.polish_d_tol) {
                act = kActFinish;
            } else {
                bool restart;
                if (c.pcount == ran) restart = true;
                else if (c.fpe0 > 0.0 && e.fpe <= o.rs_suff * c.fpe0) restart = true;
                else if (c.fpe0 > 0.0 && e.fpe <= o.rs_nec * c.fpe0 && e.fpe > c.last_fpe)
                    restart = true;
                else
                    restart = c.inner >= static_cast<int>(o.rs_art * static_cast<double>(c.pcount));
                c.last_fpe = e.fpe;
                act = kActNone;
                if (restart) {
                    ++c.restarts;
                    if (o.update_pw) updatePrimalWeight(c, e, o, true);
                    act = kActRestart;
                }
            }
        } else if (kFp32) {
            // FP32 iterations judged in FP64: hand over to FP64 once the KKT
            // ratio is within o.handover, or when FP32 stops making progress.
            const double k = kktRatio(e, o);
            bool handover = k <= o.handover;
            if (k < 0.5 * c.best_kkt) {
                c.best_kkt = k;
                c.stall = 0;
            } else if (++c.stall >= kFp32Stall) {
                handover = true;
            }
            if (handover) {
                act = o.refine ? kActToRefine : kActToFp64;
                c.best_kkt = k;
                c.stall = 0;
            }
        } else if (tp <= (o.mixed ? o.refine_p_tol : o.primal_tol) &&
                   td <= (o.mixed ? o.refine_d_tol : o.dual_tol) &&
                   e.gap <= (o.mixed ? o.refine_g_tol : o.opt_tol)) {
            act = kActConverged;
        } else if (refining) {
            // Only a ratio that stops improving (not 1% below its best for
            // kRefineStall evaluations and for half of the phase so far)
            // ends the correction phase.
            const double k = kktRatio(e, o);
            const int window = max(kRefineStall, c.refine_iters / (2 * o.N));
            if (k < 0.99 * c.best_kkt) {
                c.best_kkt = k;
                c.stall = 0;
            } else if (++c.stall >= window) {
                act = kActRefineToFp64;
            }
        }
        if (act == -1 && c.total >= 10 * o.N && c.total % (10 * o.N) == 0) act = kActCert;
        s_act = act;
    }
    __syncthreads();
    int act = s_act;
    if (act == kActCert) {
        const int verdict = certVerdict(fs, in.sense_max);
        if (tid == 0) {
            if (verdict) {
                c.phase = kPhaseDone;
                c.status = verdict == 1 ? MPX_STATUS_INFEASIBLE : MPX_STATUS_UNBOUNDED;
                c.round_iters = 0;
                stampDone(c);
                s_act = kActFinish + 100;
            } else {
                s_act = -1;
            }
        }
        __syncthreads();
        act = s_act;
        if (act == kActFinish + 100) return;
    }
    if (act == -1) {
        if (tid == 0) {
            bool restart;
            if (c.total == o.N) restart = true;
            else if (c.fpe0 > 0.0 && e.fpe <= o.rs_suff * c.fpe0) restart = true;
            else if (c.fpe0 > 0.0 && e.fpe <= o.rs_nec * c.fpe0 && e.fpe > c.last_fpe)
                restart = true;
            else
                restart = c.inner >= static_cast<int>(o.rs_art * static_cast<double>(c.total));
            c.last_fpe = e.fpe;
            s_act = kActNone;
            if (restart) {
                ++c.restarts;
                if (o.update_pw) updatePrimalWeight(c, e, o, false);
                if (o.precond_refresh && c.precond_pw > 0.0 && o.do_pc) {
                    const double drift = c.pw / c.precond_pw;
                    if (drift > o.precond_ratio || drift < 1.0 / o.precond_ratio) c.precond_pw = c.pw;
                }
                s_act = kActRestart;
            }
        }
        __syncthreads();
        act = s_act;
    }
    // The vector work of a decision (copies, the refinement's re-centring)
    // is left to kControlCopy / kControlRecenter, which split every
    // instance into the column and row slices of its evaluation parts, so
    // a large LP's restart does not hold up the round in one workgroup.
    if (act == kActConverged) {
        const bool needPolish = o.enable_polish &&
                                ((o.norm_inf ? e.rel_p_inf : e.rel_p) > o.polish_p_tol ||
                                 (o.norm_inf ? e.rel_d_inf : e.rel_d) > o.polish_d_tol);
        if (tid == 0) {
            if (needPolish) {
                c.pending = kPendPolish;
                c.phase = kPhasePolish;
                c.inner = 0;
                c.pcount = 0;
                c.fpe0 = -1.0;
                c.last_fpe = INFINITY;
            } else {
                c.phase = kPhaseDone;
                c.status = MPX_STATUS_OPTIMAL;
            }
        }
    } else if (act == kActFinish) {
        if (tid == 0) {
            c.phase = kPhaseDone;
            c.status = MPX_STATUS_OPTIMAL;
        }
    } else if (act == kActToFp64) {
        // Continue in FP64 from the FP32 point, restarted there.
        if (tid == 0) {
            c.pending = kPendToFp64;
            c.phase = kPhaseFp64;
            c.skip_eval = 1;
            c.inner = 0;
            c.fpe0 = -1.0;
            c.last_fpe = INFINITY;
            ++c.restarts;
        }
    } else if (act == kActToRefine) {
        // Refinement from the FP32 point: it becomes the FP64 base and the
        // correction starts at zero (the FP32 iteration then reads the
        // correction data).
        if (tid == 0) {
            c.pending = kPendToRefine;
            BInst& w = const_cast<BInst&>(in);
            w.f.c = in.rc;
            w.f.rl = in.rrl;
            w.f.ru = in.rru;
            w.f.q = in.rq;
            c.phase = kPhaseRefine;
            c.refined = 1;
            c.skip_eval = 1;
            c.inner = 0;
            c.fpe0 = -1.0;
            c.last_fpe = INFINITY;
            ++c.restarts;
        }
    } else if (act == kActRefineToFp64) {
        if (tid == 0) {
            c.pending = kPendRefineToFp64;
            c.phase = kPhaseFp64;
            c.inner = 0;
            c.fpe0 = -1.0;
            c.last_fpe = INFINITY;
            ++c.restarts;
        }
    } else if (act == kActRestart && refining) {
        // The base moves to the evaluated point; the correction restarts.
        if (tid == 0) {
            c.pending = kPendRecenter;
            c.inner = 0;
            c.fpe0 = -1.0;
            c.last_fpe = INFINITY;
        }
    } else if (act == kActRestart) {
        if (tid == 0) {
            c.pending = kPendRestart;
            c.inner = 0;
            c.fpe0 = -1.0;
            c.last_fpe = INFINITY;
        }
    }
    __syncthreads();
    // Iterations of the next round (the reference's batch = min(N, budget left)).
    if (tid == 0 && c.phase != kPhaseDone) {
        const bool pol = c.phase == kPhasePolish;
        const int left = pol ? o.polish_max_iter - c.pcount : o.max_iter - c.total;
        if (left <= 0) {
            // Budget spent: polish keeps the converged answer, the main loop
            // stops at the iteration limit.
            c.phase = kPhaseDone;
            c.status = pol ? MPX_STATUS_OPTIMAL : MPX_STATUS_ITER_LIMIT;
            c.round_iters = 0;
        } else {
            c.round_iters = left < o.N ? left : o.N;
        }
    }
    if (tid == 0 && c.phase == kPhaseDone) c.round_iters = 0;
    if (tid == 0) stampDone(c);
}

// ---- evaluation in parts ------------------------------------------------------
// Whether kEvalControl<T> evaluates the instance this round.
template <typename T>
__device__ inline bool evaluatedNow(const BCtrl& c) {
    constexpr bool kFp32 = sizeof(T) == 4;
    const int ph = c.phase;
    return !(ph == kPhaseDone || kFp32 != (ph == kPhaseFp32) || c.round_iters <= 0 ||
             (!kFp32 && c.skip_eval));
}

// off[i] = first part of instance i among the instances kEvalControl<T>
// evaluates this round, off[K] = all parts.  One block, instance order.
template <typename T>
__global__ __launch_bounds__(1024) void kEvalScan(const BInst* __restrict__ insts, int K,
                                                  int* __restrict__ off) {
    __shared__ int part[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < K; base += 1024) {
        const int k = base + static_cast<int>(threadIdx.x);
        int p = 0;
        if (k < K && evaluatedNow<T>(*insts[k].ctrl)) p = evalParts(insts[k]);
        part[threadIdx.x] = p;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int a = threadIdx.x >= static_cast<unsigned>(d) ? part[threadIdx.x - d] : 0;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (k < K) off[k] = carry + part[threadIdx.x] - p;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) off[K] = carry;
}

// Part p of P: columns and rows in P equal slices.
template <int GC, int GR, typename T, bool kStage>
__device__ inline void evalPart(const BInst& in, int p, int P, double* dst, int certEvery) {
    const long long cspan = (in.n + P - 1) / P, rspan = (in.m + P - 1) / P;
    const int jb = static_cast<int>(min(static_cast<long long>(in.n), p * cspan));
    const int je = static_cast<int>(min(static_cast<long long>(in.n), jb + cspan));
    const int ib = static_cast<int>(min(static_cast<long long>(in.m), p * rspan));
    const int ie = static_cast<int>(min(static_cast<long long>(in.m), ib + rspan));
    const BCtrl& c = *in.ctrl;
    bEvalFields<GC, GR, T, kStage>(in, pickVecs<T>(in), c, dst, jb, je, ib, ie,
                                   certDue(c, certEvery));
}

// Evaluations done in one part whose gathered vectors fit the LDS stage
// (bEvalFields' kStage): kEvalParts<T, true> takes these, kEvalParts<T,
// false> the rest.
__device__ inline bool evalStaged(const BInst& in) {
    return evalParts(in) == 1 && 2 * static_cast<long long>(in.m) <= kEvalStageElems &&
           in.n <= kEvalStageElems;
}

// Every part of this round's evaluations (of the kind kStage selects), over
// a fixed grid: part t belongs to the last instance i with off[i] <= t; its
// sums go to parts[(i * kEvalMaxParts + p) * kNumEval ...].
template <typename T, bool kStage>
__global__ __launch_bounds__(kBlock) void kEvalParts(const BInst* __restrict__ insts, int K,
                                                     const int* __restrict__ off,
                                                     double* __restrict__ parts,
                                                     int certEvery) {
    const int total = off[K];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = K - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (off[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const BInst& in = insts[lo];
        if (evalStaged(in) != kStage) continue;
        const int p = t - off[lo], P = evalParts(in);
        double* dst = parts + (static_cast<size_t>(lo) * kEvalMaxParts + p) * kNumEval;
        switch (in.preset) {
            case 0: evalPart<1, 1, T, kStage>(in, p, P, dst, certEvery); break;
            case 1: evalPart<2, 2, T, kStage>(in, p, P, dst, certEvery); break;
            case 2: evalPart<2, 4, T, kStage>(in, p, P, dst, certEvery); break;
            case 3: evalPart<4, 4, T, kStage>(in, p, P, dst, certEvery); break;
            case 4: evalPart<4, 8, T, kStage>(in, p, P, dst, certEvery); break;
            case 5: evalPart<8, 8, T, kStage>(in, p, P, dst, certEvery); break;
            case 6: evalPart<8, 16, T, kStage>(in, p, P, dst, certEvery); break;
            default: evalPart<16, 16, T, kStage>(in, p, P, dst, certEvery); break;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kEvalControl(const BInst* __restrict__ insts, int K,
                                                        BOptions o, const double* __restrict__ parts) {
    if (static_cast<int>(blockIdx.x) >= K) return;
    evalControl<T>(insts[blockIdx.x], o,
                   parts + static_cast<size_t>(blockIdx.x) * kEvalMaxParts * kNumEval);
}

// Calls f(instance, jb, je, ib, ie) for every part of this round's
// evaluations (offsets of kEvalScan<T>), over a fixed grid: the column and
// row slices of evalPart.
template <typename F>
__device__ inline void forEvalParts(const BInst* __restrict__ insts, int K,
                                    const int* __restrict__ off, F&& f) {
    const int total = off[K];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = K - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (off[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const BInst& in = insts[lo];
        const int p = t - off[lo], P = evalParts(in);
        const long long cspan = (in.n + P - 1) / P, rspan = (in.m + P - 1) / P;
        const int jb = static_cast<int>(min(static_cast<long long>(in.n), p * cspan));
        const int je = static_cast<int>(min(static_cast<long long>(in.n), jb + cspan));
        const int ib = static_cast<int>(min(static_cast<long long>(in.m), p * rspan));
        const int ie = static_cast<int>(min(static_cast<long long>(in.m), ib + rspan));
        f(in, jb, je, ib, ie);
    }
}

// FP64 evaluation points of the refining instances evaluated this round
// (offsets of kEvalScan<double>), slice by slice.
__global__ __launch_bounds__(kBlock) void kRefineMaterializeParts(const BInst* __restrict__ insts,
                                                                  int K,
                                                                  const int* __restrict__ off) {
    forEvalParts(insts, K, off, [&](const BInst& in, int jb, int je, int ib, int ie) {
        if (in.ctrl->phase == kPhaseRefine) refineMaterializeSlice(in, jb, je, ib, ie);
    });
}

// The copies of this round's decisions (kEvalControl<T>), slice by slice.
template <typename T>
__global__ __launch_bounds__(kBlock) void kControlCopy(const BInst* __restrict__ insts, int K,
                                                        const int* __restrict__ off) {
    forEvalParts(insts, K, off, [&](const BInst& in, int jb, int je, int ib, int ie) {
        const int code = in.ctrl->pending;
        if (code != kPendNone) pendingCopySlice<T>(in, code, jb, je, ib, ie);
    });
}

// The re-centrings of this round's decisions, once their FP64 bases are in
// place (kControlCopy), slice by slice.
__global__ __launch_bounds__(kBlock) void kControlRecenter(const BInst* __restrict__ insts, int K,
                                                           const int* __restrict__ off) {
    forEvalParts(insts, K, off, [&](const BInst& in, int jb, int je, int ib, int ie) {
        const int code = in.ctrl->pending;
        if (code == kPendToRefine || code == kPendRecenter)
            refineRecenterSlice(in, jb, je, ib, ie);
    });
}

// The evaluation and control of one round for the instances in precision T:
// part offsets, the parts, the per-instance control, then the vector work
// of its decisions in the same parts.  The FP64 evaluation also covers
// refining instances, at their materialized points.
template <typename T>
void enqueueEval(hipStream_t st, int grid, const BInst* di, int K, const BOptions& bo, int* off,
                 double* parts) {
    kEvalScan<T><<<1, 1024, 0, st>>>(di, K, off);
    if constexpr (std::is_same_v<T, double>) {
        if (bo.refine) kRefineMaterializeParts<<<grid, kBlock, 0, st>>>(di, K, off);
    }
    kEvalParts<T, true><<<grid, kBlock, 0, st>>>(di, K, off, parts, 10 * bo.N);
    kEvalParts<T, false><<<grid, kBlock, 0, st>>>(di, K, off, parts, 10 * bo.N);
    kEvalControl<T><<<K, kBlock, 0, st>>>(di, K, bo, parts);
    kControlCopy<T><<<grid, kBlock, 0, st>>>(di, K, off);
    if (bo.refine) kControlRecenter<<<grid, kBlock, 0, st>>>(di, K, off);
}

// Counts unfinished instances (one thread per instance, sum by atomics on an
// int counter is order-independent and exact).
__global__ void kCountActive(const BInst* __restrict__ insts, int count, int* __restrict__ out) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count && insts[t].ctrl->phase != kPhaseDone) atomicAdd(out, 1);
}

// Ends an unfinished instance at a time limit or cancellation: `status`
// (TIME_LIMIT / ITER_LIMIT), except that an instance in its polish phase
// has converged and ends OPTIMAL at its last evaluated point, as the
// reference's polish loop does when the clock or the stop flag breaks it
// (pdlp.cpp:1473-1478, 2545-2550 and the converged branch of 1677-1706).
__device__ inline void abortInstance(BCtrl& c, int status) {
    if (c.phase == kPhaseDone) return;
    c.status = c.phase == kPhasePolish ? MPX_STATUS_OPTIMAL : status;
    c.phase = kPhaseDone;
    c.round_iters = 0;
    stampDone(c);
}

// Ends every unfinished instance (time limit / cancellation of a batch).
__global__ void kAbortActive(const BInst* __restrict__ insts, int count, int status) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) abortInstance(*insts[t].ctrl, status);
}

// Per-instance initialisation: scaled column bounds (its own bounds when the
// instance overrides them), FP32 copies, zero iterates, control state.
struct BInitArgs {
    const double* lo_o;  // instance original column bounds
    const double* hi_o;
    double* l;           // instance scaled fp64 bounds (may alias the owner's)
    double* u;
    float* lf;
    float* uf;
    // The matrix's FP32 objective and row bounds: a refined instance of an
    // earlier run had its FP32 vectors pointed at its correction data.
    const float *cf, *rlf, *ruf;
    const double* col_scale;
    const double* scal;  // owner scal
    int n, m, own_bounds, mixed;
    // Warm start in original space, staged in the instance's FP64 x_r (n)
    // and y^ (m) arrays by the upload; consumed (and those arrays cleared)
    // by kInstInit.
    int warm;
    double offset, sgn;
};

__global__ __launch_bounds__(kBlock) void kInstInit(const BInst* __restrict__ insts,
                                                    const BInitArgs* __restrict__ args,
                                                    const BOptions o, double primal_weight) {
    const BInst in = insts[blockIdx.x];
    const BInitArgs a = args[blockIdx.x];
    const int tid = threadIdx.x;
    const double cs = a.scal[1], os = a.scal[0];
    for (int j = tid; j < a.n; j += kBlock) {
        if (a.own_bounds) {
            const double lj = a.lo_o[j], uj = a.hi_o[j];
            double l = isfinite(lj) ? lj / a.col_scale[j] : lj;
            double u = isfinite(uj) ? uj / a.col_scale[j] : uj;
            if (isfinite(l)) l *= cs;
            if (isfinite(u)) u *= cs;
            a.l[j] = l;
            a.u[j] = u;
        }
        a.lf[j] = static_cast<float>(a.l[j]);
        a.uf[j] = static_cast<float>(a.u[j]);
        // pdlp.cpp:915-930 / 1812-1827: x = x0 = warm_x / col_scale * cs.
        const double xs = a.warm ? in.d.xr[j] / a.col_scale[j] * cs : 0.0;
        in.d.x[j] = in.d.x0[j] = xs;
        in.d.xh[j] = in.d.xr[j] = 0.0;
        in.f.x[j] = in.f.x0[j] = static_cast<float>(xs);
        in.f.xh[j] = in.f.xr[j] = 0.0f;
    }
    for (int i = tid; i < a.m; i += kBlock) {
        // y = y0 = warm_y * os / (sign * row_scale).
        const double ys = a.warm ? in.d.yh[i] * os / (a.sgn * in.row_scale[i]) : 0.0;
        in.d.y[i] = in.d.y0[i] = ys;
        in.d.yh[i] = 0.0;
        in.f.y[i] = in.f.y0[i] = static_cast<float>(ys);
        in.f.yh[i] = 0.0f;
    }
    if (tid == 0) {
        BInst& w = const_cast<BInst&>(insts[blockIdx.x]);
        w.f.c = a.cf;
        w.f.rl = a.rlf;
        w.f.ru = a.ruf;
        w.f.q = nullptr;
        BCtrl& c = *in.ctrl;
        c.refined = 0;
        c.refine_iters = 0;
        c.step = 0.998 / sqrt(maxRef(a.scal[4], 1e-12));
        c.pw = primal_weight;
        c.fpe0 = -1.0;
        c.last_fpe = INFINITY;
        c.pw_err_sum = 0.0;
        c.pw_last_err = 0.0;
        c.best_pw = primal_weight;
        c.best_score = INFINITY;
        c.precond_pw = primal_weight;
        c.os = a.scal[0];
        c.cs = a.scal[1];
        c.cnorm = a.scal[2];
        c.bnorm = a.scal[3];
        c.sgn = a.sgn;
        c.offset = a.offset;
        c.best_kkt = INFINITY;
        c.rel_p = c.rel_d = c.rel_p_inf = c.rel_d_inf = c.gap = c.pobj = c.dobj = 0.0;
        c.inner = c.total = c.pcount = 0;
        c.phase = a.mixed ? kPhaseFp32 : kPhaseFp64;
        c.status = MPX_STATUS_ITER_LIMIT;
        c.restarts = c.fp32_iters = c.stall = 0;
        c.power_iters = static_cast<int>(a.scal[5]);
        c.skip_eval = 0;
        c.t_finish = 0ull;
        c.round_iters = o.max_iter < o.N ? o.max_iter : o.N;
        if (c.round_iters <= 0) {
            c.phase = kPhaseDone;
            c.round_iters = 0;
        }
    }
}

// Recovery into original space (pdlp.cpp:1677-1697) into per-instance
// device output buffers.
__global__ void kInstRecover(const BInst* __restrict__ insts, double* const* __restrict__ xo,
                             double* const* __restrict__ yo) {
    const BInst& in = insts[blockIdx.x];
    const BCtrl& c = *in.ctrl;
    const int tid = threadIdx.x;
    const double inv_cs = 1.0 / c.cs, inv_os = 1.0 / c.os;
    // Never left the plain FP32 phase: the answer is the FP32 iterate.
    const bool f32 = !c.refined && c.fp32_iters == c.total && c.total > 0;
    const bool clamp = c.status != MPX_STATUS_UNBOUNDED;
    double* x = xo[blockIdx.x];
    double* y = yo[blockIdx.x];
    for (int j = tid; j < in.n; j += kBlock) {
        const double h = f32 ? static_cast<double>(in.f.xh[j]) : in.d.xh[j];
        double xj = h * in.col_scale[j] * inv_cs;
        if (clamp) {
            if (isfinite(in.lo_o[j])) xj = maxRef(xj, in.lo_o[j]);
            if (isfinite(in.hi_o[j])) xj = minRef(xj, in.hi_o[j]);
        }
        x[j] = xj;
    }
    for (int i = tid; i < in.m; i += kBlock) {
        const double h = f32 ? static_cast<double>(in.f.yh[i]) : in.d.yh[i];
        y[i] = c.sgn * h * in.row_scale[i] * inv_os;
    }
}

// ---- host engine ----------------------------------------------------------------
void hostTranspose(int m, int n, const int* rs, const int* ci, std::vector<int>& cs,
                   std::vector<int>& ri, std::vector<int>& perm) {
    const int nnz = rs[m];
    cs.assign(static_cast<size_t>(n) + 1, 0);
    ri.resize(static_cast<size_t>(nnz));
    perm.resize(static_cast<size_t>(nnz));
    for (int k = 0; k < nnz; ++k) ++cs[ci[k] + 1];
    for (int j = 0; j < n; ++j) cs[j + 1] += cs[j];
    std::vector<int> next(cs.begin(), cs.end() - 1);
    for (int i = 0; i < m; ++i)
        for (int k = rs[i]; k < rs[i + 1]; ++k) {
            const int dst = next[ci[k]]++;
            ri[dst] = i;
            perm[dst] = k;
        }
}

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) {
        if (bytes) MPX_HIP(hipMalloc(&p, bytes));
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Part offsets and part sums of the per-round evaluation of K instances.
struct EvalBufs {
    std::unique_ptr<DevBuf> buf;
    int* off = nullptr;
    double* parts = nullptr;
    void alloc(int K) {
        const size_t offb = (static_cast<size_t>(K + 1) * sizeof(int) + 255) & ~size_t(255);
        const size_t partb = static_cast<size_t>(K) * kEvalMaxParts * kNumEval * sizeof(double);
        buf = std::make_unique<DevBuf>(offb + partb);
        off = static_cast<int*>(buf->p);
        parts = reinterpret_cast<double*>(static_cast<char*>(buf->p) + offb);
    }
};

// Bump allocator over one device allocation.
struct Carve {
    char* base;
    size_t used = 0, cap;
    template <typename T>
    T* take(size_t count) {
        const size_t b = (count * sizeof(T) + 255) & ~size_t(255);
        if (used + b > cap) {
            setError("internal: batch arena overflow");
            throw DeviceError();
        }
        T* p = reinterpret_cast<T*>(base + used);
        used += b;
        return p;
    }
};
inline size_t pad(size_t count, size_t elt) { return (count * elt + 255) & ~size_t(255); }

size_t ownerBytes(int m, int n, int nnz) {
    const size_t M = m, N = n, Z = nnz;
    return pad(M + 1, 4) + 3 * pad(Z, 4) + pad(N + 1, 4) + pad(Z, 4)  // rs ci ri perm cs
           + 3 * pad(Z, 8)                                          // val_o val valt
           + 6 * pad(N, 8) + 6 * pad(M, 8)                          // obj lo hi c l u cmax, rl ru rl_o ru_o rs sig ..
           + 4 * pad(N, 8) + 2 * pad(M, 8)                          // col_scale tau + power px0 px pxn, py
           + 2 * pad(Z, 4) + 4 * pad(N, 4) + 4 * pad(M, 4)          // fp32 copies
           + pad(8, 8) + 4096;
}

size_t instBytes(int m, int n, bool own_bounds) {
    const size_t M = m, N = n;
    size_t b = 4 * pad(N, 8) + 3 * pad(M, 8) + 4 * pad(N, 4) + 3 * pad(M, 4);  // iterates
    b += 2 * pad(N, 4);                                                     // fp32 bounds
    b += pad(N, 4) + 3 * pad(M, 4);                                         // refinement data
    if (own_bounds) b += 4 * pad(N, 8);                                     // lo_o hi_o l u
    b += pad(1, sizeof(BCtrl)) + 1024;
    return b;
}

// Host-side data of one distinct matrix that its device setup needs besides
// the LP: A' in CSR (= A in CSC) with the CSR position of every entry, and
// the power-iteration start vector.
struct OwnerHost {
    std::vector<int> cs, ri, perm;
    std::vector<double> start;
};

OwnerHost ownerHostData(const mpx_lp& lp) {
    OwnerHost h;
    hostTranspose(lp.num_rows, lp.num_cols, lp.row_starts, lp.col_indices, h.cs, h.ri, h.perm);
    // Power-iteration start vector, exactly as the reference (pdlp.cpp:756-764).
    std::mt19937 rng(42);
    std::normal_distribution<double> dist(0.0, 1.0);
    h.start.resize(static_cast<size_t>(lp.num_cols));
    for (double& v : h.start) v = dist(rng);
    double s = 0.0;
    for (double v : h.start) s += v * v;
    double xn = std::sqrt(s);
    if (xn < 1e-15) xn = 1.0;
    for (double& v : h.start) v /= xn;
    return h;
}

// Lane-group preset of a matrix: the caller's `packed_preset`, the one with
// `spmv_group` lanes per column when the caller forces a group size, the
// autotuned table entry for the matrix's shape, or the heuristic.  The
// choice depends on the matrix alone (not on timing or on the rest of a
// batch), so an instance's arithmetic is the same whatever it runs with.
int ownerPreset(const mpx_pdlp_options& o, int nnz, int m, int n) {
    if (o.packed_preset >= 0 && o.packed_preset < kNumPresets) return o.packed_preset;
    if (o.spmv_group <= 0) {
        const int tuned = tunedPreset(nnz, m, n);
        return tuned >= 0 ? tuned : heuristicPreset(nnz, m, n);
    }
    if (o.spmv_group == 1) return 0;
    int best = 0;
    for (int p = 0; p < kNumPresets; ++p)
        if (presetGC(p) <= o.spmv_group) best = p;
    return best;
}

// Carves the device arrays of one distinct matrix from `cv` and fills `b`
// (uploadOwner() then fills the arrays).
void carveOwner(const mpx_lp& lp, const mpx_pdlp_options& o, Carve& cv, BOwner& b) {
    const int m = lp.num_rows, n = lp.num_cols, nnz = lp.row_starts[m];
    std::memset(&b, 0, sizeof(b));
    b.m = m;
    b.n = n;
    b.nnz = nnz;
    b.do_ruiz = o.do_ruiz_scaling;
    b.ruiz_iters = o.ruiz_iterations;
    b.do_pc = o.do_pock_chambolle_scaling;
    b.do_rescale = o.do_bound_obj_rescaling;
    b.norm_inf = o.optimality_norm == 1;
    b.sense_max = lp.sense != 0;
    b.offset = lp.obj_offset;
    b.sv_max_iter = o.sv_max_iter;
    b.sv_tol = o.sv_tol;
    int* rs = cv.take<int>(m + 1);
    int* ci = cv.take<int>(nnz);
    int* ri = cv.take<int>(nnz);
    int* perm = cv.take<int>(nnz);
    int* cs = cv.take<int>(n + 1);
    double* val_o = cv.take<double>(nnz);
    b.val = cv.take<double>(nnz);
    b.valt = cv.take<double>(nnz);
    double* obj_o = cv.take<double>(n);
    double* lo_o = cv.take<double>(n);
    double* hi_o = cv.take<double>(n);
    b.c = cv.take<double>(n);
    b.l = cv.take<double>(n);
    b.u = cv.take<double>(n);
    b.cmax = cv.take<double>(n);
    b.col_scale = cv.take<double>(n);
    b.tau = cv.take<double>(n);
    b.px0 = cv.take<double>(n);
    b.px = cv.take<double>(n);
    b.pxn = cv.take<double>(n);
    double* rl_o = cv.take<double>(m);
    double* ru_o = cv.take<double>(m);
    b.rl = cv.take<double>(m);
    b.ru = cv.take<double>(m);
    b.row_scale = cv.take<double>(m);
    b.sig = cv.take<double>(m);
    b.py = cv.take<double>(m);
    b.valf = cv.take<float>(nnz);
    b.valtf = cv.take<float>(nnz);
    b.cf = cv.take<float>(n);
    b.tauf = cv.take<float>(n);
    b.sigf = cv.take<float>(m);
    b.rlf = cv.take<float>(m);
    b.ruf = cv.take<float>(m);
    b.scal = cv.take<double>(8);
    b.rs = rs;
    b.ci = ci;
    b.cs = cs;
    b.ri = ri;
    b.perm = perm;
    b.val_o = val_o;
    b.obj_o = obj_o;
    b.lo_o = lo_o;
    b.hi_o = hi_o;
    b.rl_o = rl_o;
    b.ru_o = ru_o;
}

// Queues the uploads of `lp` and `h` into the arrays of `b` on `st`; the
// host data must stay alive until the stream has consumed them.
void uploadOwner(const mpx_lp& lp, const OwnerHost& h, const BOwner& b, hipStream_t st) {
    const int m = lp.num_rows, n = lp.num_cols, nnz = lp.row_starts[m];
    int* rs = const_cast<int*>(b.rs);
    int* ci = const_cast<int*>(b.ci);
    int* ri = const_cast<int*>(b.ri);
    int* perm = const_cast<int*>(b.perm);
    int* cs = const_cast<int*>(b.cs);
    double* val_o = const_cast<double*>(b.val_o);
    double* obj_o = const_cast<double*>(b.obj_o);
    double* lo_o = const_cast<double*>(b.lo_o);
    double* hi_o = const_cast<double*>(b.hi_o);
    double* rl_o = const_cast<double*>(b.rl_o);
    double* ru_o = const_cast<double*>(b.ru_o);
    MPX_HIP(hipMemcpyAsync(rs, lp.row_starts, (m + 1) * 4ull, hipMemcpyHostToDevice, st));
    if (nnz) {
        MPX_HIP(hipMemcpyAsync(ci, lp.col_indices, nnz * 4ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(ri, h.ri.data(), nnz * 4ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(perm, h.perm.data(), nnz * 4ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(val_o, lp.values, nnz * 8ull, hipMemcpyHostToDevice, st));
    }
    MPX_HIP(hipMemcpyAsync(cs, h.cs.data(), (n + 1) * 4ull, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(obj_o, lp.obj, n * 8ull, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(lo_o, lp.col_lower, n * 8ull, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(hi_o, lp.col_upper, n * 8ull, hipMemcpyHostToDevice, st));
    if (m) {
        MPX_HIP(hipMemcpyAsync(rl_o, lp.row_lower, m * 8ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(ru_o, lp.row_upper, m * 8ull, hipMemcpyHostToDevice, st));
    }
    MPX_HIP(hipMemcpyAsync(b.px0, h.start.data(), n * 8ull, hipMemcpyHostToDevice, st));
}

// Carves the device arrays of one instance of matrix `w` (its own column
// bounds when lo/hi are given, queued for upload on `st`) and fills the
// instance descriptor and its initialisation arguments.
void carveInstance(const BOwner& w, const mpx_lp& lp, int owner_id, int preset, const double* lo,
                   const double* hi, bool mixed, Carve& cv, hipStream_t st, BCtrl* ctrl,
                   BInst& in, BInitArgs& a, const double* warm_x = nullptr,
                   const double* warm_y = nullptr) {
    const int m = w.m, n = w.n;
    std::memset(&in, 0, sizeof(in));
    std::memset(&a, 0, sizeof(a));
    in.m = m;
    in.n = n;
    in.nnz = w.nnz;
    in.sense_max = w.sense_max;
    in.owner = owner_id;
    in.preset = preset;
    in.rs = w.rs;
    in.ci = w.ci;
    in.cs = w.cs;
    in.ri = w.ri;
    in.rl_o = w.rl_o;
    in.ru_o = w.ru_o;
    in.obj_o = w.obj_o;
    in.val = w.val;
    in.valt = w.valt;
    in.c = w.c;
    in.rl = w.rl;
    in.ru = w.ru;
    in.row_scale = w.row_scale;
    in.col_scale = w.col_scale;
    in.sig = w.sig;
    in.tau = w.tau;
    a.n = n;
    a.m = m;
    a.col_scale = w.col_scale;
    a.scal = w.scal;
    a.mixed = mixed;
    a.offset = lp.obj_offset;
    a.sgn = lp.sense ? -1.0 : 1.0;
    if (lo) {
        double* lo_o = cv.take<double>(n);
        double* hi_o = cv.take<double>(n);
        MPX_HIP(hipMemcpyAsync(lo_o, lo, n * 8ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(hi_o, hi, n * 8ull, hipMemcpyHostToDevice, st));
        in.lo_o = lo_o;
        in.hi_o = hi_o;
        double* l = cv.take<double>(n);
        double* u = cv.take<double>(n);
        in.l = l;
        in.u = u;
        a.own_bounds = 1;
        a.lo_o = lo_o;
        a.hi_o = hi_o;
        a.l = l;
        a.u = u;
    } else {
        in.lo_o = w.lo_o;
        in.hi_o = w.hi_o;
        in.l = w.l;
        in.u = w.u;
        a.own_bounds = 0;
        a.l = w.l;
        a.u = w.u;
    }
    float* lf = cv.take<float>(n);
    float* uf = cv.take<float>(n);
    a.lf = lf;
    a.uf = uf;
    a.cf = w.cf;
    a.rlf = w.rlf;
    a.ruf = w.ruf;
    if (mixed) {
        in.rc = cv.take<float>(n);
        in.rrl = cv.take<float>(m);
        in.rru = cv.take<float>(m);
        in.rq = cv.take<float>(m);
    }
    auto fillVecs = [&](auto& v, auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        v.x = cv.take<T>(n);
        v.x0 = cv.take<T>(n);
        v.xr = cv.take<T>(n);
        v.xh = cv.take<T>(n);
        v.y = cv.take<T>(m);
        v.y0 = cv.take<T>(m);
        v.yh = cv.take<T>(m);
    };
    fillVecs(in.d, static_cast<double*>(nullptr));
    fillVecs(in.f, static_cast<float*>(nullptr));
    in.d.val = w.val;
    in.d.valt = w.valt;
    in.d.c = w.c;
    in.d.tau = w.tau;
    in.d.l = in.l;
    in.d.u = in.u;
    in.d.sig = w.sig;
    in.d.rl = w.rl;
    in.d.ru = w.ru;
    in.f.val = w.valf;
    in.f.valt = w.valtf;
    in.f.c = w.cf;
    in.f.tau = w.tauf;
    in.f.l = lf;
    in.f.u = uf;
    in.f.sig = w.sigf;
    in.f.rl = w.rlf;
    in.f.ru = w.ruf;
    in.ctrl = ctrl;
    if (warm_x && warm_y) {
        MPX_HIP(hipMemcpyAsync(in.d.xr, warm_x, n * 8ull, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(in.d.yh, warm_y, m * 8ull, hipMemcpyHostToDevice, st));
        a.warm = 1;
    }
}

// ---- lane-per-instance rounds of one shared matrix ----------------------------
// State of the interleaved kernels for up to `cap` instances of one matrix.
struct IlSet {
    IlBufs<float> f;
    IlBufs<double> d;
    // Mixed precision: the FP32 refinement set (iterates and bounds in r,
    // per-lane c' in rc and row data rrl, rru, rq).
    IlBufs<float> r;
    float *rc = nullptr, *rrl = nullptr, *rru = nullptr, *rq = nullptr;
    int *slots32 = nullptr, *slots64 = nullptr, *slotsR = nullptr, *counts = nullptr;
    int* next = nullptr;  // kNextCounts: the next round's counts, same order as counts
    ILane *lanes32 = nullptr, *lanes64 = nullptr, *lanesR = nullptr;
    // FP64 bound flags per (group, tile of 64 columns): bit j % 64 set when
    // some lane's l or u at column j differs from the owner's; the other
    // columns read the owner's bound (one cached value) instead of 64 lanes'
    // copies.
    unsigned long long* bf64 = nullptr;
    int owner = -1;  // instances taking part (-1: every instance of the table)

    static size_t boundTiles(int n) { return static_cast<size_t>((n + kLanes - 1) / kLanes); }
    static size_t bytes(int cap, int n, int m, bool mixed) {
        const size_t ng = static_cast<size_t>((cap + kLanes - 1) / kLanes);
        const size_t en = ng * kLanes * static_cast<size_t>(n);
        const size_t em = ng * kLanes * static_cast<size_t>(m);
        size_t b = 5 * pad(en, 8) + 2 * pad(em, 8);
        if (mixed) b += 2 * (5 * pad(en, 4) + 2 * pad(em, 4)) + pad(en, 4) + 3 * pad(em, 4);
        const size_t c = static_cast<size_t>(cap);
        return b + 3 * pad(c, 4) + 3 * pad(c, sizeof(ILane)) + 2 * pad(4, 4) +
               pad(ng * boundTiles(n), 8);
    }
    void carve(Carve& c, int cap, int n, int m, bool mixed) {
        const size_t ng = static_cast<size_t>((cap + kLanes - 1) / kLanes);
        const size_t N = static_cast<size_t>(n), M = static_cast<size_t>(m);
        auto take = [&](auto& b, auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            b.x = c.take<T>(ng * kLanes * N);
            b.x0 = c.take<T>(ng * kLanes * N);
            b.xr = c.take<T>(ng * kLanes * N);
            b.l = c.take<T>(ng * kLanes * N);
            b.u = c.take<T>(ng * kLanes * N);
            b.y = c.take<T>(ng * kLanes * M);
            b.y0 = c.take<T>(ng * kLanes * M);
        };
        take(d, static_cast<double*>(nullptr));
        if (mixed) {
            take(f, static_cast<float*>(nullptr));
            take(r, static_cast<float*>(nullptr));
            rc = c.take<float>(ng * kLanes * N);
            rrl = c.take<float>(ng * kLanes * M);
            rru = c.take<float>(ng * kLanes * M);
            rq = c.take<float>(ng * kLanes * M);
        }
        slots32 = c.take<int>(static_cast<size_t>(cap));
        slots64 = c.take<int>(static_cast<size_t>(cap));
        slotsR = c.take<int>(static_cast<size_t>(cap));
        lanes32 = c.take<ILane>(static_cast<size_t>(cap));
        lanes64 = c.take<ILane>(static_cast<size_t>(cap));
        lanesR = c.take<ILane>(static_cast<size_t>(cap));
        counts = c.take<int>(4);
        next = c.take<int>(4);
        bf64 = c.take<unsigned long long>(ng * boundTiles(n));
    }
};

template <typename F>
void withCW(int cw, F&& f) {
    switch (cw) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        default: f(std::integral_constant<int, 1>{}); break;
    }
}

// Grid of the lane-per-instance iteration kernels: twice that of the other
// kernels (16 workgroups per CU).  More workgroups than fit at once let the early
// finishers of the fixed work-item dealing take the rest; measured on C2
// (4,096 instances, m=3,000, n=8,000): FP32 dual 158 -> 129 us, FP32 primal
// 228 -> 211 us, FP64 dual 241 -> 219 us against 8 per CU.
inline int ilGrid(int grid) { return 2 * grid; }

// Kernel sets of a lane-per-instance round: plain FP32 phase, refinement
// (FP32 correction kernels), FP64 (FP64 phase and polish).  A round may
// leave out the set of a phase that no instance of the round is in.
constexpr int kSetPlain = 1, kSetRefine = 2, kSetFp64 = 4, kSetAll = 7;
inline int setsFor(bool f32) { return (f32 ? kSetPlain | kSetRefine : 0) | kSetFp64; }

// Slot lists of the coming round and per-instance -> interleaved copies.
void ilGather(hipStream_t st, int grid, const BInst* di, int K, const BOwner& w, IlSet& s,
              int sets) {
    const int n = w.n, m = w.m;
    kSlotLists<<<1, 1024, 0, st>>>(di, K, s.owner, s.slots32, s.slots64, s.slotsR, s.lanes32,
                                   s.lanes64, s.lanesR, s.counts);
    if (sets & kSetPlain) {
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slots32, s.counts, n, 0, s.f.x, s.f.x0,
                                                    s.f.l, s.f.u, nullptr, nullptr, nullptr);
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slots32, s.counts, m, 1, s.f.y, s.f.y0,
                                                    nullptr, nullptr, nullptr, nullptr, nullptr);
    }
    if ((sets & kSetRefine) && s.rc) {
        const int* cr = s.counts + 2;
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, cr, n, 0, s.r.x, s.r.x0, s.r.l,
                                                    s.r.u, nullptr, nullptr, nullptr);
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, cr, m, 1, s.r.y, s.r.y0,
                                                    nullptr, nullptr, nullptr, nullptr, nullptr);
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, cr, n, 2, s.rc, nullptr,
                                                    nullptr, nullptr, nullptr, nullptr, nullptr);
        kInterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, cr, m, 3, s.rrl, s.rru, s.rq,
                                                    nullptr, nullptr, nullptr, nullptr);
    }
    if (sets & kSetFp64) {
        kInterleave<double><<<grid, kBlock, 0, st>>>(di, s.slots64, s.counts + 1, n, 0, s.d.x,
                                                     s.d.x0, s.d.l, s.d.u, w.l, w.u, s.bf64);
        kInterleave<double><<<grid, kBlock, 0, st>>>(di, s.slots64, s.counts + 1, m, 1, s.d.y,
                                                     s.d.y0, nullptr, nullptr, nullptr, nullptr,
                                                     nullptr);
    }
}

// Columns (rows) per wave of the FP64 kernels of a mixed-precision batch.
// The FP32 pair is tuned in the first rounds, when every instance is still
// in FP32; one column (row) per wave is the fastest FP64 choice on C2 at
// full occupancy (primal 372 / 400 / 408 us for 1 / 2 / 4 columns, dual
// 220 / 250 / 293 us) and gives the most work items once few instances
// are left.
constexpr int kIlFp64PerWave = 1;

// Iteration k (primal then dual; FP32 instances first) with `cwp` columns
// per wave in the primal kernels and `cwd` rows per wave in the dual ones
// (`cwp64` / `cwd64` for the FP64 kernels), for the kernel sets in `sets`.
void ilIterate(hipStream_t st, int grid_base, const BInst* di, const BOwner& w, IlSet& s,
               int sets, int cwp, int cwd, int cwp64, int cwd64, int k) {
    const int n = w.n, m = w.m;
    const int grid = ilGrid(grid_base);
    auto primal = [&](bool fp32) {
        withCW(fp32 ? cwp : cwp64, [&](auto CWI) {
            constexpr int CW = decltype(CWI)::value;
            if (fp32)
                kILPrimal<float, CW><<<grid, kBlock, 0, st>>>(
                    di, s.slots32, s.counts, s.lanes32, n, m, w.cs, w.ri, w.valtf, w.cf, w.tauf,
                    s.f.y, s.f.x, s.f.x0, s.f.xr, s.f.l, s.f.u, nullptr, nullptr, nullptr, k);
            else
                kILPrimal<double, CW><<<grid, kBlock, 0, st>>>(
                    di, s.slots64, s.counts + 1, s.lanes64, n, m, w.cs, w.ri, w.valt, w.c, w.tau,
                    s.d.y, s.d.x, s.d.x0, s.d.xr, s.d.l, s.d.u, w.l, w.u, s.bf64, k);
        });
    };
    auto dual = [&](bool fp32) {
        withCW(fp32 ? cwd : cwd64, [&](auto CWI) {
            constexpr int CW = decltype(CWI)::value;
            if (fp32)
                kILDual<float, CW><<<grid, kBlock, 0, st>>>(di, s.slots32, s.counts, s.lanes32, n,
                                                            m, w.rs, w.ci, w.valf, w.sigf, w.rlf,
                                                            w.ruf, s.f.xr, s.f.y, s.f.y0, k);
            else
                kILDual<double, CW><<<grid, kBlock, 0, st>>>(
                    di, s.slots64, s.counts + 1, s.lanes64, n, m, w.rs, w.ci, w.val, w.sig, w.rl,
                    w.ru, s.d.xr, s.d.y, s.d.y0, k);
        });
    };
    if (sets & kSetPlain) {
        primal(true);
        dual(true);
    }
    if ((sets & kSetRefine) && s.rc) {
        // The refinement's correction phase: one column / row per wave.
        kILPrimal<float, 1, true><<<grid, kBlock, 0, st>>>(
            di, s.slotsR, s.counts + 2, s.lanesR, n, m, w.cs, w.ri, w.valtf, nullptr, w.tauf,
            s.r.y, s.r.x, s.r.x0, s.r.xr, s.r.l, s.r.u, nullptr, nullptr, nullptr, k, s.rc);
        kILDual<float, 1, true><<<grid, kBlock, 0, st>>>(
            di, s.slotsR, s.counts + 2, s.lanesR, n, m, w.rs, w.ci, w.valf, w.sigf, nullptr,
            nullptr, s.r.xr, s.r.y, s.r.y0, k, s.rrl, s.rru, s.rq);
    }
    if (sets & kSetFp64) {
        primal(false);
        dual(false);
    }
}

// Interleaved -> per-instance x and y after the round's last iteration.
void ilScatter(hipStream_t st, int grid, const BInst* di, const BOwner& w, IlSet& s, int sets) {
    const int n = w.n, m = w.m;
    if (sets & kSetPlain) {
        kDeinterleave<float><<<grid, kBlock, 0, st>>>(di, s.slots32, s.counts, n, 0, s.f.x);
        kDeinterleave<float><<<grid, kBlock, 0, st>>>(di, s.slots32, s.counts, m, 1, s.f.y);
    }
    if ((sets & kSetRefine) && s.rc) {
        kDeinterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, s.counts + 2, n, 0, s.r.x);
        kDeinterleave<float><<<grid, kBlock, 0, st>>>(di, s.slotsR, s.counts + 2, m, 1, s.r.y);
    }
    if (sets & kSetFp64) {
        kDeinterleave<double><<<grid, kBlock, 0, st>>>(di, s.slots64, s.counts + 1, n, 0, s.d.x);
        kDeinterleave<double><<<grid, kBlock, 0, st>>>(di, s.slots64, s.counts + 1, m, 1, s.d.y);
    }
}

// Iteration presets the online autotuner tries: (lanes per column, lanes per
// row) of the packed iteration kernels.
struct IterPreset {
    int gc, gr;
};

// Device copy of the options the per-instance control logic reads.
BOptions makeBOptions(const mpx_pdlp_options& o) {
    BOptions b{};
    b.max_iter = o.max_iter;
    b.N = o.termination_eval_frequency;
    b.polish_max_iter = o.polish_max_iter;
    b.norm_inf = o.optimality_norm == 1;
    b.update_pw = o.update_primal_weight;
    b.enable_polish = o.enable_polishing;
    b.precond_refresh = o.preconditioner_refresh;
    b.do_pc = o.do_pock_chambolle_scaling;
    b.primal_tol = o.primal_tol;
    b.dual_tol = o.dual_tol;
    b.opt_tol = o.optimality_tol;
    b.polish_p_tol = o.polish_primal_tol;
    b.polish_d_tol = o.polish_dual_tol;
    b.pid_kp = o.pid_kp;
    b.pid_ki = o.pid_ki;
    b.pid_kd = o.pid_kd;
    b.pid_i_smooth = o.pid_i_smooth;
    b.rs_suff = o.restart_sufficient_decay;
    b.rs_nec = o.restart_necessary_decay;
    b.rs_art = o.restart_artificial_fraction;
    b.precond_ratio = o.preconditioner_refresh_ratio;
    b.mixed = o.precision == MPX_PREC_MIXED;
    // FP32 refinement phase after the plain FP32 phase (else FP64 iterations).
    b.refine = b.mixed && o.mixed_refine ? 1 : 0;
    b.handover = o.mixed_handover > 0.0 ? o.mixed_handover : kFp32ExitRatio;
    // Tolerance of the refinement (or FP64) phase of mixed precision:
    // refine_tol, or the solve's own tolerances.
    b.refine_p_tol = o.refine_tol > 0.0 ? std::min(o.refine_tol, o.primal_tol) : o.primal_tol;
    b.refine_d_tol = o.refine_tol > 0.0 ? std::min(o.refine_tol, o.dual_tol) : o.dual_tol;
    b.refine_g_tol =
        o.refine_tol > 0.0 ? std::min(o.refine_tol, o.optimality_tol) : o.optimality_tol;
    return b;
}

class BatchEngine {
public:
    explicit BatchEngine(const mpx_pdlp_options& o) : o_(o) {
        MPX_HIP(hipSetDevice(o.device));
        MPX_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        MPX_HIP(hipEventCreate(&ev0_));
        MPX_HIP(hipEventCreate(&ev1_));
        bo_ = makeBOptions(o);
        mixed_ = o.precision == MPX_PREC_MIXED;
        graph_failed_ = graphsDisabled();
    }
    ~BatchEngine() {
        for (auto& kv : graphs_) kv.second.reset();
        if (ev0_) (void)hipEventDestroy(ev0_);
        if (ev1_) (void)hipEventDestroy(ev1_);
        if (st_) (void)hipStreamDestroy(st_);
    }

    // Registers a distinct LP (matrix owner).  Returns its id.
    int addOwner(const mpx_lp& lp) {
        owners_host_.push_back(&lp);
        return static_cast<int>(owners_host_.size()) - 1;
    }
    // Registers an instance of owner `ow` (col bounds override or null).
    // lo/hi: the instance's column bounds (NULL: the matrix's own); wx/wy:
    // an original-space warm start (NULL: cold), read at setup().
    int addInstance(int ow, const double* lo, const double* hi, const double* wx = nullptr,
                    const double* wy = nullptr) {
        inst_owner_.push_back(ow);
        inst_lo_.push_back(lo);
        inst_hi_.push_back(hi);
        inst_wx_.push_back(wx && wy ? wx : nullptr);
        inst_wy_.push_back(wx && wy ? wy : nullptr);
        return static_cast<int>(inst_owner_.size()) - 1;
    }

    // Upload (host -> device) followed by prepare().
    void setup();
    // Device-side setup on uploaded data (re-runnable).
    void prepare();
    // One instrumented round: every launch of the iteration kernels timed
    // with HIP events on the engine's stream.  Returns averages in us.
    // refine (mixed precision): every instance is first moved into the
    // refinement phase and the correction kernels are timed instead.
    void profileRound(int launches, double* primal_us, double* dual_us, int* slots,
                      bool refine = false);
    // Per-instance completion times (ms since the end of prepare()).
    std::vector<double> finishTimesMs();
    int rounds() const { return rounds_; }
    double uploadSeconds() const { return upload_seconds_; }
    bool interleaved() const { return interleaved_; }
    // Runs rounds until every instance is done (or the time limit / stop).
    void run(std::chrono::steady_clock::time_point t0);
    void results(double* const* x_out, double* const* y_out, mpx_pdlp_result* res);
    int count() const { return static_cast<int>(inst_owner_.size()); }
    double setupSeconds() const { return setup_seconds_; }
    // Shared-matrix batches may use the lane-per-instance kernels.
    void allowInterleaved(bool on) { allow_interleaved_ = on; }
    // Copies owner `w`'s scaled problem out (after setup()).
    void downloadOwner(int w, mpx_scaled_problem* out);
    // Runs `rounds` rounds; returns the number of unfinished instances.
    int step(int rounds);
    // Statistics of the iteration kernels (HIP events around every round's
    // iteration part): total milliseconds, rounds timed, instance-iterations.
    double iterMs() const { return iter_ms_; }
    int roundsTimed() const { return rounds_timed_; }
    IterPreset preset() const { return cur_; }
    int tunedColumnsPerWave() const { return interleaved_ ? tuned_.gc : 0; }
    int tunedRowsPerWave() const { return interleaved_ ? tuned_.gr : 0; }

private:
    struct G {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        void reset() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
            e = nullptr;
            g = nullptr;
        }
    };
    void enqueueRound(const IterPreset& p, bool timed);
    void launchRound(const IterPreset& p, bool timed);
    std::vector<IterPreset> candidates() const;

    mpx_pdlp_options o_;
    BOptions bo_{};
    bool mixed_ = false;
    hipStream_t st_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    std::vector<const mpx_lp*> owners_host_;
    std::vector<int> inst_owner_;
    std::vector<const double*> inst_lo_, inst_hi_, inst_wx_, inst_wy_;
    std::unique_ptr<DevBuf> arena_;
    std::vector<BOwner> owners_;
    std::vector<BInst> insts_;
    std::vector<int> owner_preset_;
    std::unique_ptr<DevBuf> d_insts_buf_, d_args_buf_, d_count_buf_;
    std::unique_ptr<SetupCohort> cohort_;  // device setup of every owner
    std::unique_ptr<DevBuf> d_cohort_buf_;
    std::unique_ptr<DevBuf> d_tiles_buf_, d_cnt_buf_, d_recs_buf_;
    PRec<float>* rec32_ = nullptr;
    PRec<double>* rec64_ = nullptr;
    Tile *tc32_ = nullptr, *tr32_ = nullptr, *tc64_ = nullptr, *tr64_ = nullptr;
    int* totals_ = nullptr;  // [0..3] tile counts, [4] unfinished instances
    size_t max_col_tiles_ = 0, max_row_tiles_ = 0;
    int grid_ = 2048;
    std::map<std::pair<int, int>, G> graphs_;
    bool graph_failed_ = false;
    IterPreset cur_{4, 8};
    std::vector<IterPreset> tune_left_;
    std::vector<std::pair<double, IterPreset>> tune_res_;
    double iter_ms_ = 0.0;
    int rounds_timed_ = 0;
    double setup_seconds_ = 0.0;
    double upload_seconds_ = 0.0;
    int rounds_ = 0;
    unsigned long long* stamp_ = nullptr;  // device wall clock at the end of prepare()
    BCtrl* ctrl_all_ = nullptr;
    bool allow_interleaved_ = false;
    // Lane-per-instance kernels: columns per wave (primal) and rows per wave
    // (dual) in use, and the autotuned pair ({0, 0} before tuning).
    int cwp_ = 1, cwd_ = 1;
    IterPreset tuned_{0, 0};
    bool fp32_gone_ = false;  // no instance left in its FP32 phase
    // Lane-per-instance mode (every instance shares owner 0).
    bool interleaved_ = false;
    std::unique_ptr<DevBuf> d_il_buf_;
    EvalBufs ev_;
    IlSet il_;
    // Lane-per-instance, mixed precision: the per-phase instance counts of
    // the coming round (plain FP32, FP64, refinement), known after a round.
    bool next_known_ = false;
    int next_[3] = {1, 1, 1};
    void enqueueInterleavedRound(bool timed);
    int roundSets() const;
    void readNextCounts();
};

void BatchEngine::setup() {
    const auto t0 = std::chrono::steady_clock::now();
    const int no = static_cast<int>(owners_host_.size());
    const int ni = count();
    size_t bytes = 0;
    for (int w = 0; w < no; ++w) {
        const mpx_lp& lp = *owners_host_[w];
        bytes += ownerBytes(lp.num_rows, lp.num_cols, lp.row_starts[lp.num_rows]);
    }
    for (int i = 0; i < ni; ++i) {
        const mpx_lp& lp = *owners_host_[inst_owner_[i]];
        bytes += instBytes(lp.num_rows, lp.num_cols, inst_lo_[i] != nullptr);
    }
    arena_ = std::make_unique<DevBuf>(bytes);
    ev_.alloc(std::max(1, ni));
    Carve cv{static_cast<char*>(arena_->p), 0, bytes};
    owners_.resize(static_cast<size_t>(no));
    owner_preset_.resize(static_cast<size_t>(no));
    std::vector<OwnerHost> hosts(static_cast<size_t>(no));  // alive until the stream syncs
    for (int w = 0; w < no; ++w) {
        const mpx_lp& lp = *owners_host_[w];
        hosts[w] = ownerHostData(lp);
        carveOwner(lp, o_, cv, owners_[w]);
        uploadOwner(lp, hosts[w], owners_[w], st_);
        owner_preset_[w] = ownerPreset(o_, owners_[w].nnz, owners_[w].m, owners_[w].n);
    }
    insts_.resize(static_cast<size_t>(ni));
    ctrl_all_ = cv.take<BCtrl>(static_cast<size_t>(std::max(1, ni)));
    std::vector<BInitArgs> args(static_cast<size_t>(ni));
    for (int i = 0; i < ni; ++i) {
        const int ow = inst_owner_[i];
        carveInstance(owners_[ow], *owners_host_[ow], ow, owner_preset_[ow], inst_lo_[i],
                      inst_hi_[i], mixed_, cv, st_, ctrl_all_ + i, insts_[i], args[i], inst_wx_[i],
                      inst_wy_[i]);
    }
    // Tile lists for the packed kernels, sized for the smallest tiles (16
    // columns or rows per tile at 16 lanes each).
    max_col_tiles_ = 0;
    max_row_tiles_ = 0;
    for (int i = 0; i < ni; ++i) {
        max_col_tiles_ += static_cast<size_t>((insts_[i].n + 15) / 16);
        max_row_tiles_ += static_cast<size_t>((insts_[i].m + 15) / 16);
    }
    {
        const size_t tb = 2 * (max_col_tiles_ + max_row_tiles_) * sizeof(Tile) + 16 * sizeof(int);
        d_tiles_buf_ = std::make_unique<DevBuf>(tb);
        Tile* t = static_cast<Tile*>(d_tiles_buf_->p);
        tc32_ = t;
        tr32_ = tc32_ + max_col_tiles_;
        tc64_ = tr32_ + max_row_tiles_;
        tr64_ = tc64_ + max_col_tiles_;
        totals_ = reinterpret_cast<int*>(tr64_ + max_row_tiles_);
        d_cnt_buf_ = std::make_unique<DevBuf>(sizeof(int) * 4 * static_cast<size_t>(std::max(1, ni)));
        const size_t nr = static_cast<size_t>(std::max(1, ni));
        d_recs_buf_ = std::make_unique<DevBuf>(pad(nr, sizeof(PRec<float>)) + pad(nr, sizeof(PRec<double>)));
        rec32_ = static_cast<PRec<float>*>(d_recs_buf_->p);
        rec64_ = reinterpret_cast<PRec<double>*>(static_cast<char*>(d_recs_buf_->p) +
                                                 pad(nr, sizeof(PRec<float>)));
    }
    // Lane-per-instance mode when every instance shares one matrix.
    interleaved_ = allow_interleaved_ && no == 1 && ni >= 1 && ilFits(owners_[0].n, owners_[0].m);
    if (interleaved_) {
        const size_t bytes_il = IlSet::bytes(ni, owners_[0].n, owners_[0].m, mixed_);
        d_il_buf_ = std::make_unique<DevBuf>(bytes_il);
        Carve c{static_cast<char*>(d_il_buf_->p), 0, bytes_il};
        il_.carve(c, ni, owners_[0].n, owners_[0].m, mixed_);
        il_.owner = -1;
    }
    int dev = 0, cus = 256;
    MPX_HIP(hipGetDevice(&dev));
    MPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid_ = std::max(kNumXcd, (cus * 8 / kNumXcd) * kNumXcd);
    // Online autotuning of the lane-per-instance kernels: columns per wave.
    if (interleaved_) {
        if (tuned_.gc > 0) {
            cur_ = tuned_;
        } else {
            // (columns per wave, rows per wave); tried from the back, each
            // candidate twice in two interleaved passes, the fastest single
            // round winning: the first round of a process runs cold (code
            // objects loaded, caches and translations empty), and timed once
            // it lost the one-column pair (kernels 175 / 101 us) to four
            // columns per wave (187 / 101 us).
            tune_left_ = {{4, 4}, {2, 2}, {2, 1}, {4, 1}, {1, 1},
                          {4, 4}, {2, 2}, {2, 1}, {4, 1}, {1, 1}};
            cur_ = {1, 1};
        }
    } else {
        cur_ = {0, 0};
    }
    cohort_ = std::make_unique<SetupCohort>(owners_, owner_preset_);
    d_cohort_buf_ = std::make_unique<DevBuf>(cohort_->bytes());
    cohort_->carve(static_cast<char*>(d_cohort_buf_->p));
    d_insts_buf_ = std::make_unique<DevBuf>(sizeof(BInst) * std::max(1, ni));
    d_args_buf_ = std::make_unique<DevBuf>(sizeof(BInitArgs) * std::max(1, ni));
    d_count_buf_ = std::make_unique<DevBuf>(sizeof(int) * 4 + 64);
    stamp_ = reinterpret_cast<unsigned long long*>(static_cast<char*>(d_count_buf_->p) + 32);
    MPX_HIP(hipMemcpyAsync(d_insts_buf_->p, insts_.data(), sizeof(BInst) * ni,
                           hipMemcpyHostToDevice, st_));
    MPX_HIP(hipMemcpyAsync(d_args_buf_->p, args.data(), sizeof(BInitArgs) * ni,
                           hipMemcpyHostToDevice, st_));
    MPX_HIP(hipStreamSynchronize(st_));
    upload_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    prepare();
}

// Device-side setup, re-runnable on uploaded data: scaling and power
// iteration per distinct matrix, then per-instance initialisation.
void BatchEngine::prepare() {
    const auto t0 = std::chrono::steady_clock::now();
    const int no = static_cast<int>(owners_host_.size());
    const int ni = count();
    for (auto& kv : graphs_) kv.second.reset();
    graphs_.clear();
    rounds_ = 0;
    fp32_gone_ = false;
    next_known_ = false;
    if (no) cohort_->run(st_);
    if (ni)
        kInstInit<<<ni, kBlock, 0, st_>>>(static_cast<const BInst*>(d_insts_buf_->p),
                                          static_cast<const BInitArgs*>(d_args_buf_->p), bo_,
                                          o_.primal_weight);
    kStampStart<<<1, 1, 0, st_>>>(stamp_);
    MPX_HIP(hipGetLastError());
    MPX_HIP(hipStreamSynchronize(st_));
    setup_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void BatchEngine::downloadOwner(int wi, mpx_scaled_problem* out) {
    const BOwner& w = owners_[wi];
    const size_t m = static_cast<size_t>(w.m), n = static_cast<size_t>(w.n),
                 z = static_cast<size_t>(w.nnz);
    auto get = [&](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) MPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st_));
    };
    get(out->values, w.val, z * 8);
    get(out->at_values, w.valt, z * 8);
    get(out->at_row_starts, w.cs, (n + 1) * 4);
    get(out->at_col_indices, w.ri, z * 4);
    get(out->cscaled, w.c, n * 8);
    get(out->col_lower, w.l, n * 8);
    get(out->col_upper, w.u, n * 8);
    get(out->row_lower, w.rl, m * 8);
    get(out->row_upper, w.ru, m * 8);
    get(out->row_scale, w.row_scale, m * 8);
    get(out->col_scale, w.col_scale, n * 8);
    get(out->sigma_base, w.sig, m * 8);
    get(out->tau_base, w.tau, n * 8);
    double scal[8];
    MPX_HIP(hipMemcpyAsync(scal, w.scal, sizeof(scal), hipMemcpyDeviceToHost, st_));
    MPX_HIP(hipStreamSynchronize(st_));
    out->objective_scale = scal[0];
    out->constraint_scale = scal[1];
    out->spectral_norm = std::sqrt(std::max(scal[4], 1e-12));
    out->power_iterations = static_cast<int32_t>(scal[5]);
}

std::vector<IterPreset> BatchEngine::candidates() const {
    // Around the batch's mean entries per column / row: about two entries per
    // lane, one size smaller and one larger.
    long long nnz = 0, rows = 0, cols = 0;
    for (const BInst& in : insts_) {
        nnz += in.nnz;
        rows += in.m;
        cols += in.n;
    }
    const double ac = cols ? static_cast<double>(nnz) / cols : 1.0;
    const double ar = rows ? static_cast<double>(nnz) / rows : 1.0;
    auto base = [](double a) {
        int g = 1;
        while (g < 16 && 2.0 * g < a) g <<= 1;
        return g;
    };
    const int bc = base(ac), br = base(ar);
    std::vector<IterPreset> out;
    out.push_back({bc, br});
    if (bc > 1) out.push_back({bc / 2, br});
    if (br > 1) out.push_back({bc, br / 2});
    if (bc < 16) out.push_back({bc * 2, br});
    if (br < 16) out.push_back({bc, br * 2});
    return out;
}

template <typename F>
void withG16(int g, F&& f) {
    switch (g) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: f(std::integral_constant<int, 16>{}); break;
    }
}

// One round: tile lists for the coming iterations, N iterations of the
// packed kernels (FP32 and/or FP64), then evaluation and control.
void BatchEngine::enqueueRound(const IterPreset& p, bool timed) {
    (void)p;
    const int K = count();
    const BInst* di = static_cast<const BInst*>(d_insts_buf_->p);
    int* cnt = static_cast<int*>(d_cnt_buf_->p);
    const int gk = ceilDiv(K, kBlock);
    kTileCounts<<<gk, kBlock, 0, st_>>>(di, K, cnt);
    kScan4<<<1, 1024, 0, st_>>>(cnt, K, di, totals_);
    kTileFill<<<gk, kBlock, 0, st_>>>(di, K, cnt, tc32_, tr32_, tc64_, tr64_, rec32_, rec64_);
    if (timed) MPX_HIP(hipEventRecord(ev0_, st_));
    const bool f32 = mixed_ && !fp32_gone_;
    for (int k = 1; k <= bo_.N; ++k) {
        if (f32) {
            kPackedPrimal<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tc32_, totals_ + 0, k);
            kPackedDual<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tr32_, totals_ + 1, k);
        }
        kPackedPrimal<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tc64_, totals_ + 2, k);
        kPackedDual<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tr64_, totals_ + 3, k);
    }
    if (timed) MPX_HIP(hipEventRecord(ev1_, st_));
    if (mixed_) enqueueEval<float>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
    enqueueEval<double>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
}

void BatchEngine::enqueueInterleavedRound(bool timed) {
    const int K = count();
    const BInst* di = static_cast<const BInst*>(d_insts_buf_->p);
    const BOwner& w = owners_[0];
    // FP32 kernels only while some instance is still in its FP32 phase
    // (instances never return to FP32, so this is decided on the host).
    const bool f32 = mixed_ && !fp32_gone_;
    const int sets = roundSets();
    ilGather(st_, grid_, di, K, w, il_, sets);
    if (timed) MPX_HIP(hipEventRecord(ev0_, st_));
    // FP64 kernels: the tuned pair when the batch is FP64 only.
    const int cwp64 = mixed_ ? kIlFp64PerWave : cwp_, cwd64 = mixed_ ? kIlFp64PerWave : cwd_;
    for (int k = 1; k <= bo_.N; ++k)
        ilIterate(st_, grid_, di, w, il_, sets, cwp_, cwd_, cwp64, cwd64, k);
    if (timed) MPX_HIP(hipEventRecord(ev1_, st_));
    ilScatter(st_, grid_, di, w, il_, sets);
    if (f32) enqueueEval<float>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
    enqueueEval<double>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
    if (mixed_) kNextCounts<<<1, 1024, 0, st_>>>(di, K, il_.owner, il_.next);
}

// The kernel sets of the coming lane-per-instance round: every set until
// the device has counted the phases once (kNextCounts, read back after each
// round of a mixed-precision batch), then only those with instances in them.
int BatchEngine::roundSets() const {
    if (!mixed_) return kSetFp64;
    int sets = kSetAll;
    if (next_known_)
        sets = (next_[0] > 0 ? kSetPlain : 0) | (next_[2] > 0 ? kSetRefine : 0) |
               (next_[1] > 0 ? kSetFp64 : 0);
    if (fp32_gone_) sets &= kSetFp64;
    return sets;
}

// Every active instance moved into the refinement phase around its current
// FP32 point (used to time the correction kernels at full occupancy).
__global__ __launch_bounds__(kBlock) void kForceRefine(const BInst* __restrict__ insts, int K) {
    if (static_cast<int>(blockIdx.x) >= K) return;
    const BInst& in = insts[blockIdx.x];
    BCtrl& c = *in.ctrl;
    __shared__ int go;
    if (threadIdx.x == 0) go = c.phase == kPhaseFp32 && c.round_iters > 0;
    __syncthreads();
    if (!go) return;
    for (int j = threadIdx.x; j < in.n; j += kBlock) in.d.x[j] = static_cast<double>(in.f.xh[j]);
    for (int i = threadIdx.x; i < in.m; i += kBlock) in.d.y[i] = static_cast<double>(in.f.yh[i]);
    __syncthreads();
    refineRecenter(in);
    if (threadIdx.x == 0) {
        c.phase = kPhaseRefine;
        c.refined = 1;
    }
}

void BatchEngine::profileRound(int launches, double* primal_us, double* dual_us, int* slots,
                               bool refine) {
    const int K = count();
    const BInst* di = static_cast<const BInst*>(d_insts_buf_->p);
    launches = std::max(1, std::min(launches, bo_.N));
    hipEvent_t e[4];
    for (auto& ev : e) MPX_HIP(hipEventCreate(&ev));
    double tp = 0.0, td = 0.0;
    int active = 0;
    const bool f32 = mixed_;  // the FP32 kernels carry the work in mixed mode
    refine = refine && mixed_;
    if (refine) kForceRefine<<<K, kBlock, 0, st_>>>(di, K);
    if (interleaved_ && refine) {
        const BOwner& w = owners_[0];
        ilGather(st_, grid_, di, K, w, il_, kSetAll);
        for (int k = 1; k <= launches; ++k) {
            MPX_HIP(hipEventRecord(e[0], st_));
            kILPrimal<float, 1, true><<<ilGrid(grid_), kBlock, 0, st_>>>(
                di, il_.slotsR, il_.counts + 2, il_.lanesR, w.n, w.m, w.cs, w.ri, w.valtf, nullptr,
                w.tauf, il_.r.y, il_.r.x, il_.r.x0, il_.r.xr, il_.r.l, il_.r.u, nullptr, nullptr,
                nullptr, k, il_.rc);
            MPX_HIP(hipEventRecord(e[1], st_));
            kILDual<float, 1, true><<<ilGrid(grid_), kBlock, 0, st_>>>(
                di, il_.slotsR, il_.counts + 2, il_.lanesR, w.n, w.m, w.rs, w.ci, w.valf, w.sigf,
                nullptr, nullptr, il_.r.xr, il_.r.y, il_.r.y0, k, il_.rrl, il_.rru, il_.rq);
            MPX_HIP(hipEventRecord(e[2], st_));
            MPX_HIP(hipEventSynchronize(e[2]));
            float a = 0.0f, b = 0.0f;
            MPX_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            MPX_HIP(hipEventElapsedTime(&b, e[1], e[2]));
            tp += a;
            td += b;
        }
        int c3[3];
        MPX_HIP(hipMemcpy(c3, il_.counts, sizeof(c3), hipMemcpyDeviceToHost));
        active = c3[2];
    } else if (interleaved_) {
        const BOwner& w = owners_[0];
        // Only the precision that carries the work is gathered and timed.
        ilGather(st_, grid_, di, K, w, il_, setsFor(f32));
        for (int k = 1; k <= launches; ++k) {
            MPX_HIP(hipEventRecord(e[0], st_));
            withCW(cwp_, [&](auto CWI) {
                constexpr int CW = decltype(CWI)::value;
                if (f32)
                    kILPrimal<float, CW><<<ilGrid(grid_), kBlock, 0, st_>>>(
                        di, il_.slots32, il_.counts, il_.lanes32, w.n, w.m, w.cs, w.ri, w.valtf,
                        w.cf, w.tauf, il_.f.y, il_.f.x, il_.f.x0, il_.f.xr, il_.f.l, il_.f.u,
                        nullptr, nullptr, nullptr, k);
                else
                    kILPrimal<double, CW><<<ilGrid(grid_), kBlock, 0, st_>>>(
                        di, il_.slots64, il_.counts + 1, il_.lanes64, w.n, w.m, w.cs, w.ri, w.valt,
                        w.c, w.tau, il_.d.y, il_.d.x, il_.d.x0, il_.d.xr, il_.d.l, il_.d.u,
                        w.l, w.u, il_.bf64, k);
            });
            MPX_HIP(hipEventRecord(e[1], st_));
            withCW(cwd_, [&](auto CWI) {
                constexpr int CW = decltype(CWI)::value;
                if (f32)
                    kILDual<float, CW><<<ilGrid(grid_), kBlock, 0, st_>>>(
                        di, il_.slots32, il_.counts, il_.lanes32, w.n, w.m, w.rs, w.ci, w.valf,
                        w.sigf, w.rlf, w.ruf, il_.f.xr, il_.f.y, il_.f.y0, k);
                else
                    kILDual<double, CW><<<ilGrid(grid_), kBlock, 0, st_>>>(
                        di, il_.slots64, il_.counts + 1, il_.lanes64, w.n, w.m, w.rs, w.ci, w.val,
                        w.sig, w.rl, w.ru, il_.d.xr, il_.d.y, il_.d.y0, k);
            });
            MPX_HIP(hipEventRecord(e[2], st_));
            MPX_HIP(hipEventSynchronize(e[2]));
            float a = 0.0f, b = 0.0f;
            MPX_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            MPX_HIP(hipEventElapsedTime(&b, e[1], e[2]));
            tp += a;
            td += b;
        }
        int c2[2];
        MPX_HIP(hipMemcpy(c2, il_.counts, sizeof(c2), hipMemcpyDeviceToHost));
        active = f32 ? c2[0] : c2[1];
    } else {
        int* cnt = static_cast<int*>(d_cnt_buf_->p);
        const int gk = ceilDiv(K, kBlock);
        kTileCounts<<<gk, kBlock, 0, st_>>>(di, K, cnt);
        kScan4<<<1, 1024, 0, st_>>>(cnt, K, di, totals_);
        kTileFill<<<gk, kBlock, 0, st_>>>(di, K, cnt, tc32_, tr32_, tc64_, tr64_, rec32_, rec64_);
        for (int k = 1; k <= launches; ++k) {
            MPX_HIP(hipEventRecord(e[0], st_));
            if (f32) kPackedPrimal<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tc32_, totals_ + 0, k);
            else kPackedPrimal<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tc64_, totals_ + 2, k);
            MPX_HIP(hipEventRecord(e[1], st_));
            if (f32) kPackedDual<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tr32_, totals_ + 1, k);
            else kPackedDual<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tr64_, totals_ + 3, k);
            MPX_HIP(hipEventRecord(e[2], st_));
            MPX_HIP(hipEventSynchronize(e[2]));
            float a = 0.0f, b = 0.0f;
            MPX_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            MPX_HIP(hipEventElapsedTime(&b, e[1], e[2]));
            tp += a;
            td += b;
        }
        int h5[5];
        MPX_HIP(hipMemcpy(h5, totals_, sizeof(h5), hipMemcpyDeviceToHost));
        active = h5[4];
    }
    MPX_HIP(hipGetLastError());
    for (auto& ev : e) (void)hipEventDestroy(ev);
    *primal_us = 1000.0 * tp / launches;
    *dual_us = 1000.0 * td / launches;
    *slots = active;
}

std::vector<double> BatchEngine::finishTimesMs() {
    const int ni = count();
    std::vector<BCtrl> ctrl(static_cast<size_t>(ni));
    unsigned long long t0 = 0;
    MPX_HIP(hipMemcpy(ctrl.data(), ctrl_all_, sizeof(BCtrl) * ni, hipMemcpyDeviceToHost));
    MPX_HIP(hipMemcpy(&t0, stamp_, sizeof(t0), hipMemcpyDeviceToHost));
    int dev = 0, khz = 100000;
    MPX_HIP(hipGetDevice(&dev));
    MPX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    std::vector<double> out(static_cast<size_t>(ni), 0.0);
    for (int i = 0; i < ni; ++i)
        out[i] = ctrl[i].t_finish > t0 ? static_cast<double>(ctrl[i].t_finish - t0) / khz : 0.0;
    return out;
}

void BatchEngine::launchRound(const IterPreset& p, bool timed) {
    if (interleaved_) {
        cwp_ = p.gc > 0 ? p.gc : 1;
        cwd_ = p.gr > 0 ? p.gr : 1;
    }
    if (timed || graph_failed_) {
        if (interleaved_) enqueueInterleavedRound(timed);
        else enqueueRound(p, timed);
        MPX_HIP(hipGetLastError());
        return;
    }
    G& g = graphs_[{p.gc, p.gr * 16 + (interleaved_ ? roundSets() : 0) * 2 + (fp32_gone_ ? 1 : 0)}];
    if (g.e == nullptr) {
        bool ok = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            if (interleaved_) enqueueInterleavedRound(false);
            else enqueueRound(p, false);
            ok = hipStreamEndCapture(st_, &g.g) == hipSuccess && g.g != nullptr &&
                 hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0) == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            g.reset();
            graph_failed_ = true;
            if (interleaved_) enqueueInterleavedRound(false);
            else enqueueRound(p, false);
            MPX_HIP(hipGetLastError());
            return;
        }
    }
    MPX_HIP(hipGraphLaunch(g.e, st_));
}

// After a lane-per-instance round of a mixed-precision batch: the phase
// counts of the next round, and a check that the round just run left out
// no kernel set that had instances in it.
void BatchEngine::readNextCounts() {
    if (!interleaved_ || !mixed_) return;
    const int ran = roundSets();
    int h[7];
    MPX_HIP(hipMemcpyAsync(h, il_.counts, 3 * sizeof(int), hipMemcpyDeviceToHost, st_));
    MPX_HIP(hipMemcpyAsync(h + 3, il_.next, 3 * sizeof(int), hipMemcpyDeviceToHost, st_));
    MPX_HIP(hipStreamSynchronize(st_));
    if ((h[0] > 0 && !(ran & kSetPlain)) || (h[2] > 0 && !(ran & kSetRefine)) ||
        (h[1] > 0 && !(ran & kSetFp64))) {
        setError("internal: a lane-per-instance round left out a kernel set with instances");
        throw DeviceError();
    }
    for (int p = 0; p < 3; ++p) next_[p] = h[3 + p];
    next_known_ = true;
}

int BatchEngine::step(int rounds) {
    const int K = count();
    if (K == 0) return 0;
    int* dcount = static_cast<int*>(d_count_buf_->p);
    for (int r = 0; r < rounds; ++r) {
        // Online autotuning: the first rounds try the candidate presets (all
        // rounds do real work), timing the iteration part with HIP events.
        if (!tune_left_.empty()) {
            const IterPreset p = tune_left_.back();
            tune_left_.pop_back();
            launchRound(p, true);
            float ms = 0.0f;
            MPX_HIP(hipEventSynchronize(ev1_));
            MPX_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
            tune_res_.push_back({ms, p});
            ++rounds_;
            iter_ms_ += ms;
            ++rounds_timed_;
            if (tune_left_.empty()) {
                std::sort(tune_res_.begin(), tune_res_.end(),
                          [](const auto& a, const auto& b) { return a.first < b.first; });
                cur_ = tune_res_.front().second;
                if (interleaved_) tuned_ = cur_;
            }
            readNextCounts();
            continue;
        }
        launchRound(cur_, false);
        ++rounds_;
        readNextCounts();
    }
    kZeroInt<<<1, 1, 0, st_>>>(dcount);
    kCountActive<<<ceilDiv(K, kBlock), kBlock, 0, st_>>>(static_cast<const BInst*>(d_insts_buf_->p),
                                                        K, dcount);
    MPX_HIP(hipGetLastError());
    int active = 0;
    int n32 = 1;
    MPX_HIP(hipMemcpyAsync(&active, dcount, sizeof(int), hipMemcpyDeviceToHost, st_));
    if (mixed_ && !fp32_gone_) {
        // FP32 instances at the start of the last round (0 stays 0).
        const int* src = interleaved_ ? il_.counts : totals_;
        int c3[3] = {1, 1, 1};
        MPX_HIP(hipMemcpyAsync(c3, src, sizeof(c3), hipMemcpyDeviceToHost, st_));
        MPX_HIP(hipStreamSynchronize(st_));
        // Instances in the FP32 kernels (plain or refinement).
        n32 = interleaved_ ? c3[0] + c3[2] : c3[0] + c3[1];
        if (rounds_ > 0 && n32 == 0) fp32_gone_ = true;
    }
    MPX_HIP(hipStreamSynchronize(st_));
    return active;
}

void BatchEngine::run(std::chrono::steady_clock::time_point t0) {
    const int K = count();
    if (K == 0) return;
    while (true) {
        const int active = step(2);
        if (active == 0) break;
        const bool timeUp = o_.max_solve_seconds >= 0.0 &&
                            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                    .count() >= o_.max_solve_seconds;
        const bool stop = o_.stop_fn != nullptr && o_.stop_fn(o_.stop_user) != 0;
        if (timeUp || stop) {
            kAbortActive<<<ceilDiv(K, kBlock), kBlock, 0, st_>>>(
                static_cast<const BInst*>(d_insts_buf_->p), K,
                timeUp ? MPX_STATUS_TIME_LIMIT : MPX_STATUS_ITER_LIMIT);
            MPX_HIP(hipStreamSynchronize(st_));
            break;
        }
    }
}

void BatchEngine::results(double* const* x_out, double* const* y_out, mpx_pdlp_result* res) {
    const int ni = count();
    if (ni == 0) return;
    // Recover all instances into one device staging block, then one copy out.
    std::vector<size_t> xo(static_cast<size_t>(ni)), yo(static_cast<size_t>(ni));
    size_t total = 0;
    for (int i = 0; i < ni; ++i) {
        xo[i] = total;
        total += static_cast<size_t>(insts_[i].n);
        yo[i] = total;
        total += static_cast<size_t>(insts_[i].m);
    }
    DevBuf stage(total * sizeof(double) + 256);
    double* base = static_cast<double*>(stage.p);
    std::vector<double*> xs(static_cast<size_t>(ni)), ys(static_cast<size_t>(ni));
    for (int i = 0; i < ni; ++i) {
        xs[i] = base + xo[i];
        ys[i] = base + yo[i];
    }
    DevBuf ptrs(sizeof(double*) * 2 * ni);
    MPX_HIP(hipMemcpyAsync(ptrs.p, xs.data(), sizeof(double*) * ni, hipMemcpyHostToDevice, st_));
    MPX_HIP(hipMemcpyAsync(static_cast<double**>(ptrs.p) + ni, ys.data(), sizeof(double*) * ni,
                           hipMemcpyHostToDevice, st_));
    kInstRecover<<<ni, kBlock, 0, st_>>>(static_cast<const BInst*>(d_insts_buf_->p),
                                         static_cast<double* const*>(ptrs.p),
                                         static_cast<double* const*>(ptrs.p) + ni);
    MPX_HIP(hipGetLastError());
    std::vector<BCtrl> ctrl(static_cast<size_t>(ni));
    std::vector<double> host(total);
    MPX_HIP(hipMemcpyAsync(ctrl.data(), ctrl_all_, sizeof(BCtrl) * ni, hipMemcpyDeviceToHost, st_));
    MPX_HIP(hipMemcpyAsync(host.data(), base, total * sizeof(double), hipMemcpyDeviceToHost, st_));
    MPX_HIP(hipStreamSynchronize(st_));
    for (int i = 0; i < ni; ++i) {
        const mpx_lp& lp = *owners_host_[inst_owner_[i]];
        const int n = insts_[i].n, m = insts_[i].m;
        const BCtrl& c = ctrl[i];
        mpx_pdlp_result& r = res[i];
        std::memset(&r, 0, sizeof(r));
        r.status = c.status;
        r.iterations = c.total;
        r.work_units = static_cast<double>(c.total) *
                       (4.0 * static_cast<double>(insts_[i].nnz) + 2.0 * static_cast<double>(m + n));
        r.rel_primal = o_.optimality_norm == 1 ? c.rel_p_inf : c.rel_p;
        r.rel_dual = o_.optimality_norm == 1 ? c.rel_d_inf : c.rel_d;
        r.rel_gap = c.gap;
        r.primal_obj = c.pobj;
        r.dual_obj = c.dobj;
        r.step = c.step;
        r.primal_weight = c.pw;
        r.restarts = c.restarts;
        r.iterations_fp32 = c.fp32_iters;
        r.iterations_refine = c.refine_iters;
        r.setup_seconds = setup_seconds_;
        const double* hx = host.data() + xo[i];
        const double* hy = host.data() + yo[i];
        // Optimal: solution; Infeasible: Farkas ray in y; Unbounded: primal
        // ray in x (reference pdlp.cpp:2318-2340); otherwise zeros.
        const bool keepX = c.status == MPX_STATUS_OPTIMAL || c.status == MPX_STATUS_UNBOUNDED;
        const bool keepY = c.status == MPX_STATUS_OPTIMAL || c.status == MPX_STATUS_INFEASIBLE;
        if (c.status == MPX_STATUS_OPTIMAL) {
            double objv = lp.obj_offset;
            for (int j = 0; j < n; ++j) objv += lp.obj[j] * hx[j];
            r.objective = objv;
        }
        if (x_out && x_out[i]) {
            if (keepX) std::memcpy(x_out[i], hx, n * 8ull);
            else std::memset(x_out[i], 0, n * 8ull);
        }
        if (y_out && y_out[i]) {
            if (keepY) std::memcpy(y_out[i], hy, m * 8ull);
            else std::memset(y_out[i], 0, m * 8ull);
        }
    }
}

bool validLpB(const mpx_lp* lp) {
    if (!lp || lp->num_rows <= 0 || lp->num_cols <= 0 || !lp->row_starts) return false;
    if (lp->row_starts[0] != 0) return false;
    const int nnz = lp->row_starts[lp->num_rows];
    if (nnz < 0 || (nnz > 0 && (!lp->col_indices || !lp->values))) return false;
    if (!lp->obj || !lp->col_lower || !lp->col_upper || !lp->row_lower || !lp->row_upper)
        return false;
    for (int i = 0; i < lp->num_rows; ++i)
        if (lp->row_starts[i + 1] < lp->row_starts[i]) return false;
    for (int k = 0; k < nnz; ++k)
        if (lp->col_indices[k] < 0 || lp->col_indices[k] >= lp->num_cols) return false;
    return true;
}

bool batchOptionsOk(const mpx_pdlp_options* o) {
    if (!o || o->termination_eval_frequency <= 0) return false;
    if (o->adaptive_step || (o->anderson_acceleration && o->anderson_window_size >= 2) ||
        o->matrix_zero_tol > 0.0) {
        setError("batched PDLP: adaptive_step, anderson_acceleration and matrix_zero_tol are not "
                 "supported");
        return false;
    }
    return true;
}

// ---- serving loop (continuous batching) ---------------------------------------
// One device context with a fixed table of instance slots.  The packed
// iteration kernels, the tile lists and the per-instance evaluation / control
// kernel are the batched engine's; only the table rows change between
// rounds: activation writes a new instance into a free row (and initialises
// it), retirement recovers the finished ones and leaves their rows with
// phase == done, which every kernel skips.

// First-fit allocator with coalescing over one device allocation.
class DevicePool {
public:
    explicit DevicePool(size_t bytes) : buf_(bytes), cap_(bytes) {
        if (bytes) free_[0] = bytes;
    }
    // nullptr when no free block is large enough.
    void* alloc(size_t bytes) {
        bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            live_[off] = bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return static_cast<char*>(buf_.p) + off;
        }
        return nullptr;
    }
    void release(void* p) {
        if (!p) return;
        const size_t off =
            static_cast<size_t>(static_cast<char*>(p) - static_cast<char*>(buf_.p));
        auto lv = live_.find(off);
        if (lv == live_.end()) return;
        size_t start = off, sz = lv->second;
        used_ -= sz;
        live_.erase(lv);
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + sz == nx->first) {
            sz += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                sz += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = sz;
    }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    DevBuf buf_;
    size_t cap_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_, live_;
};

__global__ void kPlaceInsts(const BInst* __restrict__ src, const int* __restrict__ slots,
                            int count, BInst* __restrict__ table) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) table[slots[t]] = src[t];
}

__global__ void kAbortSlot(BCtrl* __restrict__ c, int status) { abortInstance(*c, status); }

// An LP the server holds a copy of: a registered base, or a fresh request.
struct ServeLp {
    std::vector<int> rs, ci;
    std::vector<double> val, obj, lo, hi, rl, ru;
    mpx_lp lp{};
    BOwner owner{};
    int preset = 0;

    explicit ServeLp(const mpx_lp& src) {
        const int m = src.num_rows, n = src.num_cols, nnz = src.row_starts[m];
        rs.assign(src.row_starts, src.row_starts + m + 1);
        ci.assign(src.col_indices, src.col_indices + nnz);
        val.assign(src.values, src.values + nnz);
        obj.assign(src.obj, src.obj + n);
        lo.assign(src.col_lower, src.col_lower + n);
        hi.assign(src.col_upper, src.col_upper + n);
        rl.assign(src.row_lower, src.row_lower + m);
        ru.assign(src.row_upper, src.row_upper + m);
        lp = src;
        lp.row_starts = rs.data();
        lp.col_indices = ci.data();
        lp.values = val.data();
        lp.obj = obj.data();
        lp.col_lower = lo.data();
        lp.col_upper = hi.data();
        lp.row_lower = rl.data();
        lp.row_upper = ru.data();
    }
    size_t bytes() const {
        return ownerBytes(lp.num_rows, lp.num_cols, lp.row_starts[lp.num_rows]);
    }
    // Carves the matrix arrays (uploadOwner() fills them).
    void carve(const mpx_pdlp_options& o, Carve& cv) {
        carveOwner(lp, o, cv, owner);
        preset = ownerPreset(o, owner.nnz, owner.m, owner.n);
    }
    // Lane-per-instance state of a base's variants.
    IlSet il;
};

// Device setup of the fresh LPs admitted in one step, run by the server's
// setup thread on its own stream.
struct SetupJob {
    std::unique_ptr<SetupCohort> cohort;
    void* mem = nullptr;  // cohort scratch (pool)
    // What to upload first: each fresh LP's host data into its carved arrays
    // (the requests own the host data and outlive the job).
    std::vector<std::pair<const ServeLp*, const OwnerHost*>> uploads;
    std::atomic<bool> done{false};
    std::atomic<bool> failed{false};
    std::string error;  // set by the setup thread before `done`
};

struct ServeReq {
    int64_t ticket = 0;
    int base = -1;                  // registered base of a bounds variant
    std::unique_ptr<ServeLp> own;   // a fresh LP
    OwnerHost oh;                   // its transpose and power start vector
    std::vector<double> vlo, vhi;   // a variant's column bounds
    std::vector<double> wx, wy;     // a variant's warm start (empty: cold)
    std::chrono::steady_clock::time_point t_submit, t_active;
    void* mem = nullptr;
    BInst inst{};
    BInitArgs args{};
    double *xo = nullptr, *yo = nullptr;
    std::shared_ptr<SetupJob> job;  // fresh LPs until their setup is done
    int slot = -1;
};

struct ServeDone {
    mpx_pdlp_result res{};
    std::vector<double> x, y;
};

template <typename T>
struct PinnedArray {
    T* p = nullptr;
    explicit PinnedArray(size_t count) {
        MPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(1, count),
                              hipHostMallocDefault));
    }
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    ~PinnedArray() {
        if (p) (void)hipHostFree(p);
    }
};

class Server {
public:

    Server(const mpx_pdlp_options& o, double budget, int slots)
        : o_(o), S_(slots > 0 ? slots : 1024), h_ctrl_(S_), h_new_(S_), h_args_(S_),
          h_slots_(S_), h_ptrs_(2 * static_cast<size_t>(S_)) {
        MPX_HIP(hipSetDevice(o.device));
        MPX_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        bo_ = makeBOptions(o);
        mixed_ = o.precision == MPX_PREC_MIXED;
        graph_failed_ = graphsDisabled();
        size_t freeb = 0, total = 0;
        MPX_HIP(hipMemGetInfo(&freeb, &total));
        const size_t want = budget > 0.0 ? static_cast<size_t>(budget) : freeb / 2;
        pool_ = std::make_unique<DevicePool>(std::max<size_t>(want, 1 << 20));
        const size_t S = static_cast<size_t>(S_);
        tab_ = std::make_unique<DevBuf>(sizeof(BInst) * S);
        ctrl_ = std::make_unique<DevBuf>(sizeof(BCtrl) * S);
        ev_.alloc(S_);
        cnt_ = std::make_unique<DevBuf>(sizeof(int) * 4 * S);
        recs_ = std::make_unique<DevBuf>(pad(S, sizeof(PRec<float>)) + pad(S, sizeof(PRec<double>)));
        rec32_ = static_cast<PRec<float>*>(recs_->p);
        rec64_ = reinterpret_cast<PRec<double>*>(static_cast<char*>(recs_->p) +
                                                 pad(S, sizeof(PRec<float>)));
        newi_ = std::make_unique<DevBuf>(sizeof(BInst) * S);
        newa_ = std::make_unique<DevBuf>(sizeof(BInitArgs) * S);
        news_ = std::make_unique<DevBuf>(sizeof(int) * S);
        ptrs_ = std::make_unique<DevBuf>(sizeof(double*) * 2 * S);
        // Empty rows: no data, control state done.
        BCtrl* dc = static_cast<BCtrl*>(ctrl_->p);
        for (int s = 0; s < S_; ++s) {
            std::memset(&h_new_.p[s], 0, sizeof(BInst));
            h_new_.p[s].ctrl = dc + s;
            std::memset(&h_ctrl_.p[s], 0, sizeof(BCtrl));
            h_ctrl_.p[s].phase = kPhaseDone;
        }
        MPX_HIP(hipMemcpyAsync(tab_->p, h_new_.p, sizeof(BInst) * S, hipMemcpyHostToDevice, st_));
        MPX_HIP(hipMemcpyAsync(ctrl_->p, h_ctrl_.p, sizeof(BCtrl) * S, hipMemcpyHostToDevice, st_));
        MPX_HIP(hipStreamSynchronize(st_));
        for (int s = S_ - 1; s >= 0; --s) free_slots_.push_back(s);
        slot_req_.resize(S);
        starting_.assign(S, 0);
        int dev = 0, cus = 256;
        MPX_HIP(hipGetDevice(&dev));
        MPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid_ = std::max(kNumXcd, (cus * 8 / kNumXcd) * kNumXcd);
        // Initial tile capacity per list (8 MB of tiles for pools of 1 GB and
        // more); it grows on demand.
        const size_t t0 = pool_->capacity() >= (size_t(1) << 30) ? size_t(1) << 20 : 1024;
        growTiles(t0, t0);
        setup_thread_ = std::thread([this] { setupLoop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        job_cv_.notify_all();
        if (setup_thread_.joinable()) setup_thread_.join();
        (void)hipStreamSynchronize(st_);
        for (auto& kv : graphs_) kv.second.reset();
        (void)hipStreamDestroy(st_);
    }

    int addBase(const mpx_lp& lp) {
        auto b = std::make_unique<ServeLp>(lp);
        const OwnerHost h = ownerHostData(b->lp);
        const size_t il_bytes = IlSet::bytes(S_, b->lp.num_cols, b->lp.num_rows, mixed_);
        const size_t need = b->bytes() + il_bytes;
        void* mem = pool_->alloc(need);
        if (!mem) {
            setError("mpx_server_add_base: the base LP does not fit the server's memory pool");
            throw OutOfMemory();
        }
        Carve cv{static_cast<char*>(mem), 0, need};
        b->carve(o_, cv);
        uploadOwner(b->lp, h, b->owner, st_);
        b->il.carve(cv, S_, b->lp.num_cols, b->lp.num_rows, mixed_);
        b->il.owner = static_cast<int>(bases_.size());
        SetupCohort cohort({b->owner}, {b->preset});
        void* scratch = pool_->alloc(cohort.bytes());
        if (!scratch) {
            pool_->release(mem);
            setError("mpx_server_add_base: the base LP does not fit the server's memory pool");
            throw OutOfMemory();
        }
        cohort.carve(static_cast<char*>(scratch));
        cohort.run(st_);
        pool_->release(scratch);
        bases_.push_back(std::move(b));
        // The round graphs include every base's lane-per-instance kernels.
        for (auto& kv : graphs_) kv.second.reset();
        graphs_.clear();
        return static_cast<int>(bases_.size()) - 1;
    }
    int numBases() const { return static_cast<int>(bases_.size()); }
    int numCols(int base) const { return bases_[base]->lp.num_cols; }

    int64_t submit(const mpx_lp& lp) {
        auto r = std::make_unique<ServeReq>();
        r->own = std::make_unique<ServeLp>(lp);
        r->oh = ownerHostData(r->own->lp);
        return enqueue(std::move(r));
    }
    int64_t submitBounds(int base, const double* lo, const double* hi,
                         const double* wx = nullptr, const double* wy = nullptr) {
        auto r = std::make_unique<ServeReq>();
        const int n = bases_[base]->lp.num_cols, m = bases_[base]->lp.num_rows;
        r->base = base;
        r->vlo.assign(lo, lo + n);
        r->vhi.assign(hi, hi + n);
        if (wx && wy) {
            r->wx.assign(wx, wx + n);
            r->wy.assign(wy, wy + m);
        }
        return enqueue(std::move(r));
    }

    int step() {
        using clk = std::chrono::steady_clock;
        auto ms = [](clk::time_point a, clk::time_point b) {
            return 1000.0 * std::chrono::duration<double>(b - a).count();
        };
        const auto t0 = clk::now();
        admit();
        const auto t1 = clk::now();
        activate();
        const auto t2 = clk::now();
        iterate();
        const auto t3 = clk::now();
        retire();
        enforceTimeLimit();
        const auto t4 = clk::now();
        admit_ms_ += ms(t0, t1);
        activate_ms_ += ms(t1, t2);
        retire_ms_ += ms(t3, t4);
        max_step_ms_ = std::max(max_step_ms_, ms(t0, t4));
        return static_cast<int>(queued_.size() + preparing_.size()) + running_;
    }

    // 1: done (and forgotten), 0: not finished, -1: unknown ticket.
    int result(int64_t ticket, double* x, double* y, mpx_pdlp_result* res) {
        auto it = done_.find(ticket);
        if (it == done_.end()) return ticket >= 0 && ticket < next_ticket_ ? 0 : -1;
        if (res) *res = it->second.res;
        if (x && !it->second.x.empty())
            std::memcpy(x, it->second.x.data(), it->second.x.size() * sizeof(double));
        if (y && !it->second.y.empty())
            std::memcpy(y, it->second.y.data(), it->second.y.size() * sizeof(double));
        done_.erase(it);
        return 1;
    }

    int poll(int64_t* out, int cap) {
        int k = 0;
        while (k < cap && !newly_done_.empty()) {
            out[k++] = newly_done_.front();
            newly_done_.pop_front();
        }
        return k;
    }

    void stats(mpx_server_stats* s) const {
        std::memset(s, 0, sizeof(*s));
        s->submitted = next_ticket_;
        s->finished = finished_;
        s->queued = static_cast<int32_t>(queued_.size());
        s->preparing = static_cast<int32_t>(preparing_.size());
        s->running = running_;
        s->slots = S_;
        s->rounds = rounds_;
        s->pool_bytes = static_cast<double>(pool_->capacity());
        s->pool_used_bytes = static_cast<double>(pool_->used());
        s->pool_peak_bytes = static_cast<double>(pool_->peak());
        s->iterate_ms = iterate_ms_;
        s->instance_iterations = inst_iters_;
        s->admit_ms = admit_ms_;
        s->activate_ms = activate_ms_;
        s->retire_ms = retire_ms_;
        s->max_step_ms = max_step_ms_;
        s->setup_query_ms = query_ms_;
        s->setup_wait_ms = wait_ms_;
    }

private:
    struct OutOfMemory {};
    struct G {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        void reset() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
            e = nullptr;
            g = nullptr;
        }
    };

    const mpx_lp& lpOf(const ServeReq& r) const {
        return r.own ? r.own->lp : bases_[r.base]->lp;
    }

    int64_t enqueue(std::unique_ptr<ServeReq> r) {
        r->ticket = next_ticket_++;
        r->t_submit = std::chrono::steady_clock::now();
        const int64_t t = r->ticket;
        queued_.push_back(std::move(r));
        return t;
    }

    size_t bytesFor(const ServeReq& r) const {
        const mpx_lp& lp = lpOf(r);
        const size_t m = lp.num_rows, n = lp.num_cols;
        size_t b = instBytes(lp.num_rows, lp.num_cols, !r.own) + pad(n, 8) + pad(m, 8);
        if (r.own) b += r.own->bytes();
        return b;
    }

    void finishNow(std::unique_ptr<ServeReq> r, int status) {
        ServeDone d;
        d.res.status = status;
        d.res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                      r->t_submit).count();
        d.res.setup_seconds = d.res.seconds;
        done_[r->ticket] = std::move(d);
        newly_done_.push_back(r->ticket);
        ++finished_;
    }

    // Queued -> preparing: memory, a slot, uploads on the main stream (idle
    // between rounds, so the pageable copies do not wait behind kernels),
    // then for fresh LPs the scaling and power iteration on a side stream
    // that waits for the uploads.  Bounds variants need no setup.
    void admit() {
        std::vector<ServeReq*> fresh;
        while (!queued_.empty() && !free_slots_.empty()) {
            ServeReq& r = *queued_.front();
            const size_t need = bytesFor(r);
            if (need > pool_->capacity()) {
                setError("serving loop: request larger than the memory pool");
                std::unique_ptr<ServeReq> q = std::move(queued_.front());
                queued_.pop_front();
                finishNow(std::move(q), MPX_STATUS_ERROR);
                continue;
            }
            void* mem = pool_->alloc(need);
            if (!mem) break;  // first in, first out: wait for memory
            std::unique_ptr<ServeReq> q = std::move(queued_.front());
            queued_.pop_front();
            q->mem = mem;
            q->slot = free_slots_.back();
            free_slots_.pop_back();
            Carve cv{static_cast<char*>(mem), 0, need};
            BCtrl* ctrl = static_cast<BCtrl*>(ctrl_->p) + q->slot;
            if (q->own) {
                // Uploads and setup are the setup thread's (its own stream).
                q->own->carve(o_, cv);
                carveInstance(q->own->owner, q->own->lp, -1, q->own->preset, nullptr, nullptr,
                              mixed_, cv, st_, ctrl, q->inst, q->args);
                fresh.push_back(q.get());
            } else {
                const ServeLp& b = *bases_[q->base];
                carveInstance(b.owner, b.lp, q->base, b.preset, q->vlo.data(), q->vhi.data(),
                              mixed_, cv, st_, ctrl, q->inst, q->args,
                              q->wx.empty() ? nullptr : q->wx.data(),
                              q->wy.empty() ? nullptr : q->wy.data());
                q->inst.il = 1;
            }
            q->xo = cv.take<double>(static_cast<size_t>(q->inst.n));
            q->yo = cv.take<double>(static_cast<size_t>(q->inst.m));
            preparing_.push_back(std::move(q));
        }
        if (fresh.empty()) return;
        std::vector<BOwner> owners;
        std::vector<int> presets;
        for (ServeReq* q : fresh) {
            owners.push_back(q->own->owner);
            presets.push_back(q->own->preset);
        }
        auto job = std::make_shared<SetupJob>();
        job->cohort = std::make_unique<SetupCohort>(std::move(owners), std::move(presets));
        job->mem = pool_->alloc(job->cohort->bytes());
        if (!job->mem) {
            // No room for the setup scratch: put the fresh requests back at
            // the head of the queue, in order.
            MPX_HIP(hipStreamSynchronize(st_));
            for (auto it = preparing_.end(); it != preparing_.begin();) {
                --it;
                if (!(*it)->own || (*it)->job) continue;
                if (std::find(fresh.begin(), fresh.end(), it->get()) == fresh.end()) continue;
                std::unique_ptr<ServeReq> q = std::move(*it);
                it = preparing_.erase(it);
                pool_->release(q->mem);
                q->mem = nullptr;
                free_slots_.push_back(q->slot);
                q->slot = -1;
                queued_.push_front(std::move(q));
            }
            return;
        }
        job->cohort->carve(static_cast<char*>(job->mem));
        for (ServeReq* q : fresh) {
            q->job = job;
            job->uploads.emplace_back(q->own.get(), &q->oh);
        }
        jobs_live_.push_back(job);
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(job);
        }
        job_cv_.notify_one();
    }

    // Setup thread: runs each job's cohort on its own stream.
    void setupLoop() {
        hipStream_t ws = nullptr;
        bool ok = hipSetDevice(o_.device) == hipSuccess &&
                  hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) == hipSuccess;
        while (true) {
            std::shared_ptr<SetupJob> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                job_cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) break;  // stop_ and nothing left
                job = jobs_.front();
                jobs_.pop_front();
            }
            bool good = ok;
            if (good) {
                try {
                    for (const auto& u : job->uploads)
                        uploadOwner(u.first->lp, *u.second, u.first->owner, ws);
                    job->cohort->run(ws);
                } catch (const DeviceError&) {
                    good = false;
                    job->error = lastError();
                }
            } else {
                job->error = "setup thread: no device stream";
            }
            job->failed.store(!good);
            {
                std::lock_guard<std::mutex> lk(mu_);
                job->done.store(true);
            }
            done_cv_.notify_all();
        }
        if (ws) (void)hipStreamDestroy(ws);
    }

    // Preparing -> running for every request whose setup has completed:
    // table rows written and instances initialised on the main stream.
    void activate() {
        if (preparing_.empty()) {
            reapJobs();
            return;
        }
        if (running_ == 0) {
            // Nothing to iterate meanwhile: wait for the first setup.
            bool any = false;
            for (auto& q : preparing_) any = any || !q->job || q->job->done.load();
            if (!any) {
                const auto tq = std::chrono::steady_clock::now();
                std::shared_ptr<SetupJob> j = preparing_.front()->job;
                std::unique_lock<std::mutex> lk(mu_);
                done_cv_.wait(lk, [&] { return j->done.load(); });
                wait_ms_ += 1000.0 * std::chrono::duration<double>(
                                         std::chrono::steady_clock::now() - tq).count();
            }
        }
        int k = 0;
        size_t need_c = 0, need_r = 0;
        std::vector<std::unique_ptr<ServeReq>> still;
        const auto now = std::chrono::steady_clock::now();
        for (auto& q : preparing_) {
            if (q->job) {
                if (!q->job->done.load()) {
                    still.push_back(std::move(q));
                    continue;
                }
                const bool failed = q->job->failed.load();
                const std::string why = failed ? q->job->error : std::string();
                q->job.reset();
                if (failed) {
                    setError("serving loop: device setup of a fresh LP failed: " + why);
                    pool_->release(q->mem);
                    free_slots_.push_back(q->slot);
                    finishNow(std::move(q), MPX_STATUS_ERROR);
                    continue;
                }
            }
            // The pageable uploads have been consumed: drop the host copies.
            q->oh = OwnerHost();
            h_new_.p[k] = q->inst;
            h_args_.p[k] = q->args;
            h_slots_.p[k] = q->slot;
            ++k;
            q->t_active = now;
            if (!q->inst.il) {
                need_c += static_cast<size_t>((q->inst.n + 15) / 16);
                need_r += static_cast<size_t>((q->inst.m + 15) / 16);
            }
            const int s = q->slot;
            slot_req_[s] = std::move(q);
            starting_[s] = 1;
            ++running_;
        }
        preparing_ = std::move(still);
        reapJobs();
        if (k == 0) return;
        tiles_c_ += need_c;
        tiles_r_ += need_r;
        if (tiles_c_ > cap_c_ || tiles_r_ > cap_r_) growTiles(tiles_c_ * 3 / 2, tiles_r_ * 3 / 2);
        MPX_HIP(hipMemcpyAsync(newi_->p, h_new_.p, sizeof(BInst) * k, hipMemcpyHostToDevice, st_));
        MPX_HIP(hipMemcpyAsync(newa_->p, h_args_.p, sizeof(BInitArgs) * k, hipMemcpyHostToDevice,
                               st_));
        MPX_HIP(hipMemcpyAsync(news_->p, h_slots_.p, sizeof(int) * k, hipMemcpyHostToDevice, st_));
        kPlaceInsts<<<ceilDiv(k, kBlock), kBlock, 0, st_>>>(static_cast<const BInst*>(newi_->p),
                                                           static_cast<const int*>(news_->p), k,
                                                           static_cast<BInst*>(tab_->p));
        kInstInit<<<k, kBlock, 0, st_>>>(static_cast<const BInst*>(newi_->p),
                                         static_cast<const BInitArgs*>(newa_->p), bo_,
                                         o_.primal_weight);
        MPX_HIP(hipGetLastError());
        // The staging arrays are reused next step, after this stream's sync.
        fp32_gone_ = false;
    }

    // Setup jobs that are done and no longer referenced by a request: their
    // scratch goes back to the pool.
    void reapJobs() {
        for (auto it = jobs_live_.begin(); it != jobs_live_.end();) {
            if ((*it)->done.load() && it->use_count() == 1) {
                pool_->release((*it)->mem);
                it = jobs_live_.erase(it);
            } else {
                ++it;
            }
        }
    }

    // Tile lists of the packed kernels, carved from the pool: no hipMalloc /
    // hipFree while serving (hipFree synchronises the whole device, i.e. it
    // would wait for every setup in flight on the side streams).
    void growTiles(size_t c, size_t r) {
        MPX_HIP(hipStreamSynchronize(st_));
        for (auto& kv : graphs_) kv.second.reset();
        graphs_.clear();
        const size_t nc = std::max(c, cap_c_), nr = std::max(r, cap_r_);
        const size_t bytes =
            2 * pad(nc, sizeof(Tile)) + 2 * pad(nr, sizeof(Tile)) + pad(16, sizeof(int));
        void* mem = pool_->alloc(bytes);
        if (!mem) {
            setError("serving loop: memory pool too small for the tile lists");
            throw DeviceError();
        }
        pool_->release(tiles_mem_);
        tiles_mem_ = mem;
        cap_c_ = nc;
        cap_r_ = nr;
        Carve cv{static_cast<char*>(mem), 0, bytes};
        tc32_ = cv.take<Tile>(cap_c_);
        tr32_ = cv.take<Tile>(cap_r_);
        tc64_ = cv.take<Tile>(cap_c_);
        tr64_ = cv.take<Tile>(cap_r_);
        totals_ = cv.take<int>(16);
    }

    void enqueueRound() {
        const int K = S_;
        const BInst* di = static_cast<const BInst*>(tab_->p);
        int* cnt = static_cast<int*>(cnt_->p);
        const int gk = ceilDiv(K, kBlock);
        kTileCounts<<<gk, kBlock, 0, st_>>>(di, K, cnt);
        kScan4<<<1, 1024, 0, st_>>>(cnt, K, di, totals_);
        kTileFill<<<gk, kBlock, 0, st_>>>(di, K, cnt, tc32_, tr32_, tc64_, tr64_, rec32_, rec64_);
        const bool f32 = mixed_ && !fp32_gone_;
        const int sets = roundSets();
        const int il = sets & kSetAll;
        const bool p32 = f32 && (sets & kPacked32), p64 = (sets & kPacked64) != 0;
        for (auto& b : bases_) ilGather(st_, grid_, di, K, b->owner, b->il, il);
        for (int k = 1; k <= bo_.N; ++k) {
            if (p32) {
                kPackedPrimal<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tc32_, totals_ + 0, k);
                kPackedDual<float><<<grid_, kBlock, 0, st_>>>(di, rec32_, tr32_, totals_ + 1, k);
            }
            if (p64) {
                kPackedPrimal<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tc64_, totals_ + 2, k);
                kPackedDual<double><<<grid_, kBlock, 0, st_>>>(di, rec64_, tr64_, totals_ + 3, k);
            }
            // One column / row per wave in both precisions: the pair the
            // batched engine's autotuner settles on for C2-sized bases (FP32
            // primal 174.5 us against 187 us at four columns per wave).
            for (auto& b : bases_)
                ilIterate(st_, grid_, di, b->owner, b->il, il, 1, 1, kIlFp64PerWave,
                          kIlFp64PerWave, k);
        }
        for (auto& b : bases_) ilScatter(st_, grid_, di, b->owner, b->il, il);
        if (mixed_) enqueueEval<float>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
        enqueueEval<double>(st_, grid_, di, K, bo_, ev_.off, ev_.parts);
    }

    // Kernel sets of the coming round: packed FP32 / FP64 for fresh LPs,
    // the lane-per-instance sets for bounds variants, from each running
    // slot's phase after the last round (a slot activated since then starts
    // in FP32 when mixed, else FP64).
    static constexpr int kPacked32 = 8, kPacked64 = 16;
    int roundSets() const {
        int sets = 0;
        for (int s = 0; s < S_; ++s) {
            if (!slot_req_[s]) continue;
            const int ph = starting_[s] ? (mixed_ ? kPhaseFp32 : kPhaseFp64) : h_ctrl_.p[s].phase;
            if (ph == kPhaseDone) continue;
            if (slot_req_[s]->inst.il)
                sets |= ph == kPhaseFp32 ? kSetPlain : ph == kPhaseRefine ? kSetRefine : kSetFp64;
            else
                sets |= iteratesFp32(ph) ? kPacked32 : kPacked64;
        }
        return sets;
    }

    void launchRound() {
        if (graph_failed_) {
            enqueueRound();
            MPX_HIP(hipGetLastError());
            return;
        }
        G& g = graphs_[roundSets() * 2 + (fp32_gone_ ? 1 : 0)];
        if (g.e == nullptr) {
            bool ok = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                enqueueRound();
                ok = hipStreamEndCapture(st_, &g.g) == hipSuccess && g.g != nullptr &&
                     hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0) == hipSuccess;
            }
            if (!ok) {
                (void)hipGetLastError();
                g.reset();
                graph_failed_ = true;
                enqueueRound();
                MPX_HIP(hipGetLastError());
                return;
            }
        }
        MPX_HIP(hipGraphLaunch(g.e, st_));
    }

    // One round of every running instance, then the control rows back.
    void iterate() {
        if (running_ == 0) return;
        const auto t0 = std::chrono::steady_clock::now();
        const size_t need = 4 + 3 * bases_.size();
        if (need > h_cnt_cap_) {
            h_cnt_ = std::make_unique<PinnedArray<int>>(need);
            h_cnt_cap_ = need;
        }
        const int ran = roundSets();
        launchRound();
        ++rounds_;
        MPX_HIP(hipMemcpyAsync(h_ctrl_.p, ctrl_->p, sizeof(BCtrl) * S_, hipMemcpyDeviceToHost, st_));
        // The round's own tile and slot counts: no kernel set that had work
        // may have been left out.
        int* hc = h_cnt_->p;
        MPX_HIP(hipMemcpyAsync(hc, totals_, 4 * sizeof(int), hipMemcpyDeviceToHost, st_));
        for (size_t b = 0; b < bases_.size(); ++b)
            MPX_HIP(hipMemcpyAsync(hc + 4 + 3 * b, bases_[b]->il.counts, 3 * sizeof(int),
                                   hipMemcpyDeviceToHost, st_));
        MPX_HIP(hipStreamSynchronize(st_));
        bool missed = (hc[0] > 0 && !(ran & kPacked32)) || (hc[2] > 0 && !(ran & kPacked64));
        for (size_t b = 0; b < bases_.size(); ++b) {
            const int* c = hc + 4 + 3 * b;
            missed = missed || (c[0] > 0 && !(ran & kSetPlain)) ||
                     (c[2] > 0 && !(ran & kSetRefine)) || (c[1] > 0 && !(ran & kSetFp64));
        }
        if (missed) {
            setError("internal: a serving round left out a kernel set with instances");
            throw DeviceError();
        }
        std::fill(starting_.begin(), starting_.end(), 0);
        if (mixed_ && !fp32_gone_) {
            // FP32 kernels stay in the round graph while a running instance
            // is in its FP32 phase (activation of new ones resets this).
            bool any = false;
            for (int s = 0; s < S_ && !any; ++s)
                any = slot_req_[s] && iteratesFp32(h_ctrl_.p[s].phase);
            fp32_gone_ = !any;
        }
        iterate_ms_ += 1000.0 * std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                              t0).count();
    }

    // Running instances that finished: recovery into their own output
    // arrays, copies out, memory and slot recycled.
    void retire() {
        if (running_ == 0) return;
        std::vector<int> fin;
        for (int s = 0; s < S_; ++s)
            if (slot_req_[s] && h_ctrl_.p[s].phase == kPhaseDone) fin.push_back(s);
        if (fin.empty()) return;
        const int k = static_cast<int>(fin.size());
        for (int i = 0; i < k; ++i) {
            const ServeReq& q = *slot_req_[fin[i]];
            h_new_.p[i] = q.inst;
            h_ptrs_.p[i] = q.xo;
            h_ptrs_.p[k + i] = q.yo;
        }
        MPX_HIP(hipMemcpyAsync(newi_->p, h_new_.p, sizeof(BInst) * k, hipMemcpyHostToDevice, st_));
        MPX_HIP(hipMemcpyAsync(ptrs_->p, h_ptrs_.p, sizeof(double*) * 2 * k, hipMemcpyHostToDevice,
                               st_));
        double* const* xp = static_cast<double* const*>(ptrs_->p);
        kInstRecover<<<k, kBlock, 0, st_>>>(static_cast<const BInst*>(newi_->p), xp, xp + k);
        MPX_HIP(hipGetLastError());
        std::vector<ServeDone> out(static_cast<size_t>(k));
        for (int i = 0; i < k; ++i) {
            const ServeReq& q = *slot_req_[fin[i]];
            out[i].x.resize(static_cast<size_t>(q.inst.n));
            out[i].y.resize(static_cast<size_t>(q.inst.m));
            MPX_HIP(hipMemcpyAsync(out[i].x.data(), q.xo, q.inst.n * 8ull, hipMemcpyDeviceToHost,
                                   st_));
            MPX_HIP(hipMemcpyAsync(out[i].y.data(), q.yo, q.inst.m * 8ull, hipMemcpyDeviceToHost,
                                   st_));
        }
        MPX_HIP(hipStreamSynchronize(st_));
        const auto now = std::chrono::steady_clock::now();
        for (int i = 0; i < k; ++i) {
            const int s = fin[i];
            std::unique_ptr<ServeReq> q = std::move(slot_req_[s]);
            const BCtrl& c = h_ctrl_.p[s];
            const mpx_lp& lp = lpOf(*q);
            const int n = q->inst.n, m = q->inst.m;
            ServeDone& d = out[i];
            mpx_pdlp_result& r = d.res;
            r.status = c.status;
            r.iterations = c.total;
            r.work_units = static_cast<double>(c.total) *
                           (4.0 * static_cast<double>(q->inst.nnz) + 2.0 * static_cast<double>(m + n));
            r.rel_primal = o_.optimality_norm == 1 ? c.rel_p_inf : c.rel_p;
            r.rel_dual = o_.optimality_norm == 1 ? c.rel_d_inf : c.rel_d;
            r.rel_gap = c.gap;
            r.primal_obj = c.pobj;
            r.dual_obj = c.dobj;
            r.step = c.step;
            r.primal_weight = c.pw;
            r.restarts = c.restarts;
            r.iterations_fp32 = c.fp32_iters;
            r.iterations_refine = c.refine_iters;
            r.seconds = std::chrono::duration<double>(now - q->t_submit).count();
            r.setup_seconds = std::chrono::duration<double>(q->t_active - q->t_submit).count();
            // Same conventions as the batched results: Optimal keeps both,
            // Infeasible the Farkas ray in y, Unbounded the ray in x.
            const bool keepX = c.status == MPX_STATUS_OPTIMAL || c.status == MPX_STATUS_UNBOUNDED;
            const bool keepY = c.status == MPX_STATUS_OPTIMAL || c.status == MPX_STATUS_INFEASIBLE;
            if (c.status == MPX_STATUS_OPTIMAL) {
                double objv = lp.obj_offset;
                for (int j = 0; j < n; ++j) objv += lp.obj[j] * d.x[j];
                r.objective = objv;
            }
            if (!keepX) std::fill(d.x.begin(), d.x.end(), 0.0);
            if (!keepY) std::fill(d.y.begin(), d.y.end(), 0.0);
            inst_iters_ += c.total;
            if (!q->inst.il) {
                tiles_c_ -= static_cast<size_t>((n + 15) / 16);
                tiles_r_ -= static_cast<size_t>((m + 15) / 16);
            }
            pool_->release(q->mem);
            free_slots_.push_back(s);
            --running_;
            done_[q->ticket] = std::move(d);
            newly_done_.push_back(q->ticket);
            ++finished_;
        }
    }

    // Per-request time limits, and the stop callback: while it returns
    // nonzero every running instance ends with ITER_LIMIT, as a stopped
    // single solve does (reference pdlp.cpp:2136-2140).
    void enforceTimeLimit() {
        if (running_ == 0) return;
        const bool stop = o_.stop_fn != nullptr && o_.stop_fn(o_.stop_user) != 0;
        if (o_.max_solve_seconds < 0.0 && !stop) return;
        const auto now = std::chrono::steady_clock::now();
        BCtrl* dc = static_cast<BCtrl*>(ctrl_->p);
        for (int s = 0; s < S_; ++s) {
            const ServeReq* q = slot_req_[s].get();
            if (!q || h_ctrl_.p[s].phase == kPhaseDone) continue;
            if (o_.max_solve_seconds >= 0.0 &&
                std::chrono::duration<double>(now - q->t_active).count() >= o_.max_solve_seconds)
                kAbortSlot<<<1, 1, 0, st_>>>(dc + s, MPX_STATUS_TIME_LIMIT);
            else if (stop)
                kAbortSlot<<<1, 1, 0, st_>>>(dc + s, MPX_STATUS_ITER_LIMIT);
        }
        MPX_HIP(hipGetLastError());
    }

    mpx_pdlp_options o_;
    BOptions bo_{};
    bool mixed_ = false;
    int S_;
    hipStream_t st_ = nullptr;
    std::unique_ptr<DevicePool> pool_;
    std::unique_ptr<DevBuf> tab_, ctrl_, cnt_, newi_, newa_, news_, ptrs_, recs_;
    EvalBufs ev_;
    PRec<float>* rec32_ = nullptr;
    PRec<double>* rec64_ = nullptr;
    void* tiles_mem_ = nullptr;
    PinnedArray<BCtrl> h_ctrl_;
    // A round's tile counts, then three slot counts per base (grown with
    // the bases).
    std::unique_ptr<PinnedArray<int>> h_cnt_;
    size_t h_cnt_cap_ = 0;
    std::vector<char> starting_;  // per slot: activated since the last round
    PinnedArray<BInst> h_new_;
    PinnedArray<BInitArgs> h_args_;
    PinnedArray<int> h_slots_;
    PinnedArray<double*> h_ptrs_;
    // Setup thread and its job queue.
    std::thread setup_thread_;
    std::mutex mu_;
    std::condition_variable job_cv_, done_cv_;
    std::deque<std::shared_ptr<SetupJob>> jobs_;
    std::vector<std::shared_ptr<SetupJob>> jobs_live_;
    bool stop_ = false;
    Tile *tc32_ = nullptr, *tr32_ = nullptr, *tc64_ = nullptr, *tr64_ = nullptr;
    int* totals_ = nullptr;
    size_t cap_c_ = 0, cap_r_ = 0, tiles_c_ = 0, tiles_r_ = 0;
    int grid_ = 2048;
    std::map<int, G> graphs_;
    bool graph_failed_ = false;
    bool fp32_gone_ = false;
    std::vector<std::unique_ptr<ServeLp>> bases_;
    std::deque<std::unique_ptr<ServeReq>> queued_;
    std::vector<std::unique_ptr<ServeReq>> preparing_;
    std::vector<std::unique_ptr<ServeReq>> slot_req_;
    std::vector<int> free_slots_;
    int running_ = 0;
    std::map<int64_t, ServeDone> done_;
    std::deque<int64_t> newly_done_;
    int64_t next_ticket_ = 0, finished_ = 0, rounds_ = 0, inst_iters_ = 0;
    double iterate_ms_ = 0.0, admit_ms_ = 0.0, activate_ms_ = 0.0, retire_ms_ = 0.0;
    double max_step_ms_ = 0.0, query_ms_ = 0.0, wait_ms_ = 0.0;

public:
    using OOM = OutOfMemory;
};

}  // namespace mpx

using namespace mpx;

extern "C" {

int mpx_pdlp_solve_batch_shared(const mpx_lp* base, int32_t count, const double* col_lower,
                                const double* col_upper, const mpx_pdlp_options* opts,
                                double* x_out, double* y_out, mpx_pdlp_result* results) {
    return mpx_pdlp_solve_batch_shared_warm(base, count, col_lower, col_upper, nullptr, nullptr,
                                            opts, x_out, y_out, results);
}

int mpx_pdlp_solve_batch_shared_warm(const mpx_lp* base, int32_t count, const double* col_lower,
                                     const double* col_upper, const double* warm_x,
                                     const double* warm_y, const mpx_pdlp_options* opts,
                                     double* x_out, double* y_out, mpx_pdlp_result* results) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!validLpB(base) || count < 0 || !results || (col_lower == nullptr) != (col_upper == nullptr) ||
        (warm_x == nullptr) != (warm_y == nullptr)) {
        setError("mpx_pdlp_solve_batch_shared: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (!batchOptionsOk(opts)) {
        return MPX_ERR_UNSUPPORTED;
    }
    const int n = base->num_cols, m = base->num_rows;
    try {
        BatchEngine eng(*opts);
        eng.allowInterleaved(true);
        const int ow = eng.addOwner(*base);
        for (int i = 0; i < count; ++i)
            eng.addInstance(ow, col_lower ? col_lower + static_cast<size_t>(i) * n : nullptr,
                            col_upper ? col_upper + static_cast<size_t>(i) * n : nullptr,
                            warm_x ? warm_x + static_cast<size_t>(i) * n : nullptr,
                            warm_y ? warm_y + static_cast<size_t>(i) * m : nullptr);
        eng.setup();
        eng.run(t0);
        std::vector<double*> xs(static_cast<size_t>(count)), ys(static_cast<size_t>(count));
        for (int i = 0; i < count; ++i) {
            xs[i] = x_out ? x_out + static_cast<size_t>(i) * n : nullptr;
            ys[i] = y_out ? y_out + static_cast<size_t>(i) * m : nullptr;
        }
        eng.results(xs.data(), ys.data(), results);
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < count; ++i) results[i].seconds = secs;
    } catch (const DeviceError&) {
        for (int i = 0; i < count; ++i) results[i].status = MPX_STATUS_ERROR;
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_pdlp_solve_batch(const mpx_lp* lps, int32_t count, const mpx_pdlp_options* opts,
                         double* const* x_out, double* const* y_out, mpx_pdlp_result* results) {
    const auto t0 = std::chrono::steady_clock::now();
    if (count < 0 || (count > 0 && (!lps || !results))) {
        setError("mpx_pdlp_solve_batch: invalid arguments");
        return MPX_ERR_INVALID;
    }
    for (int i = 0; i < count; ++i)
        if (!validLpB(&lps[i])) {
            setError("mpx_pdlp_solve_batch: invalid LP " + std::to_string(i) +
                     " (batched LPs need m > 0 and n > 0)");
            return MPX_ERR_INVALID;
        }
    if (!batchOptionsOk(opts)) return MPX_ERR_UNSUPPORTED;
    try {
        BatchEngine eng(*opts);
        for (int i = 0; i < count; ++i) eng.addInstance(eng.addOwner(lps[i]), nullptr, nullptr);
        eng.setup();
        eng.run(t0);
        eng.results(x_out, y_out, results);
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < count; ++i) results[i].seconds = secs;
    } catch (const DeviceError&) {
        for (int i = 0; i < count; ++i) results[i].status = MPX_STATUS_ERROR;
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

}  // extern "C"

struct mpx_server {
    std::unique_ptr<mpx::Server> s;
    bool broken = false;
};

struct mpx_batch {
    mpx_pdlp_options opts;
    std::vector<mpx_lp> lps;                 // descriptors (objective copied)
    std::vector<std::vector<double>> objs;
    std::unique_ptr<mpx::BatchEngine> eng;
    mpx_batch_stats stats{};
};

namespace {

int createBatch(mpx_batch* b) {
    try {
        b->eng->setup();
        b->stats.upload_ms = 1000.0 * b->eng->uploadSeconds();
        b->stats.setup_ms = 1000.0 * b->eng->setupSeconds();
        b->stats.interleaved = b->eng->interleaved() ? 1 : 0;
    } catch (const mpx::DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

void copyLps(mpx_batch* b, const mpx_lp* lps, int count) {
    b->lps.assign(lps, lps + count);
    b->objs.resize(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        b->objs[i].assign(lps[i].obj, lps[i].obj + lps[i].num_cols);
        b->lps[i].obj = b->objs[i].data();
    }
}

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size()))) - 1;
    return v[std::min(v.size() - 1, k)];
}

}  // namespace

extern "C" {

int mpx_pdlp_scale_problem_batched(const mpx_lp* lp, const mpx_pdlp_options* opts,
                                   mpx_scaled_problem* out) {
    if (!validLpB(lp) || !out) {
        setError("mpx_pdlp_scale_problem_batched: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (!batchOptionsOk(opts)) return MPX_ERR_UNSUPPORTED;
    try {
        BatchEngine eng(*opts);
        eng.addInstance(eng.addOwner(*lp), nullptr, nullptr);
        eng.setup();
        eng.downloadOwner(0, out);
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_batch_create_shared(const mpx_lp* base, int32_t count, const double* col_lower,
                            const double* col_upper, const mpx_pdlp_options* opts,
                            mpx_batch** out) {
    if (!out || !validLpB(base) || count <= 0 || (col_lower == nullptr) != (col_upper == nullptr)) {
        setError("mpx_batch_create_shared: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (!batchOptionsOk(opts)) return MPX_ERR_UNSUPPORTED;
    *out = nullptr;
    auto b = std::make_unique<mpx_batch>();
    b->opts = *opts;
    copyLps(b.get(), base, 1);
    try {
        b->eng = std::make_unique<BatchEngine>(b->opts);
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    b->eng->allowInterleaved(true);
    const int ow = b->eng->addOwner(b->lps[0]);
    const size_t n = static_cast<size_t>(base->num_cols);
    for (int i = 0; i < count; ++i)
        b->eng->addInstance(ow, col_lower ? col_lower + i * n : nullptr,
                            col_upper ? col_upper + i * n : nullptr);
    const int rc = createBatch(b.get());
    if (rc != MPX_OK) return rc;
    *out = b.release();
    return MPX_OK;
}

int mpx_batch_create(const mpx_lp* lps, int32_t count, const mpx_pdlp_options* opts,
                     mpx_batch** out) {
    if (!out || count <= 0 || !lps) {
        setError("mpx_batch_create: invalid arguments");
        return MPX_ERR_INVALID;
    }
    for (int i = 0; i < count; ++i)
        if (!validLpB(&lps[i])) {
            setError("mpx_batch_create: invalid LP " + std::to_string(i));
            return MPX_ERR_INVALID;
        }
    if (!batchOptionsOk(opts)) return MPX_ERR_UNSUPPORTED;
    *out = nullptr;
    auto b = std::make_unique<mpx_batch>();
    b->opts = *opts;
    copyLps(b.get(), lps, count);
    try {
        b->eng = std::make_unique<BatchEngine>(b->opts);
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    for (int i = 0; i < count; ++i) b->eng->addInstance(b->eng->addOwner(b->lps[i]), nullptr, nullptr);
    const int rc = createBatch(b.get());
    if (rc != MPX_OK) return rc;
    *out = b.release();
    return MPX_OK;
}

int mpx_batch_run(mpx_batch* b) {
    if (!b) return MPX_ERR_INVALID;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        if (b->stats.rounds > 0 || b->eng->rounds() > 0) b->eng->prepare();
        b->eng->run(t0);
        b->stats.run_ms =
            1000.0 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        b->stats.setup_ms = 1000.0 * b->eng->setupSeconds();
        b->stats.rounds = b->eng->rounds();
        const std::vector<double> ft = b->eng->finishTimesMs();
        b->stats.latency_p50_ms = pct(ft, 0.50);
        b->stats.latency_p99_ms = pct(ft, 0.99);
        b->stats.latency_max_ms = pct(ft, 1.0);
        b->stats.columns_per_wave = b->eng->tunedColumnsPerWave();
        b->stats.rows_per_wave = b->eng->tunedRowsPerWave();
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_batch_results(mpx_batch* b, double* const* x_out, double* const* y_out,
                      mpx_pdlp_result* results) {
    if (!b || !results) return MPX_ERR_INVALID;
    try {
        b->eng->results(x_out, y_out, results);
        int64_t it = 0;
        for (int i = 0; i < b->eng->count(); ++i) {
            it += results[i].iterations;
            results[i].seconds = b->stats.run_ms / 1000.0;
        }
        b->stats.instance_iterations = it;
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_batch_get_stats(mpx_batch* b, mpx_batch_stats* stats) {
    if (!b || !stats) return MPX_ERR_INVALID;
    *stats = b->stats;
    return MPX_OK;
}

static int profileBatch(mpx_batch* b, int32_t launches, mpx_batch_stats* stats, bool refine) {
    if (!b || launches <= 0) return MPX_ERR_INVALID;
    if (refine && b->opts.precision != MPX_PREC_MIXED) {
        mpx::setError("mpx_batch_profile_refine: the batch is not mixed precision");
        return MPX_ERR_INVALID;
    }
    try {
        if (b->eng->rounds() > 0 || refine) b->eng->prepare();  // fresh after create otherwise
        double p = 0.0, d = 0.0;
        int slots = 0;
        b->eng->profileRound(launches, &p, &d, &slots, refine);
        b->stats.primal_kernel_us = p;
        b->stats.dual_kernel_us = d;
        b->stats.profiled_slots = slots;
        b->stats.profiled_launches = std::min(launches, b->opts.termination_eval_frequency);
        b->stats.columns_per_wave = b->eng->tunedColumnsPerWave();
        b->stats.rows_per_wave = b->eng->tunedRowsPerWave();
        b->stats.rounds = 1;  // state is mid-run: the next run re-prepares
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    if (stats) *stats = b->stats;
    return MPX_OK;
}

int mpx_batch_profile(mpx_batch* b, int32_t launches, mpx_batch_stats* stats) {
    return profileBatch(b, launches, stats, false);
}

int mpx_batch_profile_refine(mpx_batch* b, int32_t launches, mpx_batch_stats* stats) {
    return profileBatch(b, launches, stats, true);
}

int mpx_batch_destroy(mpx_batch* b) {
    delete b;
    return MPX_OK;
}

int mpx_server_create(const mpx_pdlp_options* opts, double hbm_budget_bytes, int32_t max_slots,
                      mpx_server** out) {
    if (!out) {
        setError("mpx_server_create: invalid arguments");
        return MPX_ERR_INVALID;
    }
    *out = nullptr;
    if (!batchOptionsOk(opts)) return MPX_ERR_UNSUPPORTED;
    try {
        auto s = std::make_unique<mpx_server>();
        s->s = std::make_unique<Server>(*opts, hbm_budget_bytes, max_slots);
        *out = s.release();
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_server_add_base(mpx_server* srv, const mpx_lp* base, int32_t* base_id) {
    if (!srv || !base_id || !validLpB(base)) {
        setError("mpx_server_add_base: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (srv->broken) return MPX_ERR_DEVICE;
    if (!ilFits(base->num_cols, base->num_rows)) {
        setError("mpx_server_add_base: a base LP has at most 4194303 rows and columns");
        return MPX_ERR_UNSUPPORTED;
    }
    try {
        *base_id = srv->s->addBase(*base);
    } catch (const Server::OOM&) {
        return MPX_ERR_OUT_OF_MEMORY;
    } catch (const DeviceError&) {
        srv->broken = true;
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_server_submit(mpx_server* srv, const mpx_lp* lp, int64_t* ticket) {
    if (!srv || !ticket || !validLpB(lp)) {
        setError("mpx_server_submit: invalid arguments (served LPs need m > 0 and n > 0)");
        return MPX_ERR_INVALID;
    }
    if (srv->broken) return MPX_ERR_DEVICE;
    *ticket = srv->s->submit(*lp);
    return MPX_OK;
}

int mpx_server_submit_bounds(mpx_server* srv, int32_t base_id, const double* col_lower,
                             const double* col_upper, int64_t* ticket) {
    if (!srv || !ticket || !col_lower || !col_upper || base_id < 0 ||
        base_id >= srv->s->numBases()) {
        setError("mpx_server_submit_bounds: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (srv->broken) return MPX_ERR_DEVICE;
    *ticket = srv->s->submitBounds(base_id, col_lower, col_upper);
    return MPX_OK;
}

int mpx_server_submit_bounds_warm(mpx_server* srv, int32_t base_id, const double* col_lower,
                                  const double* col_upper, const double* warm_x,
                                  const double* warm_y, int64_t* ticket) {
    if (!srv || !ticket || !col_lower || !col_upper || !warm_x || !warm_y || base_id < 0 ||
        base_id >= srv->s->numBases()) {
        setError("mpx_server_submit_bounds_warm: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (srv->broken) return MPX_ERR_DEVICE;
    *ticket = srv->s->submitBounds(base_id, col_lower, col_upper, warm_x, warm_y);
    return MPX_OK;
}

int mpx_server_step(mpx_server* srv, int32_t* active) {
    if (!srv) return MPX_ERR_INVALID;
    if (srv->broken) return MPX_ERR_DEVICE;
    try {
        const int a = srv->s->step();
        if (active) *active = a;
    } catch (const DeviceError&) {
        srv->broken = true;
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_server_result(mpx_server* srv, int64_t ticket, int32_t* done, double* x_out,
                      double* y_out, mpx_pdlp_result* res) {
    if (!srv || !done) return MPX_ERR_INVALID;
    const int r = srv->s->result(ticket, x_out, y_out, res);
    if (r < 0) {
        setError("mpx_server_result: unknown or already fetched ticket");
        return MPX_ERR_INVALID;
    }
    *done = r;
    return MPX_OK;
}

int mpx_server_poll(mpx_server* srv, int64_t* tickets, int32_t capacity, int32_t* count) {
    if (!srv || !count || capacity < 0 || (capacity > 0 && !tickets)) return MPX_ERR_INVALID;
    *count = srv->s->poll(tickets, capacity);
    return MPX_OK;
}

int mpx_server_get_stats(mpx_server* srv, mpx_server_stats* stats) {
    if (!srv || !stats) return MPX_ERR_INVALID;
    srv->s->stats(stats);
    return MPX_OK;
}

int mpx_server_destroy(mpx_server* srv) {
    delete srv;
    return MPX_OK;
}

}  // extern "C"
