// batch_setup.h -- device setup of a cohort of distinct matrices ("owners"):
// Ruiz equilibration, objective / bound scaling, Pock-Chambolle bases,
// bound / objective rescaling, termination norms, FP32 copies
// (src/lp/pdlp.cpp:469-669, 53-81) and the power iteration for the step size
// (src/lp/pdlp.cpp:748-795).
//
// Every pass is spread over the whole GPU: it walks a tile list of
// (owner, first row / column / entry), one thread per row, column or entry
// (the SpMVs of the power iteration use the owner's lane-group preset);
// passes that reduce run one workgroup per tile with the reduction at the
// top level of the kernel.
// Per-element arithmetic is the reference's (each row's or column's entries
// in storage order).  Reductions are two-level in a fixed order: a partial
// per tile (fixed tree inside the workgroup), then one workgroup per owner
// over its own tiles.  Tiles never straddle owners, so an owner's results do
// not depend on which other owners share the cohort.
//
// The power iteration runs in groups of kPowerGroup steps; every kernel
// skips owners that have stopped, so extra steps change nothing.  The host
// checks the number of unfinished owners after each group.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "batch_kernels.h"
#include "device_common.h"

namespace mpx {

struct SuTile {
    int32_t o;      // owner (index in the cohort)
    int32_t first;  // first row / column / entry
};

// Tile ranges of one owner in each list.
struct SuRange {
    int32_t row0, nrow, col0, ncol, prow0, nprow, pcol0, npcol;
};

enum SuStatus { kPwRun = 0, kPwLast = 1, kPwDone = 2 };

struct SuPower {
    double s2, xn;
    int32_t it, status, active, pad;
};

constexpr int kSuEntries = 4 * kBlock;  // entries per entry tile
constexpr int kPowerGroup = 16;         // power steps between host checks

struct SuArgs {
    BOwner* owners;
    const int* presets;
    const SuRange* ranges;
    SuPower* pw;
    const SuTile *rowT, *colT, *entT, *prowT, *pcolT;
    int nrowT, ncolT, nentT, nprowT, npcolT;
    double* part;  // [4][maxT] per-tile partials
    int maxT;
    int* remaining;
};

template <typename F>
__device__ inline void suTiles(const SuTile* __restrict__ tiles, int cnt, F&& f) {
    for (int t = blockIdx.x; t < cnt; t += gridDim.x) f(t, tiles[t]);
}

// val = val_o, row and column scales = 1, power iterate = start vector.
__global__ __launch_bounds__(kBlock) void kSuInit(SuArgs a) {
    suTiles(a.entT, a.nentT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int e = min(w.nnz, t.first + kSuEntries);
        for (int k = t.first + static_cast<int>(threadIdx.x); k < e; k += kBlock)
            w.val[k] = w.val_o[k];
    });
    suTiles(a.rowT, a.nrowT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int i = t.first + static_cast<int>(threadIdx.x);
        if (i < w.m) w.row_scale[i] = 1.0;
    });
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (j < w.n) {
            w.col_scale[j] = 1.0;
            w.px[j] = w.px0[j];
        }
    });
}

__device__ inline bool suRuiz(const BOwner& w) { return w.do_ruiz && w.nnz > 0; }

// Ruiz, row half: 1/sqrt(row inf-norm), skipped at <= 1e-12 (pdlp.cpp:520-557).
__global__ __launch_bounds__(kBlock) void kSuRuizRows(SuArgs a) {
    suTiles(a.rowT, a.nrowT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int i = t.first + static_cast<int>(threadIdx.x);
        if (!suRuiz(w) || i >= w.m) return;
        double rmax = 0.0;
        for (int k = w.rs[i]; k < w.rs[i + 1]; ++k) rmax = maxRef(rmax, fabs(w.val[k]));
        if (rmax <= 1e-12) return;
        const double f = 1.0 / sqrt(rmax);
        w.row_scale[i] *= f;
        for (int k = w.rs[i]; k < w.rs[i + 1]; ++k) w.val[k] *= f;
    });
}

// Ruiz, column half: inf-norm of each column (through the CSC permutation).
__global__ __launch_bounds__(kBlock) void kSuRuizCols(SuArgs a) {
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (!suRuiz(w) || j >= w.n) return;
        double mx = 0.0;
        for (int q = w.cs[j]; q < w.cs[j + 1]; ++q) mx = maxRef(mx, fabs(w.val[w.perm[q]]));
        w.cmax[j] = mx;
        if (mx > 1e-12) w.col_scale[j] *= 1.0 / sqrt(mx);
    });
}

__global__ __launch_bounds__(kBlock) void kSuRuizApply(SuArgs a) {
    suTiles(a.entT, a.nentT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        if (!suRuiz(w)) return;
        const int e = min(w.nnz, t.first + kSuEntries);
        for (int k = t.first + static_cast<int>(threadIdx.x); k < e; k += kBlock) {
            const double mx = w.cmax[w.ci[k]];
            if (mx > 1e-12) w.val[k] *= 1.0 / sqrt(mx);
        }
    });
}

// Scaled objective and bounds (pdlp.cpp:560-573), A' values, unit PC bases.
__global__ __launch_bounds__(kBlock) void kSuVectors(SuArgs a) {
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (j >= w.n) return;
        const double sgn = w.sense_max ? -1.0 : 1.0;
        const double cj = sgn * w.obj_o[j];
        w.c[j] = cj * w.col_scale[j];
        const double lj = w.lo_o[j], uj = w.hi_o[j];
        w.l[j] = isfinite(lj) ? lj / w.col_scale[j] : lj;
        w.u[j] = isfinite(uj) ? uj / w.col_scale[j] : uj;
        w.tau[j] = 1.0;
    });
    suTiles(a.rowT, a.nrowT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int i = t.first + static_cast<int>(threadIdx.x);
        if (i >= w.m) return;
        const double lo = w.rl_o[i], hi = w.ru_o[i];
        w.rl[i] = isfinite(lo) ? lo * w.row_scale[i] : lo;
        w.ru[i] = isfinite(hi) ? hi * w.row_scale[i] : hi;
        w.sig[i] = 1.0;
    });
    suTiles(a.entT, a.nentT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int e = min(w.nnz, t.first + kSuEntries);
        for (int k = t.first + static_cast<int>(threadIdx.x); k < e; k += kBlock)
            w.valt[k] = w.val[w.perm[k]];
    });
}

// Pock-Chambolle bases: 1 / max(||row||_2, 1e-8), same for columns
// (pdlp.cpp:581-604).
__global__ __launch_bounds__(kBlock) void kSuPockChambolle(SuArgs a) {
    suTiles(a.rowT, a.nrowT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int i = t.first + static_cast<int>(threadIdx.x);
        if (!w.do_pc || i >= w.m) return;
        double s = 0.0;
        for (int k = w.rs[i]; k < w.rs[i + 1]; ++k) s += w.val[k] * w.val[k];
        w.sig[i] = 1.0 / maxRef(sqrt(s), 1e-8);
    });
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (!w.do_pc || j >= w.n) return;
        double s = 0.0;
        for (int q = w.cs[j]; q < w.cs[j + 1]; ++q) s += w.valt[q] * w.valt[q];
        w.tau[j] = 1.0 / maxRef(sqrt(s), 1e-8);
    });
}

// Tile partials: [0] sum c^2, [1] sum of finite scaled row bounds squared,
// [2] objective norm term, [3] rhs norm term (sum of squares, or max for the
// infinity norm; pdlp.cpp:53-81, 606-640).  One workgroup per tile: column
// tiles first, then row tiles; the reduction runs at the top level of the
// kernel with both the sum and the maximum formed, the norm picks one.
__global__ __launch_bounds__(kBlock) void kSuNormParts(SuArgs a) {
    const int b = blockIdx.x;
    const bool colTile = b < a.ncolT;
    const int tix = colTile ? b : b - a.ncolT;
    if (!colTile && tix >= a.nrowT) return;
    const SuTile t = colTile ? a.colT[tix] : a.rowT[tix];
    const BOwner& w = a.owners[t.o];
    const int e = t.first + static_cast<int>(threadIdx.x);
    double v[3] = {0.0, 0.0, 0.0};  // squares for os/cs, norm sum, norm max
    if (colTile) {
        if (e < w.n) {
            v[0] = w.c[e] * w.c[e];
            v[1] = w.obj_o[e] * w.obj_o[e];
            v[2] = fabs(w.obj_o[e]);
        }
    } else if (e < w.m) {
        if (isfinite(w.rl[e])) v[0] += w.rl[e] * w.rl[e];
        if (isfinite(w.ru[e])) v[0] += w.ru[e] * w.ru[e];
        const double lo = w.rl_o[e], hi = w.ru_o[e];
        if (isfinite(hi)) {
            v[1] += hi * hi;
            v[2] = maxRef(v[2], fabs(hi));
        }
        if (isfinite(lo) && lo != hi) {
            v[1] += lo * lo;
            v[2] = maxRef(v[2], fabs(lo));
        }
    }
    const bool isMax[3] = {false, false, true};
    double out[3];
    blockReduce<3>(v, isMax, out);
    if (threadIdx.x == 0) {
        const size_t T = static_cast<size_t>(a.maxT);
        a.part[(colTile ? 0 : 1) * T + tix] = out[0];
        a.part[(colTile ? 2 : 3) * T + tix] = w.norm_inf ? out[2] : out[1];
    }
}

// os = 1/(||c~|| + 1), cs = 1/(||b~|| + 1) (pdlp.cpp:606-640) and the
// original-problem termination norms; one workgroup per owner, its tiles in
// order.
__global__ __launch_bounds__(kBlock) void kSuNorms(SuArgs a) {
    const int o = blockIdx.x;
    const BOwner& w = a.owners[o];
    const SuRange r = a.ranges[o];
    const size_t T = static_cast<size_t>(a.maxT);
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < r.ncol; t += kBlock) {
        v[0] += a.part[0 * T + r.col0 + t];
        v[2] += a.part[2 * T + r.col0 + t];
        v[3] = maxRef(v[3], a.part[2 * T + r.col0 + t]);
    }
    for (int t = threadIdx.x; t < r.nrow; t += kBlock) {
        v[1] += a.part[1 * T + r.row0 + t];
        v[4] += a.part[3 * T + r.row0 + t];
        v[5] = maxRef(v[5], a.part[3 * T + r.row0 + t]);
    }
    const bool isMax[6] = {false, false, false, true, false, true};
    double out[6];
    blockReduce<6>(v, isMax, out);
    if (threadIdx.x == 0) {
        double os = 1.0, cs = 1.0;
        if (w.do_rescale) {
            os = 1.0 / (sqrt(out[0]) + 1.0);
            cs = 1.0 / (sqrt(out[1]) + 1.0);
        }
        w.scal[0] = os;
        w.scal[1] = cs;
        w.scal[2] = w.norm_inf ? out[3] : sqrt(out[2]);
        w.scal[3] = w.norm_inf ? out[5] : sqrt(out[4]);
    }
}

// Apply os / cs, then FP32 copies of the iteration data.
__global__ __launch_bounds__(kBlock) void kSuRescale(SuArgs a) {
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (j >= w.n) return;
        if (w.do_rescale) {
            const double os = w.scal[0], cs = w.scal[1];
            w.c[j] *= os;
            if (isfinite(w.l[j])) w.l[j] *= cs;
            if (isfinite(w.u[j])) w.u[j] *= cs;
        }
        w.cf[j] = static_cast<float>(w.c[j]);
        w.tauf[j] = static_cast<float>(w.tau[j]);
    });
    suTiles(a.rowT, a.nrowT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int i = t.first + static_cast<int>(threadIdx.x);
        if (i >= w.m) return;
        if (w.do_rescale) {
            const double cs = w.scal[1];
            if (isfinite(w.rl[i])) w.rl[i] *= cs;
            if (isfinite(w.ru[i])) w.ru[i] *= cs;
        }
        w.sigf[i] = static_cast<float>(w.sig[i]);
        w.rlf[i] = static_cast<float>(w.rl[i]);
        w.ruf[i] = static_cast<float>(w.ru[i]);
    });
    suTiles(a.entT, a.nentT, [&](int, SuTile t) {
        const BOwner& w = a.owners[t.o];
        const int e = min(w.nnz, t.first + kSuEntries);
        for (int k = t.first + static_cast<int>(threadIdx.x); k < e; k += kBlock) {
            w.valf[k] = static_cast<float>(w.val[k]);
            w.valtf[k] = static_cast<float>(w.valt[k]);
        }
    });
}

// ---- power iteration (pdlp.cpp:748-795) ------------------------------------
// Per step: y = A x; x' = A'y; s2 = x.x'; res = ||x' - s2 x||; x = x'/||x'||;
// stop when ||x'|| < 1e-15 (before normalising) or res < sv_tol |s2|
// (after), or after sv_max_iter steps.
__global__ void kPwInit(SuArgs a, int count) {
    const int o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= count) return;
    const BOwner& w = a.owners[o];
    SuPower p{};
    p.s2 = 1.0;
    p.status = (w.sv_max_iter > 0 && w.m > 0 && w.n > 0) ? kPwRun : kPwDone;
    a.pw[o] = p;
}

template <int G>
__device__ inline void pwRow(const BOwner& w, int first) {
    const int i = first + static_cast<int>(threadIdx.x) / G, lane = static_cast<int>(threadIdx.x) % G;
    if (i >= w.m) return;
    const double v = bDot<G>(w.rs, w.ci, w.val, w.px, i, lane);
    if (lane == 0) w.py[i] = v;
}

__global__ __launch_bounds__(kBlock) void kPwRows(SuArgs a) {
    suTiles(a.prowT, a.nprowT, [&](int, SuTile t) {
        if (a.pw[t.o].status != kPwRun) return;
        const BOwner& w = a.owners[t.o];
        switch (presetGR(a.presets[t.o])) {
            case 1: pwRow<1>(w, t.first); break;
            case 2: pwRow<2>(w, t.first); break;
            case 4: pwRow<4>(w, t.first); break;
            case 8: pwRow<8>(w, t.first); break;
            default: pwRow<16>(w, t.first); break;
        }
    });
}

template <int G>
__device__ inline double pwCol(const BOwner& w, int first) {
    const int j = first + static_cast<int>(threadIdx.x) / G, lane = static_cast<int>(threadIdx.x) % G;
    double d = 0.0;
    if (j < w.n) {
        const double v = bDot<G>(w.cs, w.ri, w.valt, w.py, j, lane);
        if (lane == 0) {
            w.pxn[j] = v;
            d = w.px[j] * v;
        }
    }
    return d;
}

// x' = A'y and the tile partials of x.x' (one workgroup per tile).
__global__ __launch_bounds__(kBlock) void kPwCols(SuArgs a) {
    const int tix = blockIdx.x;
    if (tix >= a.npcolT) return;
    const SuTile t = a.pcolT[tix];
    if (a.pw[t.o].status != kPwRun) return;
    const BOwner& w = a.owners[t.o];
    double v[1];
    switch (presetGC(a.presets[t.o])) {
        case 1: v[0] = pwCol<1>(w, t.first); break;
        case 2: v[0] = pwCol<2>(w, t.first); break;
        case 4: v[0] = pwCol<4>(w, t.first); break;
        case 8: v[0] = pwCol<8>(w, t.first); break;
        default: v[0] = pwCol<16>(w, t.first); break;
    }
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) a.part[tix] = out[0];
}

// s2 = x.x' over the owner's tiles in order (one workgroup per owner).
__global__ __launch_bounds__(kBlock) void kPwDot(SuArgs a) {
    const int o = blockIdx.x;
    if (a.pw[o].status != kPwRun) return;
    const SuRange r = a.ranges[o];
    double v[1] = {0.0};
    for (int t = threadIdx.x; t < r.npcol; t += kBlock) v[0] += a.part[r.pcol0 + t];
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) a.pw[o].s2 = out[0];
}

// Tile partials of ||x' - s2 x||^2 and ||x'||^2 (one workgroup per tile).
__global__ __launch_bounds__(kBlock) void kPwResid(SuArgs a) {
    const int tix = blockIdx.x;
    if (tix >= a.ncolT) return;
    const SuTile t = a.colT[tix];
    if (a.pw[t.o].status != kPwRun) return;
    const BOwner& w = a.owners[t.o];
    const double s2 = a.pw[t.o].s2;
    const int j = t.first + static_cast<int>(threadIdx.x);
    double v[2] = {0.0, 0.0};
    if (j < w.n) {
        const double d = w.pxn[j] - s2 * w.px[j];
        v[0] = d * d;
        v[1] = w.pxn[j] * w.pxn[j];
    }
    const bool isMax[2] = {false, false};
    double out[2];
    blockReduce<2>(v, isMax, out);
    if (threadIdx.x == 0) {
        a.part[1 * static_cast<size_t>(a.maxT) + tix] = out[0];
        a.part[2 * static_cast<size_t>(a.maxT) + tix] = out[1];
    }
}

// Residual and norm over the owner's tiles, stopping rules, next state.
__global__ __launch_bounds__(kBlock) void kPwNorm(SuArgs a) {
    const int o = blockIdx.x;
    SuPower& p = a.pw[o];
    __shared__ int s_st;
    if (threadIdx.x == 0) s_st = p.status;
    __syncthreads();
    const int st = s_st;
    if (st != kPwRun) {
        if (threadIdx.x == 0) {
            p.active = 0;
            if (st == kPwLast) p.status = kPwDone;
        }
        return;
    }
    const SuRange r = a.ranges[o];
    const size_t T = static_cast<size_t>(a.maxT);
    double v[2] = {0.0, 0.0};
    for (int t = threadIdx.x; t < r.ncol; t += kBlock) {
        v[0] += a.part[1 * T + r.col0 + t];
        v[1] += a.part[2 * T + r.col0 + t];
    }
    const bool isMax[2] = {false, false};
    double out[2];
    blockReduce<2>(v, isMax, out);
    if (threadIdx.x == 0) {
        const BOwner& w = a.owners[o];
        const double res = sqrt(out[0]), xn = sqrt(out[1]);
        p.it += 1;
        if (xn < 1e-15) {
            p.status = kPwDone;
            p.active = 0;
        } else {
            p.xn = xn;
            p.active = 1;
            p.status = (res < w.sv_tol * fabs(p.s2) || p.it >= w.sv_max_iter) ? kPwLast : kPwRun;
        }
    }
}

__global__ __launch_bounds__(kBlock) void kPwScale(SuArgs a) {
    suTiles(a.colT, a.ncolT, [&](int, SuTile t) {
        const SuPower& p = a.pw[t.o];
        if (!p.active) return;
        const BOwner& w = a.owners[t.o];
        const int j = t.first + static_cast<int>(threadIdx.x);
        if (j < w.n) w.px[j] = w.pxn[j] / p.xn;
    });
}

__global__ __launch_bounds__(kBlock) void kPwCount(SuArgs a, int count) {
    double v[1] = {0.0};
    for (int o = threadIdx.x; o < count; o += kBlock) v[0] += a.pw[o].status != kPwDone ? 1.0 : 0.0;
    const bool isMax[1] = {false};
    double out[1];
    blockReduce<1>(v, isMax, out);
    if (threadIdx.x == 0) *a.remaining = static_cast<int>(out[0]);
}

__global__ void kPwFinish(SuArgs a, int count) {
    const int o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= count) return;
    const BOwner& w = a.owners[o];
    w.scal[4] = a.pw[o].s2;
    w.scal[5] = static_cast<double>(a.pw[o].it);
}

// ---- host side ------------------------------------------------------------------
// Tile lists, ranges and device scratch of one cohort.  build() on the host,
// bytes() to size the scratch, carve() into a caller-provided block, run()
// enqueues everything on `st` and waits for it (it synchronises `st` after
// every group of power steps).
class SetupCohort {
public:
    SetupCohort(std::vector<BOwner> owners, std::vector<int> presets)
        : owners_(std::move(owners)), presets_(std::move(presets)) {
        const int k = count();
        ranges_.resize(static_cast<size_t>(k));
        for (int o = 0; o < k; ++o) {
            const BOwner& w = owners_[o];
            SuRange& r = ranges_[o];
            r.row0 = static_cast<int>(rowT_.size());
            for (int i = 0; i < w.m; i += kBlock) rowT_.push_back({o, i});
            r.nrow = static_cast<int>(rowT_.size()) - r.row0;
            r.col0 = static_cast<int>(colT_.size());
            for (int j = 0; j < w.n; j += kBlock) colT_.push_back({o, j});
            r.ncol = static_cast<int>(colT_.size()) - r.col0;
            for (int e = 0; e < w.nnz; e += kSuEntries) entT_.push_back({o, e});
            const int rpt = kBlock / presetGR(presets_[o]), cpt = kBlock / presetGC(presets_[o]);
            r.prow0 = static_cast<int>(prowT_.size());
            for (int i = 0; i < w.m; i += rpt) prowT_.push_back({o, i});
            r.nprow = static_cast<int>(prowT_.size()) - r.prow0;
            r.pcol0 = static_cast<int>(pcolT_.size());
            for (int j = 0; j < w.n; j += cpt) pcolT_.push_back({o, j});
            r.npcol = static_cast<int>(pcolT_.size()) - r.pcol0;
        }
        maxT_ = std::max<int>({1, static_cast<int>(rowT_.size()), static_cast<int>(colT_.size()),
                               static_cast<int>(pcolT_.size())});
    }
    int count() const { return static_cast<int>(owners_.size()); }

    static size_t padb(size_t count, size_t elt) { return (count * elt + 255) & ~size_t(255); }
    size_t bytes() const {
        const size_t k = static_cast<size_t>(count());
        return padb(k, sizeof(BOwner)) + padb(k, sizeof(int)) + padb(k, sizeof(SuRange)) +
               padb(k, sizeof(SuPower)) + padb(rowT_.size(), sizeof(SuTile)) +
               padb(colT_.size(), sizeof(SuTile)) + padb(entT_.size(), sizeof(SuTile)) +
               padb(prowT_.size(), sizeof(SuTile)) + padb(pcolT_.size(), sizeof(SuTile)) +
               padb(4 * static_cast<size_t>(maxT_), sizeof(double)) + padb(1, sizeof(int)) + 256;
    }
    // `base` must hold bytes() bytes of device memory.
    void carve(char* base) {
        size_t off = 0;
        auto take = [&](size_t count, size_t elt) {
            void* p = base + off;
            off += padb(count, elt);
            return p;
        };
        const size_t k = static_cast<size_t>(count());
        a_.owners = static_cast<BOwner*>(take(k, sizeof(BOwner)));
        a_.presets = static_cast<int*>(take(k, sizeof(int)));
        a_.ranges = static_cast<SuRange*>(take(k, sizeof(SuRange)));
        a_.pw = static_cast<SuPower*>(take(k, sizeof(SuPower)));
        a_.rowT = static_cast<SuTile*>(take(rowT_.size(), sizeof(SuTile)));
        a_.colT = static_cast<SuTile*>(take(colT_.size(), sizeof(SuTile)));
        a_.entT = static_cast<SuTile*>(take(entT_.size(), sizeof(SuTile)));
        a_.prowT = static_cast<SuTile*>(take(prowT_.size(), sizeof(SuTile)));
        a_.pcolT = static_cast<SuTile*>(take(pcolT_.size(), sizeof(SuTile)));
        a_.part = static_cast<double*>(take(4 * static_cast<size_t>(maxT_), sizeof(double)));
        a_.remaining = static_cast<int*>(take(1, sizeof(int)));
        a_.nrowT = static_cast<int>(rowT_.size());
        a_.ncolT = static_cast<int>(colT_.size());
        a_.nentT = static_cast<int>(entT_.size());
        a_.nprowT = static_cast<int>(prowT_.size());
        a_.npcolT = static_cast<int>(pcolT_.size());
        a_.maxT = maxT_;
    }

    // Uploads the cohort tables and runs the whole setup on `st`, launching
    // every kernel directly (sixteen power steps per host check cost about
    // what one graph launch of them does).
    void run(hipStream_t st) {
        const int k = count();
        if (k == 0) return;
        auto up = [&](void* dst, const void* src, size_t bytes) {
            if (bytes) MPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        };
        up(a_.owners, owners_.data(), sizeof(BOwner) * k);
        up(const_cast<int*>(a_.presets), presets_.data(), sizeof(int) * k);
        up(const_cast<SuRange*>(a_.ranges), ranges_.data(), sizeof(SuRange) * k);
        up(const_cast<SuTile*>(a_.rowT), rowT_.data(), sizeof(SuTile) * rowT_.size());
        up(const_cast<SuTile*>(a_.colT), colT_.data(), sizeof(SuTile) * colT_.size());
        up(const_cast<SuTile*>(a_.entT), entT_.data(), sizeof(SuTile) * entT_.size());
        up(const_cast<SuTile*>(a_.prowT), prowT_.data(), sizeof(SuTile) * prowT_.size());
        up(const_cast<SuTile*>(a_.pcolT), pcolT_.data(), sizeof(SuTile) * pcolT_.size());
        auto grid = [](int tiles) { return std::max(1, std::min(tiles, 4096)); };
        const int gAll = grid(std::max({a_.nrowT, a_.ncolT, a_.nentT}));
        kSuInit<<<gAll, kBlock, 0, st>>>(a_);
        int ruiz = 0;
        for (const BOwner& w : owners_) ruiz = std::max(ruiz, w.do_ruiz ? w.ruiz_iters : 0);
        for (int it = 0; it < ruiz; ++it) {
            // Owners with fewer Ruiz iterations never occur: one options
            // struct per cohort.
            kSuRuizRows<<<grid(a_.nrowT), kBlock, 0, st>>>(a_);
            kSuRuizCols<<<grid(a_.ncolT), kBlock, 0, st>>>(a_);
            kSuRuizApply<<<grid(a_.nentT), kBlock, 0, st>>>(a_);
        }
        kSuVectors<<<gAll, kBlock, 0, st>>>(a_);
        kSuPockChambolle<<<grid(std::max(a_.nrowT, a_.ncolT)), kBlock, 0, st>>>(a_);
        kSuNormParts<<<std::max(1, a_.ncolT + a_.nrowT), kBlock, 0, st>>>(a_);
        kSuNorms<<<k, kBlock, 0, st>>>(a_);
        kSuRescale<<<gAll, kBlock, 0, st>>>(a_);
        kPwInit<<<ceilDiv(k, kBlock), kBlock, 0, st>>>(a_, k);
        MPX_HIP(hipGetLastError());
        // Power iteration in groups of kPowerGroup steps; the host checks
        // after each.
        SuArgs pa = a_;  // tile lists of the power steps (compacted below)
        int left = 1, running = k;
        int maxIter = 0;
        for (const BOwner& w : owners_) maxIter = std::max(maxIter, w.sv_max_iter);
        std::vector<SuPower> hp;
        for (int done = 0; left > 0 && done <= maxIter + kPowerGroup; done += kPowerGroup) {
            powerGroup(st, k, pa);
            kPwCount<<<1, kBlock, 0, st>>>(pa, k);
            MPX_HIP(hipMemcpyAsync(&left, a_.remaining, sizeof(int), hipMemcpyDeviceToHost, st));
            MPX_HIP(hipStreamSynchronize(st));
            // Once at most half of the listed matrices still iterate, the
            // power steps get tile lists of those alone (same tiles in the
            // same order per matrix, so every matrix's sums are unchanged).
            if (left > 0 && 2 * left <= running) {
                hp.resize(static_cast<size_t>(k));
                MPX_HIP(hipMemcpyAsync(hp.data(), a_.pw, sizeof(SuPower) * k, hipMemcpyDeviceToHost, st));
                MPX_HIP(hipStreamSynchronize(st));
                running = compactPower(st, k, hp, pa);
            }
        }
        kPwFinish<<<ceilDiv(k, kBlock), kBlock, 0, st>>>(a_, k);
        MPX_HIP(hipGetLastError());
        MPX_HIP(hipStreamSynchronize(st));
    }

private:
    void powerGroup(hipStream_t st, int k, const SuArgs& a) {
        auto grid = [](int tiles) { return std::max(1, std::min(tiles, 4096)); };
        for (int s = 0; s < kPowerGroup; ++s) {
            kPwRows<<<grid(a.nprowT), kBlock, 0, st>>>(a);
            kPwCols<<<std::max(1, a.npcolT), kBlock, 0, st>>>(a);
            kPwDot<<<k, kBlock, 0, st>>>(a);
            kPwResid<<<std::max(1, a.ncolT), kBlock, 0, st>>>(a);
            kPwNorm<<<k, kBlock, 0, st>>>(a);
            kPwScale<<<grid(a.ncolT), kBlock, 0, st>>>(a);
        }
    }

    // Rewrites the power steps' tile lists (row-product, column-product and
    // column tiles) and the per-matrix ranges for the matrices still in
    // kPwRun, in cohort order; the full lists are uploaded again by the next
    // run().  Returns the number of those matrices.
    int compactPower(hipStream_t st, int k, const std::vector<SuPower>& hp, SuArgs& pa) {
        std::vector<SuTile> prow, pcol, col;
        std::vector<SuRange> rng = ranges_;
        int running = 0;
        for (int o = 0; o < k; ++o) {
            if (hp[static_cast<size_t>(o)].status != kPwRun) continue;
            ++running;
            SuRange& r = rng[static_cast<size_t>(o)];
            const SuRange& f = ranges_[static_cast<size_t>(o)];
            r.prow0 = static_cast<int32_t>(prow.size());
            prow.insert(prow.end(), prowT_.begin() + f.prow0, prowT_.begin() + f.prow0 + f.nprow);
            r.pcol0 = static_cast<int32_t>(pcol.size());
            pcol.insert(pcol.end(), pcolT_.begin() + f.pcol0, pcolT_.begin() + f.pcol0 + f.npcol);
            r.col0 = static_cast<int32_t>(col.size());
            col.insert(col.end(), colT_.begin() + f.col0, colT_.begin() + f.col0 + f.ncol);
        }
        auto up = [&](const void* dst, const void* src, size_t bytes) {
            if (bytes)
                MPX_HIP(hipMemcpyAsync(const_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice, st));
        };
        up(pa.prowT, prow.data(), sizeof(SuTile) * prow.size());
        up(pa.pcolT, pcol.data(), sizeof(SuTile) * pcol.size());
        up(pa.colT, col.data(), sizeof(SuTile) * col.size());
        up(pa.ranges, rng.data(), sizeof(SuRange) * rng.size());
        // The copies come from pageable vectors that end with this call.
        MPX_HIP(hipStreamSynchronize(st));
        pa.nprowT = static_cast<int>(prow.size());
        pa.npcolT = static_cast<int>(pcol.size());
        pa.ncolT = static_cast<int>(col.size());
        return running;
    }

    std::vector<BOwner> owners_;
    std::vector<int> presets_;
    std::vector<SuRange> ranges_;
    std::vector<SuTile> rowT_, colT_, entT_, prowT_, pcolT_;
    int maxT_ = 1;
    SuArgs a_{};
};

}  // namespace mpx
