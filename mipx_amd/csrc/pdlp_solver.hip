// pdlp_solver.hip -- single-LP PDLP on one MI355X, and the C-ABI entry points.
//
// mpx_pdlp_solve() replaces the reference's PdlpSolver::solveGpu()
// (src/lp/pdlp.cpp:1739-2792).  Flow:
//   1. upload the raw CSR problem once; the CSC structure (A' in CSR) and the
//      CSR->CSC permutation are built on the host (index work only);
//   2. scaling on the device (pdlp_setup.h), bit-identical per element to the
//      reference host code;
//   3. power iteration on the device with one host check per captured group
//      of steps (the reference syncs three times per step, pdlp.cpp:1988-2041);
//   4. Halpern PDHG in hipGraph-captured blocks of N = termination_eval_frequency
//      iterations (2 fused kernels per iteration); step, primal weight and the
//      inner count live in device memory so one graph survives restarts
//      (pdlp.cpp:2043-2051,2145-2190);
//   5. after every block a deterministic evaluation (2 fused SpMV + reduction
//      kernels, no atomics) and one 80-byte device->host copy; the restart,
//      PI primal-weight, termination and certificate logic then runs on the
//      host exactly as in the reference (pdlp.cpp:2228-2469);
//   6. polish phase (pdlp.cpp:2477-2676) and recovery (2678-2712).
// MPX_PREC_MIXED runs phase 4 in FP32 (FP32 matrix, iterates and
// accumulation) while every evaluation stays FP64, then refines in FP64.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mipx_pdlp.h"
#include "device_common.h"
#include "lp_filter.h"
#include "pdlp_accel.h"
#include "pdlp_kernels.h"
#include "pdlp_refine.h"
#include "pdlp_setup.h"

namespace mpx {

namespace {
thread_local std::string g_error;
}
void setError(const std::string& msg) { g_error = msg; }
const char* lastError() { return g_error.c_str(); }

template <typename F>
void withGroup(int g, F&& f) {
    switch (g) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        default: f(std::integral_constant<int, 64>{}); break;
    }
}

// Lanes per row from the mean entries per row (autotuned choice when the
// caller leaves spmv_group at 0): about two entries per lane.
int pickGroup(long long nnz, int rows, int forced) {
    if (forced > 0) {
        int g = 1;
        while (g < forced && g < 64) g <<= 1;
        return g;
    }
    if (rows <= 0) return 1;
    const double avg = static_cast<double>(nnz) / rows;
    int g = 1;
    while (g < 64 && 2.0 * g < avg) g <<= 1;
    return g;
}

inline bool isFin(double v) { return std::isfinite(v); }

// ---- device arena ----------------------------------------------------------
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void reserve(size_t bytes) {
        cap = bytes;
        if (bytes) MPX_HIP(hipMalloc(&base, bytes));
    }
    template <typename T>
    T* take(size_t count) {
        const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (used + bytes > cap) {
            setError("internal: device arena overflow");
            throw DeviceError();
        }
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
};

inline size_t padBytes(size_t count, size_t elt) { return (count * elt + 255) & ~size_t(255); }

struct PinnedHost {
    void* p = nullptr;
    explicit PinnedHost(size_t bytes) { MPX_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
    ~PinnedHost() {
        if (p) (void)hipHostFree(p);
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { MPX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
    }
    ~Graph() { reset(); }
};

// Host CSC structure of A plus the CSR->CSC permutation: rows visited in
// increasing order, so each column lists its rows increasing (the order of
// the reference's buildTransposeCSR, pdlp.cpp:671-707).
void buildTransposeStructure(int m, int n, const int* rs, const int* ci, std::vector<int>& cs,
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

// ---- original-space certificate checks (host) ------------------------------
// Same tests as the reference's checkInfeasibilityCertificate /
// checkUnboundednessCertificate (src/lp/pdlp.cpp:119-284), on the original LP.
double l2(const std::vector<double>& v) {
    double s = 0.0;
    for (double e : v) s += e * e;
    return std::sqrt(s);
}

bool farkasHolds(const mpx_lp& lp, const std::vector<double>& y, double tol) {
    const int m = lp.num_rows, n = lp.num_cols;
    if (m == 0) return false;
    const double ny = l2(y);
    if (ny < 1e-15) return false;
    std::vector<double> r(y);
    for (double& v : r) v /= ny;
    for (int i = 0; i < m; ++i) {
        const bool lo = isFin(lp.row_lower[i]), hi = isFin(lp.row_upper[i]);
        if (hi && !lo && r[i] < -tol) return false;
        if (lo && !hi && r[i] > tol) return false;
    }
    std::vector<double> atr(static_cast<size_t>(n), 0.0);
    for (int i = 0; i < m; ++i) {
        if (r[i] == 0.0) continue;
        for (int k = lp.row_starts[i]; k < lp.row_starts[i + 1]; ++k)
            atr[lp.col_indices[k]] += lp.values[k] * r[i];
    }
    for (int j = 0; j < n; ++j) {
        if (isFin(lp.col_lower[j]) && atr[j] < -tol) return false;
        if (isFin(lp.col_upper[j]) && atr[j] > tol) return false;
    }
    double rowpart = 0.0;
    for (int i = 0; i < m; ++i) {
        const double lo = lp.row_lower[i], hi = lp.row_upper[i];
        if (lo == hi) rowpart += r[i] * lo;
        else if (r[i] > 0.0 && isFin(hi)) rowpart += r[i] * hi;
        else if (r[i] < 0.0 && isFin(lo)) rowpart += r[i] * lo;
    }
    double colpart = 0.0;
    for (int j = 0; j < n; ++j) {
        if (atr[j] > 0.0 && isFin(lp.col_lower[j])) colpart += atr[j] * lp.col_lower[j];
        else if (atr[j] < 0.0 && isFin(lp.col_upper[j])) colpart += atr[j] * lp.col_upper[j];
    }
    return rowpart - colpart < -tol;
}

bool rayHolds(const mpx_lp& lp, const std::vector<double>& x, double tol) {
    const int m = lp.num_rows, n = lp.num_cols;
    if (n == 0) return false;
    const double nx = l2(x);
    if (nx < 1e-15) return false;
    std::vector<double> r(x);
    for (double& v : r) v /= nx;
    for (int j = 0; j < n; ++j) {
        if (isFin(lp.col_lower[j]) && r[j] < -tol) return false;
        if (isFin(lp.col_upper[j]) && r[j] > tol) return false;
    }
    for (int i = 0; i < m; ++i) {
        double ar = 0.0;
        for (int k = lp.row_starts[i]; k < lp.row_starts[i + 1]; ++k)
            ar += lp.values[k] * r[lp.col_indices[k]];
        const bool lo = isFin(lp.row_lower[i]), hi = isFin(lp.row_upper[i]);
        if (lo && hi) {
            if (std::abs(ar) > tol) return false;
        } else if (hi) {
            if (ar > tol) return false;
        } else if (lo) {
            if (ar < -tol) return false;
        }
    }
    const double sgn = lp.sense ? -1.0 : 1.0;
    double cr = 0.0;
    for (int j = 0; j < n; ++j) cr += lp.obj[j] * r[j];
    return sgn * cr < -tol;
}

// ---- the solver ------------------------------------------------------------
// Largest Anderson window accepted (the Gram pass launches one grid row per
// pair, w (w + 1) / 2 of them).
constexpr int kMaxAndersonWindow = 64;

struct EvalOut {
    double fpe, rel_p, rel_d, rel_p_inf, rel_d_inf, gap, pobj, dobj, dx2, dy2;
};

template <typename T>
struct IterVecs {
    T *x = nullptr, *x0 = nullptr, *xr = nullptr, *xh = nullptr;
    T *y = nullptr, *y0 = nullptr, *yh = nullptr;
    T *val = nullptr, *valt = nullptr;
    T *c = nullptr, *tau = nullptr, *l = nullptr, *u = nullptr;
    T *sig = nullptr, *rl = nullptr, *ru = nullptr;
};

class SingleSolver {
public:
    // `orig` (default: `lp`) is the problem as given, for the certificate
    // checks, when `lp` is its copy without near-zero entries.
    SingleSolver(const mpx_lp& lp, const mpx_pdlp_options& o, const mpx_lp* orig = nullptr)
        : lp_(lp), orig_(orig ? *orig : lp), o_(o) {
        m_ = lp.num_rows;
        n_ = lp.num_cols;
        nnz_ = lp.row_starts[m_];
        sgn_ = lp.sense ? -1.0 : 1.0;
        mixed_ = o.precision == MPX_PREC_MIXED;
        gr_ = pickGroup(nnz_, m_, o.spmv_group);
        gc_ = pickGroup(nnz_, n_, o.spmv_group);
        adaptive_ = o.adaptive_step != 0;
        aa_w_ = (o.anderson_acceleration && o.anderson_window_size >= 2) ? o.anderson_window_size : 0;
    }

    // Device setup: arena, uploads, scaling, power iteration.
    void setup(const double* warm_x, const double* warm_y);
    // Scaled problem download (mpx_pdlp_scale_problem).
    void download(mpx_scaled_problem* out);
    // Full solve.
    void solve(const std::chrono::steady_clock::time_point t0, double* x_out, double* y_out,
               mpx_pdlp_result* res);
    // Plain iterations from a given state (mpx_pdhg_iterate).
    void iterateFrom(const mpx_scaled_problem* sp, int iters, int inner, double step, double pw,
                     int precision, double* cx, double* ix, double* px, double* cy, double* iy,
                     double* py);

    double step() const { return step_; }
    int powerIters() const { return power_iters_; }

private:
    template <typename T>
    IterVecs<T>& vecs();
    template <typename T>
    void launchIteration(int k, bool storeHat);
    template <typename T>
    void runBlock(int batch, Graph& graph, bool allowGraph);
    template <typename T>
    EvalOut evaluate();
    template <typename T>
    void restartCopy();
    template <typename T>
    void recover(double* x_out, double* y_out, double* objective);
    template <typename T>
    void downloadHats(std::vector<double>& xh, std::vector<double>& yh);
    void pushScalars();
    void powerIteration();
    void convertIterates64to32();
    void convertIterates32to64();
    // FP32 copy of the chunk-major values (mixed precision).
    void convertChunks() {
        if (nch_ && valch32_ && nnz_ > 0)
            kConvert<float, double><<<ceilDiv(nnz_, kBlock), kBlock, 0, stream_.s>>>(nnz_, valch64_,
                                                                                     valch32_);
    }
    template <typename T>
    void adaptStep(double lb, double ub);
    template <typename T>
    bool andersonRestart();

    const mpx_lp& lp_;
    const mpx_lp& orig_;
    const mpx_pdlp_options& o_;
    int m_, n_, nnz_;
    double sgn_;
    bool mixed_;
    int gr_, gc_;
    // Long columns / rows handled one wave each (counts 0 when not used).
    int *long_c_ = nullptr, *long_r_ = nullptr;
    int n_long_c_ = 0, n_long_r_ = 0;
    int skip_c_ = std::numeric_limits<int>::max(), skip_r_ = std::numeric_limits<int>::max();
    Stream stream_;
    Arena arena_;
    // structure and original data
    int *rs_ = nullptr, *ci_ = nullptr, *cs_ = nullptr, *ri_ = nullptr, *perm_ = nullptr;
    double *obj_o_ = nullptr, *l_o_ = nullptr, *u_o_ = nullptr, *rl_o_ = nullptr, *ru_o_ = nullptr;
    double *row_scale_ = nullptr, *col_scale_ = nullptr, *cmax_ = nullptr;
    double* scales_ = nullptr;  // [os, cs]
    double* sums_ = nullptr;
    double *pcol_ = nullptr, *prow_ = nullptr, *metrics_ = nullptr, *red_ = nullptr;
    int nbc_ = 1, nbr_ = 1, nbred_ = 1;
    IterVecs<double> v64_;
    IterVecs<float> v32_;
    IterScalars* sc_dev_ = nullptr;
    IterScalars* sc_host_ = nullptr;
    double* metrics_host_ = nullptr;
    std::unique_ptr<PinnedHost> pinned_;
    double os_ = 1.0, cs_host_ = 1.0;
    double step_ = 1.0;
    int power_iters_ = 0;
    double pw_ = 1.0;
    int inner_ = 0;
    double cnorm_ = 0.0, bnorm_ = 0.0;
    // Adaptive step of the GPU host loop (pdlp.cpp:2271-2307): previous x^
    // and A x^, reduction partials and the two sums.
    bool adaptive_ = false, have_prev_ = false;
    double *prev_ax_ = nullptr, *prev_x_ = nullptr, *acc_part_ = nullptr, *acc_out_ = nullptr;
    int nbn_ = 1;
    // Column chunks of the dual step's SpMV (0: one kDualStep launch): chunk-
    // major CSR (row starts per chunk, then the entries of chunk 0, 1, ...),
    // its position map into the scaled CSR values, and the partial sums.
    int nch_ = 0, gch_ = 1;
    int *rsch_ = nullptr, *cich_ = nullptr, *permch_ = nullptr;
    double *valch64_ = nullptr, *part64_ = nullptr;
    float *valch32_ = nullptr, *part32_ = nullptr;
    // Anderson acceleration (pdlp.cpp:293-437): window of aa_w_ restart
    // candidates and residuals in HBM, aa_count_ stored so far.
    int aa_w_ = 0, aa_count_ = 0, nbaa_ = 1;
    double *aa_z_ = nullptr, *aa_r_ = nullptr, *aa_alpha_ = nullptr;
    std::unique_ptr<PinnedHost> acc_pinned_;
    // FP64 iterative refinement of mixed precision (pdlp_refine.h): while
    // refine_ is set the FP32 iterates are a correction around the FP64 base
    // (x_b in v64_.x, y_b in v64_.y), v32_.c / l / u / rl / ru hold the
    // correction's data and q32_ the row shift.
    bool refine_ = false;
    float* q32_ = nullptr;
    void recenter();
    void materialize();
};

template <>
IterVecs<double>& SingleSolver::vecs<double>() { return v64_; }
template <>
IterVecs<float>& SingleSolver::vecs<float>() { return v32_; }

void SingleSolver::pushScalars() {
    sc_host_->step = step_;
    sc_host_->primal_weight = pw_;
    sc_host_->inner = inner_;
    MPX_HIP(hipMemcpyAsync(sc_dev_, sc_host_, sizeof(IterScalars), hipMemcpyHostToDevice,
                           stream_.s));
}

void SingleSolver::setup(const double* warm_x, const double* warm_y) {
    const size_t m = static_cast<size_t>(m_), n = static_cast<size_t>(n_),
                 nnz = static_cast<size_t>(nnz_);
    nbc_ = std::max(1, ceilDiv(static_cast<long long>(n_) * gc_, kBlock));
    nbr_ = std::max(1, ceilDiv(static_cast<long long>(m_) * gr_, kBlock));
    nbred_ = std::max(1, std::min(1024, ceilDiv(std::max(n_, m_), kBlock)));
    const int nbAt = std::max(1, ceilDiv(static_cast<long long>(n_) * gc_, kBlock));
    // Arena size.
    size_t bytes = 0;
    bytes += padBytes(m + 1, 4) + 3 * padBytes(nnz, 4) + padBytes(n + 1, 4) + padBytes(nnz, 4);
    bytes += 2 * padBytes(nnz, 8);                            // val, valt (fp64)
    bytes += 6 * padBytes(n, 8) + 6 * padBytes(m, 8);         // originals, scales, cmax
    bytes += 7 * padBytes(n, 8) + 6 * padBytes(m, 8);         // fp64 iterate + constants
    bytes += padBytes(16, 8) * 4 + padBytes(2048, 8);         // scales, sums, metrics, red
    bytes += padBytes(static_cast<size_t>(nbc_) * kNumColFields, 8);
    bytes += padBytes(static_cast<size_t>(nbr_) * kNumRowFields, 8);
    bytes += padBytes(static_cast<size_t>(std::max(nbAt, 2 * nbred_)), 8);
    bytes += 3 * padBytes(std::max(n, m), 8);                 // power iteration vectors
    bytes += padBytes(1, sizeof(IterScalars)) + padBytes(1, sizeof(PowerState));
    bytes += padBytes(n, 4) + padBytes(m, 4);                 // long column / row lists
    nbn_ = std::max(1, ceilDiv(n_, kBlock));
    const size_t dim = n + m;
    nbaa_ = std::max(1, std::min(1024, ceilDiv(static_cast<long long>(dim), kBlock)));
    const size_t pairs = static_cast<size_t>(aa_w_) * static_cast<size_t>(aa_w_ + 1) / 2;
    const size_t accParts = std::max({static_cast<size_t>(nbr_), static_cast<size_t>(nbn_),
                                      static_cast<size_t>(nbaa_) * std::max<size_t>(pairs, 1)});
    // Chunked dual SpMV: forced by spmv_chunk_cols > 0, automatic (2^19
    // columns) when x_r is larger than 8 MB in FP64; not with rows long
    // enough for the one-wave row kernel.
    {
        long long cc = o_.spmv_chunk_cols;
        if (cc == 0) cc = n_ > (1 << 20) ? (1 << 19) : -1;
        bool longRows = false;
        if (o_.spmv_group <= 0 && gr_ < kWave) {
            long long inLong = 0;
            for (int i = 0; i < m_; ++i) {
                const int len = lp_.row_starts[i + 1] - lp_.row_starts[i];
                if (len > 4 * gr_) inLong += len;
            }
            longRows = inLong > 0 && 20 * inLong >= static_cast<long long>(nnz_);
        }
        if (cc > 0 && cc < n_ && !longRows && m_ > 0) {
            nch_ = static_cast<int>((n_ + cc - 1) / cc);
            gch_ = pickGroup(nnz_ / nch_, m_, o_.spmv_group);
            bytes += padBytes(static_cast<size_t>(nch_) * (m + 1), 4) + 2 * padBytes(nnz, 4) +
                     padBytes(nnz, 8) + padBytes(m, 8);
            if (mixed_) bytes += padBytes(nnz, 4) + padBytes(m, 4);
        }
    }
    if (adaptive_) bytes += padBytes(m, 8) + padBytes(n, 8);
    if (aa_w_) bytes += 2 * padBytes(static_cast<size_t>(aa_w_) * dim, 8) + padBytes(aa_w_, 8);
    if (adaptive_ || aa_w_) bytes += padBytes(accParts, 8) + padBytes(std::max<size_t>(pairs, 2), 8);
    if (mixed_) {
        bytes += 2 * padBytes(nnz, 4);
        bytes += 8 * padBytes(n, 4) + 7 * padBytes(m, 4);
    }
    bytes += 4096;
    arena_.reserve(bytes);

    rs_ = arena_.take<int>(m + 1);
    ci_ = arena_.take<int>(nnz);
    cs_ = arena_.take<int>(n + 1);
    ri_ = arena_.take<int>(nnz);
    perm_ = arena_.take<int>(nnz);
    v64_.val = arena_.take<double>(nnz);
    v64_.valt = arena_.take<double>(nnz);
    obj_o_ = arena_.take<double>(n);
    l_o_ = arena_.take<double>(n);
    u_o_ = arena_.take<double>(n);
    rl_o_ = arena_.take<double>(m);
    ru_o_ = arena_.take<double>(m);
    row_scale_ = arena_.take<double>(m);
    col_scale_ = arena_.take<double>(n);
    cmax_ = arena_.take<double>(n);
    v64_.x = arena_.take<double>(n);
    v64_.x0 = arena_.take<double>(n);
    v64_.xr = arena_.take<double>(n);
    v64_.xh = arena_.take<double>(n);
    v64_.c = arena_.take<double>(n);
    v64_.tau = arena_.take<double>(n);
    v64_.l = arena_.take<double>(n);
    v64_.u = arena_.take<double>(n);
    v64_.y = arena_.take<double>(m);
    v64_.y0 = arena_.take<double>(m);
    v64_.yh = arena_.take<double>(m);
    v64_.sig = arena_.take<double>(m);
    v64_.rl = arena_.take<double>(m);
    v64_.ru = arena_.take<double>(m);
    scales_ = arena_.take<double>(16);
    sums_ = arena_.take<double>(16);
    metrics_ = arena_.take<double>(16);
    red_ = arena_.take<double>(2 * 1024);
    pcol_ = arena_.take<double>(static_cast<size_t>(nbc_) * kNumColFields);
    prow_ = arena_.take<double>(static_cast<size_t>(nbr_) * kNumRowFields);
    sc_dev_ = arena_.take<IterScalars>(1);
    if (adaptive_) {
        prev_ax_ = arena_.take<double>(m);
        prev_x_ = arena_.take<double>(n);
    }
    if (nch_) {
        rsch_ = arena_.take<int>(static_cast<size_t>(nch_) * (m + 1));
        cich_ = arena_.take<int>(nnz);
        permch_ = arena_.take<int>(nnz);
        valch64_ = arena_.take<double>(nnz);
        part64_ = arena_.take<double>(m);
        if (mixed_) {
            valch32_ = arena_.take<float>(nnz);
            part32_ = arena_.take<float>(m);
        }
    }
    if (aa_w_) {
        aa_z_ = arena_.take<double>(static_cast<size_t>(aa_w_) * dim);
        aa_r_ = arena_.take<double>(static_cast<size_t>(aa_w_) * dim);
        aa_alpha_ = arena_.take<double>(aa_w_);
    }
    if (adaptive_ || aa_w_) {
        acc_part_ = arena_.take<double>(accParts);
        acc_out_ = arena_.take<double>(std::max<size_t>(pairs, 2));
        acc_pinned_ = std::make_unique<PinnedHost>(sizeof(double) * (std::max<size_t>(pairs, 2) + aa_w_));
    }
    if (mixed_) {
        v32_.val = arena_.take<float>(nnz);
        v32_.valt = arena_.take<float>(nnz);
        v32_.x = arena_.take<float>(n);
        v32_.x0 = arena_.take<float>(n);
        v32_.xr = arena_.take<float>(n);
        v32_.xh = arena_.take<float>(n);
        v32_.c = arena_.take<float>(n);
        v32_.tau = arena_.take<float>(n);
        v32_.l = arena_.take<float>(n);
        v32_.u = arena_.take<float>(n);
        v32_.y = arena_.take<float>(m);
        v32_.y0 = arena_.take<float>(m);
        v32_.yh = arena_.take<float>(m);
        v32_.sig = arena_.take<float>(m);
        v32_.rl = arena_.take<float>(m);
        v32_.ru = arena_.take<float>(m);
        q32_ = arena_.take<float>(m);
    }
    pinned_ = std::make_unique<PinnedHost>(sizeof(IterScalars) + 32 * sizeof(double));
    sc_host_ = static_cast<IterScalars*>(pinned_->p);
    metrics_host_ = reinterpret_cast<double*>(static_cast<char*>(pinned_->p) + sizeof(IterScalars));

    hipStream_t st = stream_.s;
    // Host transpose structure.
    std::vector<int> cs, ri, perm;
    buildTransposeStructure(m_, n_, lp_.row_starts, lp_.col_indices, cs, ri, perm);
    // Long columns / rows (more than four entries per lane of their group)
    // go to the one-wave kernels when they carry at least 5% of the entries.
    // Only when the group size is chosen here: a group forced by the caller
    // (spmv_group = 1 is the reference's summation order) is kept for all.
    auto longList = [&](const int* starts, int count, int g, std::vector<int>& list) {
        if (o_.spmv_group > 0) return std::numeric_limits<int>::max();
        const int thresh = 4 * g;
        long long inLong = 0;
        for (int t = 0; t < count; ++t)
            if (starts[t + 1] - starts[t] > thresh) {
                list.push_back(t);
                inLong += starts[t + 1] - starts[t];
            }
        if (g >= kWave || list.empty() || 20 * inLong < static_cast<long long>(nnz_)) {
            list.clear();
            return std::numeric_limits<int>::max();
        }
        return thresh;
    };
    std::vector<int> longC, longR;
    skip_c_ = longList(cs.data(), n_, gc_, longC);
    skip_r_ = longList(lp_.row_starts, m_, gr_, longR);
    n_long_c_ = static_cast<int>(longC.size());
    n_long_r_ = static_cast<int>(longR.size());
    long_c_ = arena_.take<int>(std::max<size_t>(1, longC.size()));
    long_r_ = arena_.take<int>(std::max<size_t>(1, longR.size()));
    if (n_long_c_)
        MPX_HIP(hipMemcpyAsync(long_c_, longC.data(), longC.size() * 4, hipMemcpyHostToDevice, st));
    if (n_long_r_)
        MPX_HIP(hipMemcpyAsync(long_r_, longR.data(), longR.size() * 4, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(rs_, lp_.row_starts, (m + 1) * 4, hipMemcpyHostToDevice, st));
    if (nnz) {
        MPX_HIP(hipMemcpyAsync(ci_, lp_.col_indices, nnz * 4, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(ri_, ri.data(), nnz * 4, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(perm_, perm.data(), nnz * 4, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(v64_.val, lp_.values, nnz * 8, hipMemcpyHostToDevice, st));
    }
    MPX_HIP(hipMemcpyAsync(cs_, cs.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
    // Chunk-major copy of the CSR structure for the chunked dual SpMV (host
    // index work; the values follow through permch_ after scaling).
    std::vector<int> rsch, cich, permch;
    if (nch_ && n_long_r_ > 0) nch_ = 0;  // never with the one-wave row kernel
    if (nch_) {
        const long long cw = (static_cast<long long>(n_) + nch_ - 1) / nch_;
        std::vector<long long> pos(static_cast<size_t>(nch_) + 1, 0);
        for (size_t q = 0; q < nnz; ++q) ++pos[static_cast<size_t>(lp_.col_indices[q] / cw) + 1];
        for (int c = 0; c < nch_; ++c) pos[static_cast<size_t>(c) + 1] += pos[static_cast<size_t>(c)];
        rsch.resize(static_cast<size_t>(nch_) * (m + 1));
        cich.resize(nnz);
        permch.resize(nnz);
        for (int i = 0; i < m_; ++i) {
            for (int c = 0; c < nch_; ++c)
                rsch[static_cast<size_t>(c) * (m + 1) + i] = static_cast<int>(pos[static_cast<size_t>(c)]);
            for (int q = lp_.row_starts[i]; q < lp_.row_starts[i + 1]; ++q) {
                const size_t c = static_cast<size_t>(lp_.col_indices[q] / cw);
                const long long at = pos[c]++;
                cich[static_cast<size_t>(at)] = lp_.col_indices[q];
                permch[static_cast<size_t>(at)] = q;
            }
        }
        for (int c = 0; c < nch_; ++c)
            rsch[static_cast<size_t>(c) * (m + 1) + m] = static_cast<int>(pos[static_cast<size_t>(c)]);
        MPX_HIP(hipMemcpyAsync(rsch_, rsch.data(), rsch.size() * 4, hipMemcpyHostToDevice, st));
        if (nnz) {
            MPX_HIP(hipMemcpyAsync(cich_, cich.data(), nnz * 4, hipMemcpyHostToDevice, st));
            MPX_HIP(hipMemcpyAsync(permch_, permch.data(), nnz * 4, hipMemcpyHostToDevice, st));
        }
    }
    MPX_HIP(hipMemcpyAsync(obj_o_, lp_.obj, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(l_o_, lp_.col_lower, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(u_o_, lp_.col_upper, n * 8, hipMemcpyHostToDevice, st));
    if (m) {
        MPX_HIP(hipMemcpyAsync(rl_o_, lp_.row_lower, m * 8, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(ru_o_, lp_.row_upper, m * 8, hipMemcpyHostToDevice, st));
    }
    // Scales start at 1 (filled by a conversion from a host vector of ones is
    // wasteful; a kernel would do, but hipMemcpy of a constant buffer is
    // simplest and happens once).
    {
        std::vector<double> ones(std::max(n, m), 1.0);
        MPX_HIP(hipMemcpyAsync(row_scale_, ones.data(), m * 8, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(col_scale_, ones.data(), n * 8, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(v64_.sig, ones.data(), m * 8, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(v64_.tau, ones.data(), n * 8, hipMemcpyHostToDevice, st));
        const double one2[2] = {1.0, 1.0};
        MPX_HIP(hipMemcpyAsync(scales_, one2, 16, hipMemcpyHostToDevice, st));
        MPX_HIP(hipStreamSynchronize(st));
    }
    const int gm = std::max(1, ceilDiv(m_, kBlock)), gn = std::max(1, ceilDiv(n_, kBlock));
    const int gz = std::max(1, ceilDiv(nnz_, kBlock));
    // Ruiz (pdlp.cpp:520-558).
    if (o_.do_ruiz_scaling && m_ > 0 && n_ > 0 && nnz_ > 0) {
        for (int it = 0; it < o_.ruiz_iterations; ++it) {
            kRuizRows<<<gm, kBlock, 0, st>>>(m_, rs_, v64_.val, row_scale_);
            kRuizColMax<<<gn, kBlock, 0, st>>>(n_, cs_, perm_, v64_.val, cmax_, col_scale_);
            kRuizColApply<<<gz, kBlock, 0, st>>>(nnz_, ci_, cmax_, v64_.val);
        }
    }
    kScaleColumns<<<gn, kBlock, 0, st>>>(n_, sgn_, obj_o_, col_scale_, l_o_, u_o_, v64_.c, v64_.l,
                                          v64_.u);
    if (m_) kScaleRows<<<gm, kBlock, 0, st>>>(m_, row_scale_, rl_o_, ru_o_, v64_.rl, v64_.ru);
    if (nnz_) kGather<<<gz, kBlock, 0, st>>>(nnz_, perm_, v64_.val, v64_.valt);
    if (nnz_ && nch_) kGather<<<gz, kBlock, 0, st>>>(nnz_, permch_, v64_.val, valch64_);
    // Pock-Chambolle (pdlp.cpp:581-604).
    if (o_.do_pock_chambolle_scaling && m_ > 0 && n_ > 0) {
        kPockChambolle<<<gm, kBlock, 0, st>>>(m_, rs_, v64_.val, v64_.sig);
        kPockChambolle<<<gn, kBlock, 0, st>>>(n_, cs_, v64_.valt, v64_.tau);
    }
    // Bound / objective rescaling (pdlp.cpp:606-640).
    if (o_.do_bound_obj_rescaling && m_ > 0 && n_ > 0) {
        double* part = red_;
        kRescaleNorms<<<nbred_, kBlock, 0, st>>>(n_, m_, v64_.c, v64_.rl, v64_.ru, part, nbred_);
        kFinalize<<<1, kBlock, 0, st>>>(part, nbred_, 2, 0u, sums_);
        kRescaleFactors<<<1, 1, 0, st>>>(sums_, scales_);
        kApplyRescale<<<std::max(gn, gm), kBlock, 0, st>>>(n_, m_, scales_, v64_.c, v64_.l,
                                                             v64_.u, v64_.rl, v64_.ru);
    }
    MPX_HIP(hipGetLastError());
    double sc[2];
    MPX_HIP(hipMemcpyAsync(sc, scales_, 16, hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    os_ = sc[0];
    cs_host_ = sc[1];

    // Termination norms of the original problem (pdlp.cpp:53-81), host.
    if (o_.optimality_norm == 1) {
        double cn = 0.0, bn = 0.0;
        for (int j = 0; j < n_; ++j) cn = std::max(cn, std::abs(lp_.obj[j]));
        for (int i = 0; i < m_; ++i) {
            const double lo = lp_.row_lower[i], hi = lp_.row_upper[i];
            if (isFin(hi)) bn = std::max(bn, std::abs(hi));
            if (isFin(lo) && lo != hi) bn = std::max(bn, std::abs(lo));
        }
        cnorm_ = cn;
        bnorm_ = bn;
    } else {
        double cq = 0.0, bq = 0.0;
        for (int j = 0; j < n_; ++j) cq += lp_.obj[j] * lp_.obj[j];
        for (int i = 0; i < m_; ++i) {
            const double lo = lp_.row_lower[i], hi = lp_.row_upper[i];
            if (isFin(hi)) bq += hi * hi;
            if (isFin(lo) && lo != hi) bq += lo * lo;
        }
        cnorm_ = std::sqrt(cq);
        bnorm_ = std::sqrt(bq);
    }

    // Iterates start at zero or at the scaled warm start.
    MPX_HIP(hipMemsetAsync(v64_.x, 0, n * 8, st));
    MPX_HIP(hipMemsetAsync(v64_.x0, 0, n * 8, st));
    MPX_HIP(hipMemsetAsync(v64_.xh, 0, n * 8, st));
    MPX_HIP(hipMemsetAsync(v64_.xr, 0, n * 8, st));
    if (m) {
        MPX_HIP(hipMemsetAsync(v64_.y, 0, m * 8, st));
        MPX_HIP(hipMemsetAsync(v64_.y0, 0, m * 8, st));
        MPX_HIP(hipMemsetAsync(v64_.yh, 0, m * 8, st));
    }
    if (adaptive_) {
        MPX_HIP(hipMemsetAsync(prev_ax_, 0, m * 8, st));
        MPX_HIP(hipMemsetAsync(prev_x_, 0, n * 8, st));
    }
    if (warm_x && warm_y) {
        double* wx = cmax_;   // n scratch
        double* wy = prow_;   // needs m doubles: use y-hat as staging instead
        (void)wy;
        MPX_HIP(hipMemcpyAsync(wx, warm_x, n * 8, hipMemcpyHostToDevice, st));
        MPX_HIP(hipMemcpyAsync(v64_.yh, warm_y, m * 8, hipMemcpyHostToDevice, st));
        kWarmStart<<<std::max(gn, gm), kBlock, 0, st>>>(n_, m_, sgn_, wx, v64_.yh, col_scale_,
                                                         row_scale_, scales_, v64_.x, v64_.x0,
                                                         v64_.y, v64_.y0);
        MPX_HIP(hipMemsetAsync(v64_.yh, 0, m * 8, st));
    }
    MPX_HIP(hipGetLastError());
    pw_ = o_.primal_weight;
    powerIteration();
}

void SingleSolver::powerIteration() {
    hipStream_t st = stream_.s;
    if (n_ == 0 || m_ == 0) {
        step_ = 0.998;
        return;
    }
    // Start vector exactly as the reference (pdlp.cpp:756-764).
    std::mt19937 rng(42);
    std::normal_distribution<double> dist(0.0, 1.0);
    std::vector<double> xs(static_cast<size_t>(n_));
    for (double& v : xs) v = dist(rng);
    double xn = l2(xs);
    if (xn < 1e-15) xn = 1.0;
    for (double& v : xs) v /= xn;
    // Scratch: x in v64_.xr, xn in v64_.xh, y in v64_.yh (overwritten later).
    double* x = v64_.xr;
    double* xnew = v64_.xh;
    double* y = v64_.yh;
    MPX_HIP(hipMemcpyAsync(x, xs.data(), static_cast<size_t>(n_) * 8, hipMemcpyHostToDevice, st));
    PowerState* ps = reinterpret_cast<PowerState*>(sums_ + 8);
    MPX_HIP(hipMemsetAsync(ps, 0, sizeof(PowerState), st));
    double* s2 = sums_ + 4;
    double* rn = sums_ + 5;  // rn[0], rn[1]
    const int nbAt = nbc_;
    double* part = pcol_;
    const int gm = std::max(1, ceilDiv(static_cast<long long>(m_) * gr_, kBlock));
    const int gnAt = std::max(1, ceilDiv(static_cast<long long>(n_) * gc_, kBlock));
    const int gn = std::max(1, ceilDiv(n_, kBlock));
    auto oneStep = [&]() {
        withGroup(gr_, [&](auto GI) {
            constexpr int G = decltype(GI)::value;
            kPowerSpmvA<G><<<gm, kBlock, 0, st>>>(m_, rs_, ci_, v64_.val, x, y, ps);
        });
        withGroup(gc_, [&](auto GI) {
            constexpr int G = decltype(GI)::value;
            kPowerSpmvAt<G><<<gnAt, kBlock, 0, st>>>(n_, cs_, ri_, v64_.valt, y, x, xnew, part,
                                                      nbAt, ps);
        });
        kFinalize<<<1, kBlock, 0, st>>>(part, nbAt, 1, 0u, s2);
        kPowerResid<<<nbred_, kBlock, 0, st>>>(n_, x, xnew, s2, red_, nbred_, ps);
        kFinalize<<<1, kBlock, 0, st>>>(red_, nbred_, 2, 0u, rn);
        kPowerUpdate<<<1, 1, 0, st>>>(s2, rn, o_.sv_tol, o_.sv_max_iter, ps);
        kPowerNormalize<<<gn, kBlock, 0, st>>>(n_, xnew, x, ps);
        kPowerCheck<<<1, 1, 0, st>>>(o_.sv_tol, o_.sv_max_iter, ps);
    };
    // Steps run in captured groups; one host check per group.
    const int per = 16;
    Graph g;
    bool useGraph = !graphsDisabled();
    PowerState hs{};
    int launched = 0;
    while (launched < o_.sv_max_iter) {
        if (useGraph && g.e == nullptr) {
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                for (int s = 0; s < per; ++s) oneStep();
                if (hipStreamEndCapture(st, &g.g) != hipSuccess || g.g == nullptr ||
                    hipGraphInstantiate(&g.e, g.g, nullptr, nullptr, 0) != hipSuccess) {
                    g.reset();
                    useGraph = false;
                    (void)hipGetLastError();
                }
            } else {
                useGraph = false;
                (void)hipGetLastError();
            }
        }
        if (useGraph) {
            MPX_HIP(hipGraphLaunch(g.e, st));
        } else {
            for (int s = 0; s < per; ++s) oneStep();
        }
        launched += per;
        MPX_HIP(hipMemcpyAsync(&hs, ps, sizeof(PowerState), hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
        if (hs.done) break;
    }
    MPX_HIP(hipGetLastError());
    power_iters_ = hs.iters;
    step_ = 0.998 / std::sqrt(std::max(hs.sigma_sq, 1e-12));
}

template <typename T>
void SingleSolver::launchIteration(int k, bool storeHat) {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    // The refinement's row shift (FP32 correction phase only).
    const T* q = nullptr;
    if constexpr (std::is_same_v<T, float>) q = refine_ ? q32_ : nullptr;
    const int gridC = std::max(1, ceilDiv(static_cast<long long>(n_) * gc_, kBlock));
    const int gridR = std::max(1, ceilDiv(static_cast<long long>(m_) * gr_, kBlock));
    withGroup(gc_, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        if (storeHat)
            kPrimalStep<G, T, true><<<gridC, kBlock, 0, st>>>(n_, cs_, ri_, v.valt, v.y, v.x, v.x0,
                                                             v.xr, v.xh, v.c, v.tau, v.l, v.u,
                                                             sc_dev_, k, skip_c_);
        else
            kPrimalStep<G, T, false><<<gridC, kBlock, 0, st>>>(n_, cs_, ri_, v.valt, v.y, v.x,
                                                              v.x0, v.xr, v.xh, v.c, v.tau, v.l,
                                                              v.u, sc_dev_, k, skip_c_);
    });
    if (n_long_c_) {
        const int g = ceilDiv(static_cast<long long>(n_long_c_) * kWave, kBlock);
        if (storeHat)
            kPrimalStepLong<T, true><<<g, kBlock, 0, st>>>(n_long_c_, long_c_, cs_, ri_, v.valt,
                                                           v.y, v.x, v.x0, v.xr, v.xh, v.c, v.tau,
                                                           v.l, v.u, sc_dev_, k);
        else
            kPrimalStepLong<T, false><<<g, kBlock, 0, st>>>(n_long_c_, long_c_, cs_, ri_, v.valt,
                                                            v.y, v.x, v.x0, v.xr, v.xh, v.c,
                                                            v.tau, v.l, v.u, sc_dev_, k);
    }
    if (nch_) {
        // A x_r chunk by chunk (no long rows in this mode).
        const T* valc;
        T* part;
        if constexpr (std::is_same_v<T, double>) {
            valc = valch64_;
            part = part64_;
        } else {
            valc = valch32_;
            part = part32_;
        }
        const int gridCh = std::max(1, ceilDiv(static_cast<long long>(m_) * gch_, kBlock));
        withGroup(gch_, [&](auto GI) {
            constexpr int G = decltype(GI)::value;
            for (int c = 0; c < nch_; ++c) {
                const int* rsc = rsch_ + static_cast<size_t>(c) * (m_ + 1);
                if (storeHat)
                    kDualChunk<G, T, true><<<gridCh, kBlock, 0, st>>>(
                        m_, rsc, cich_, valc, v.xr, part, c, nch_, v.y, v.y0, v.yh, v.sig, v.rl,
                        v.ru, q, sc_dev_, k);
                else
                    kDualChunk<G, T, false><<<gridCh, kBlock, 0, st>>>(
                        m_, rsc, cich_, valc, v.xr, part, c, nch_, v.y, v.y0, v.yh, v.sig, v.rl,
                        v.ru, q, sc_dev_, k);
            }
        });
        return;
    }
    withGroup(gr_, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        if (storeHat)
            kDualStep<G, T, true><<<gridR, kBlock, 0, st>>>(m_, rs_, ci_, v.val, v.xr, v.y, v.y0,
                                                           v.yh, v.sig, v.rl, v.ru, q, sc_dev_, k,
                                                           skip_r_);
        else
            kDualStep<G, T, false><<<gridR, kBlock, 0, st>>>(m_, rs_, ci_, v.val, v.xr, v.y,
                                                            v.y0, v.yh, v.sig, v.rl, v.ru, q,
                                                            sc_dev_, k, skip_r_);
    });
    if (n_long_r_) {
        const int g = ceilDiv(static_cast<long long>(n_long_r_) * kWave, kBlock);
        if (storeHat)
            kDualStepLong<T, true><<<g, kBlock, 0, st>>>(n_long_r_, long_r_, rs_, ci_, v.val, v.xr,
                                                         v.y, v.y0, v.yh, v.sig, v.rl, v.ru, q,
                                                         sc_dev_, k);
        else
            kDualStepLong<T, false><<<g, kBlock, 0, st>>>(n_long_r_, long_r_, rs_, ci_, v.val,
                                                          v.xr, v.y, v.y0, v.yh, v.sig, v.rl,
                                                          v.ru, q, sc_dev_, k);
    }
}

template <typename T>
void SingleSolver::runBlock(int batch, Graph& graph, bool allowGraph) {
    hipStream_t st = stream_.s;
    pushScalars();
    const int N = o_.termination_eval_frequency;
    if (allowGraph && batch == N && !graphsDisabled()) {
        if (graph.e == nullptr) {
            bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                for (int k = 1; k <= batch; ++k) launchIteration<T>(k, k == batch);
                ok = hipStreamEndCapture(st, &graph.g) == hipSuccess && graph.g != nullptr &&
                     hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0) == hipSuccess;
            }
            if (!ok) {
                // Permanent fallback to direct launches (reference pdlp.cpp:2159-2168).
                (void)hipGetLastError();
                graph.reset();
                for (int k = 1; k <= batch; ++k) launchIteration<T>(k, k == batch);
                MPX_HIP(hipGetLastError());
                return;
            }
        }
        MPX_HIP(hipGraphLaunch(graph.e, st));
        return;
    }
    for (int k = 1; k <= batch; ++k) launchIteration<T>(k, k == batch);
    MPX_HIP(hipGetLastError());
}

template <typename T>
EvalOut SingleSolver::evaluate() {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    withGroup(gc_, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        kEvalCol<G, T><<<nbc_, kBlock, 0, st>>>(n_, cs_, ri_, v64_.valt, v.yh, v.y0, v.xh, v.x0,
                                                v64_.c, v64_.l, v64_.u, col_scale_, os_, pcol_,
                                                nbc_);
    });
    withGroup(gr_, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        kEvalRow<G, T><<<nbr_, kBlock, 0, st>>>(m_, rs_, ci_, v64_.val, v.xh, v.yh, v.y0, v64_.rl,
                                                v64_.ru, row_scale_, cs_host_, prow_, nbr_);
    });
    kFinalize<<<1, kBlock, 0, st>>>(pcol_, nbc_, kNumColFields, 1u << kDmax, metrics_);
    kFinalize<<<1, kBlock, 0, st>>>(prow_, nbr_, kNumRowFields, 1u << kPmax,
                                    metrics_ + kNumColFields);
    MPX_HIP(hipGetLastError());
    MPX_HIP(hipMemcpyAsync(metrics_host_, metrics_, 10 * sizeof(double), hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    const double* mc = metrics_host_;
    const double* mr = metrics_host_ + kNumColFields;
    EvalOut e{};
    e.dx2 = mc[kDx2];
    e.dy2 = mr[kDy2];
    const double movement = e.dx2 * pw_ + e.dy2 / pw_;
    const double interaction = mc[kInter] * (2.0 * step_);
    e.fpe = std::sqrt(std::max(movement + interaction, 0.0));
    e.rel_p = std::sqrt(mr[kPsq]) / (1.0 + bnorm_);
    e.rel_p_inf = mr[kPmax] / (1.0 + bnorm_);
    e.rel_d = std::sqrt(mc[kDsq]) / (1.0 + cnorm_);
    e.rel_d_inf = mc[kDmax] / (1.0 + cnorm_);
    const double unscale = 1.0 / (os_ * cs_host_);
    e.pobj = lp_.obj_offset + sgn_ * mc[kPobj] * unscale;
    e.dobj = lp_.obj_offset + sgn_ * (mc[kDcol] + mr[kDrow]) * unscale;
    e.gap = std::abs(e.pobj - e.dobj) / (1.0 + std::abs(e.pobj) + std::abs(e.dobj));
    return e;
}

template <typename T>
void SingleSolver::restartCopy() {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    kRestartCopy<T><<<std::max(1, ceilDiv(n_, kBlock)), kBlock, 0, st>>>(n_, v.xh, v.x, v.x0);
    if (m_) kRestartCopy<T><<<ceilDiv(m_, kBlock), kBlock, 0, st>>>(m_, v.yh, v.y, v.y0);
    MPX_HIP(hipGetLastError());
}

template <typename T>
void SingleSolver::downloadHats(std::vector<double>& xh, std::vector<double>& yh) {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    xh.resize(static_cast<size_t>(n_));
    yh.resize(static_cast<size_t>(m_));
    if constexpr (std::is_same_v<T, double>) {
        MPX_HIP(hipMemcpyAsync(xh.data(), v.xh, xh.size() * 8, hipMemcpyDeviceToHost, st));
        if (m_) MPX_HIP(hipMemcpyAsync(yh.data(), v.yh, yh.size() * 8, hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
    } else {
        std::vector<float> a(xh.size()), b(yh.size());
        MPX_HIP(hipMemcpyAsync(a.data(), v.xh, a.size() * 4, hipMemcpyDeviceToHost, st));
        if (m_) MPX_HIP(hipMemcpyAsync(b.data(), v.yh, b.size() * 4, hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < a.size(); ++i) xh[i] = a[i];
        for (size_t i = 0; i < b.size(); ++i) yh[i] = b[i];
    }
}

template <typename T>
void SingleSolver::recover(double* x_out, double* y_out, double* objective) {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    double* xo = cmax_;  // n-sized scratch
    std::vector<double> hx(static_cast<size_t>(n_)), hy(static_cast<size_t>(m_));
    kRecover<T><<<std::max(1, ceilDiv(std::max(n_, m_), kBlock)), kBlock, 0, st>>>(
        n_, 0, sgn_, v.xh, v.yh, col_scale_, row_scale_, scales_, l_o_, u_o_, 1, xo, nullptr);
    MPX_HIP(hipMemcpyAsync(hx.data(), xo, hx.size() * 8, hipMemcpyDeviceToHost, st));
    std::vector<double> xh_unused, yh;
    MPX_HIP(hipStreamSynchronize(st));
    downloadHats<T>(xh_unused, yh);
    std::vector<double> rsc(static_cast<size_t>(m_));
    if (m_) {
        MPX_HIP(hipMemcpyAsync(rsc.data(), row_scale_, rsc.size() * 8, hipMemcpyDeviceToHost, st));
        MPX_HIP(hipStreamSynchronize(st));
    }
    const double inv_os = 1.0 / os_;
    for (int i = 0; i < m_; ++i) hy[i] = sgn_ * yh[i] * rsc[i] * inv_os;
    double objv = lp_.obj_offset;
    for (int j = 0; j < n_; ++j) objv += lp_.obj[j] * hx[j];
    *objective = objv;
    if (x_out) std::memcpy(x_out, hx.data(), hx.size() * 8);
    if (y_out) std::memcpy(y_out, hy.data(), hy.size() * 8);
}

// Moves the refinement's base to the evaluated point (v64_.xh / yh, already
// materialized in FP64) and rebuilds the correction's data around it.
void SingleSolver::recenter() {
    hipStream_t st = stream_.s;
    MPX_HIP(hipMemcpyAsync(v64_.x, v64_.xh, static_cast<size_t>(n_) * 8, hipMemcpyDeviceToDevice, st));
    if (m_)
        MPX_HIP(hipMemcpyAsync(v64_.y, v64_.yh, static_cast<size_t>(m_) * 8, hipMemcpyDeviceToDevice,
                               st));
    kRefineCols<float><<<std::max(1, ceilDiv(n_, kBlock)), kBlock, 0, st>>>(
        n_, cs_, ri_, v64_.valt, v64_.c, v64_.l, v64_.u, v64_.x, v64_.y, v32_.c, v32_.l, v32_.u,
        v32_.x, v32_.x0, v32_.xh);
    if (m_)
        kRefineRows<float><<<ceilDiv(m_, kBlock), kBlock, 0, st>>>(
            m_, rs_, ci_, v64_.val, v64_.rl, v64_.ru, v64_.x, v64_.y, v32_.rl, v32_.ru, q32_,
            v32_.y, v32_.y0, v32_.yh);
    MPX_HIP(hipGetLastError());
}

// FP64 evaluation point of the correction phase: x^ = x_b + dx^ (snapped to
// the FP64 bounds), y^ = y_b + dy^ and the anchors, into v64_.xh / x0 / yh / y0.
void SingleSolver::materialize() {
    hipStream_t st = stream_.s;
    kRefineMaterialize<float><<<std::max(1, ceilDiv(std::max(n_, m_), kBlock)), kBlock, 0, st>>>(
        n_, m_, v64_.x, v64_.y, v32_.xh, v32_.x0, v32_.yh, v32_.y0, v32_.l, v32_.u, v64_.l,
        v64_.u, q32_, v32_.rl, v32_.ru, v64_.xh, v64_.x0, v64_.yh, v64_.y0);
    MPX_HIP(hipGetLastError());
}

void SingleSolver::convertIterates64to32() {
    hipStream_t st = stream_.s;
    auto cv = [&](float* d, const double* s, int count) {
        if (count > 0)
            kConvert<float, double><<<ceilDiv(count, kBlock), kBlock, 0, st>>>(count, s, d);
    };
    cv(v32_.x, v64_.x, n_);
    cv(v32_.x0, v64_.x0, n_);
    cv(v32_.xh, v64_.xh, n_);
    cv(v32_.y, v64_.y, m_);
    cv(v32_.y0, v64_.y0, m_);
    cv(v32_.yh, v64_.yh, m_);
    MPX_HIP(hipGetLastError());
}

void SingleSolver::convertIterates32to64() {
    hipStream_t st = stream_.s;
    auto cv = [&](double* d, const float* s, int count) {
        if (count > 0)
            kConvert<double, float><<<ceilDiv(count, kBlock), kBlock, 0, st>>>(count, s, d);
    };
    cv(v64_.xh, v32_.xh, n_);
    cv(v64_.yh, v32_.yh, m_);
    cv(v64_.x0, v32_.x0, n_);
    cv(v64_.y0, v32_.y0, m_);
    MPX_HIP(hipGetLastError());
}

void SingleSolver::download(mpx_scaled_problem* out) {
    hipStream_t st = stream_.s;
    const size_t m = static_cast<size_t>(m_), n = static_cast<size_t>(n_),
                 nnz = static_cast<size_t>(nnz_);
    auto get = [&](double* dst, const double* src, size_t count) {
        if (dst && count) MPX_HIP(hipMemcpyAsync(dst, src, count * 8, hipMemcpyDeviceToHost, st));
    };
    auto geti = [&](int32_t* dst, const int* src, size_t count) {
        if (dst && count) MPX_HIP(hipMemcpyAsync(dst, src, count * 4, hipMemcpyDeviceToHost, st));
    };
    get(out->values, v64_.val, nnz);
    get(out->at_values, v64_.valt, nnz);
    geti(out->at_row_starts, cs_, n + 1);
    geti(out->at_col_indices, ri_, nnz);
    get(out->cscaled, v64_.c, n);
    get(out->col_lower, v64_.l, n);
    get(out->col_upper, v64_.u, n);
    get(out->row_lower, v64_.rl, m);
    get(out->row_upper, v64_.ru, m);
    get(out->row_scale, row_scale_, m);
    get(out->col_scale, col_scale_, n);
    get(out->sigma_base, v64_.sig, m);
    get(out->tau_base, v64_.tau, n);
    MPX_HIP(hipStreamSynchronize(st));
    out->objective_scale = os_;
    out->constraint_scale = cs_host_;
    out->spectral_norm = 0.998 / step_;
    out->power_iterations = power_iters_;
}

// Converts the fp64 scaled problem into the fp32 iteration copies.
template <typename T>
void convertProblem(hipStream_t st, IterVecs<T>& d, const IterVecs<double>& s, int m, int n,
                    int nnz) {
    auto cv = [&](T* dst, const double* src, long long count) {
        if (count > 0)
            kConvert<T, double><<<ceilDiv(count, kBlock), kBlock, 0, st>>>(count, src, dst);
    };
    cv(d.val, s.val, nnz);
    cv(d.valt, s.valt, nnz);
    cv(d.c, s.c, n);
    cv(d.tau, s.tau, n);
    cv(d.l, s.l, n);
    cv(d.u, s.u, n);
    cv(d.sig, s.sig, m);
    cv(d.rl, s.rl, m);
    cv(d.ru, s.ru, m);
}

// Adaptive step of the GPU host loop (pdlp.cpp:2271-2307), after the
// termination test of an evaluation: from this x^ and A x^ and the previous
// evaluation's, local = 0.998 / sqrt(sum sigma (A dx)^2 / sum dx^2 / tau);
// the step shrinks (adaptive_step_down) above it and grows
// (adaptive_step_up) otherwise, within [lb, ub].  The first evaluation after
// a restart only records its point.
template <typename T>
void SingleSolver::adaptStep(double lb, double ub) {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    withGroup(gr_, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        kAdaptRows<G, T><<<nbr_, kBlock, 0, st>>>(m_, rs_, ci_, v64_.val, v.xh, v64_.sig, prev_ax_,
                                                  acc_part_);
    });
    kFinalize<<<1, kBlock, 0, st>>>(acc_part_, nbr_, 1, 0u, acc_out_);
    kAdaptCols<T><<<nbn_, kBlock, 0, st>>>(n_, v.xh, v64_.tau, prev_x_, acc_part_);
    kFinalize<<<1, kBlock, 0, st>>>(acc_part_, nbn_, 1, 0u, acc_out_ + 1);
    MPX_HIP(hipGetLastError());
    if (!have_prev_) {
        have_prev_ = true;
        return;
    }
    double* h = static_cast<double*>(acc_pinned_->p);
    MPX_HIP(hipMemcpyAsync(h, acc_out_, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    const double adiff = h[0], xdiff = h[1];
    if (xdiff > 1e-30) {
        const double local = 0.998 / std::sqrt(maxRef(adiff / xdiff, 1e-30));
        const double ns = step_ > local ? step_ * o_.adaptive_step_down : step_ * o_.adaptive_step_up;
        step_ = clampRef(ns, lb, ub);
    }
}

// Anderson restart (pdlp.cpp:2418-2466 with the accelerator of 293-437):
// stores [x^; y^] and its residual against [x0; y0], then extrapolates over
// the newest d = min(count, window) points: Gram matrix of the residuals on
// the device, (G + mu I) a = 1 by Cholesky on the host (mu = reg * trace /
// d), a normalised to sum 1, accepted only if |sum a_i r_i| < |r_newest|.
// Returns true when x0 = x and y0 = y now hold the extrapolated point
// (x projected onto the scaled bounds); false means a plain restart is due.
template <typename T>
bool SingleSolver::andersonRestart() {
    IterVecs<T>& v = vecs<T>();
    hipStream_t st = stream_.s;
    const long long dim = static_cast<long long>(n_) + m_;
    const int slot = aa_count_ % aa_w_;
    kAaPush<T><<<ceilDiv(dim, kBlock), kBlock, 0, st>>>(n_, m_, v.xh, v.yh, v.x0, v.y0,
                                                        aa_z_ + slot * dim, aa_r_ + slot * dim);
    ++aa_count_;
    const int d = std::min(aa_count_, aa_w_);
    if (d < 2) {
        MPX_HIP(hipGetLastError());
        return false;
    }
    const int first = (aa_count_ - d) % aa_w_;
    const int pairs = d * (d + 1) / 2;
    kAaGram<<<dim3(nbaa_, pairs), kBlock, 0, st>>>(dim, d, first, aa_w_, aa_r_, acc_part_);
    kFinalize<<<1, kBlock, 0, st>>>(acc_part_, nbaa_, pairs, 0u, acc_out_);
    MPX_HIP(hipGetLastError());
    double* h = static_cast<double*>(acc_pinned_->p);
    MPX_HIP(hipMemcpyAsync(h, acc_out_, pairs * sizeof(double), hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    std::vector<double> g(static_cast<size_t>(d * d));
    for (int i = 0, p = 0; i < d; ++i)
        for (int j = i; j < d; ++j, ++p) g[i * d + j] = g[j * d + i] = h[p];
    const double newest = g[(d - 1) * d + (d - 1)];
    double tr = 0.0;
    for (int i = 0; i < d; ++i) tr += g[i * d + i];
    const double mu = o_.anderson_regularization * tr / static_cast<double>(d);
    for (int i = 0; i < d; ++i) g[i * d + i] += mu;
    std::vector<double> l(static_cast<size_t>(d * d), 0.0), t(static_cast<size_t>(d)),
        a(static_cast<size_t>(d));
    for (int i = 0; i < d; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = g[i * d + j];
            for (int q = 0; q < j; ++q) s -= l[i * d + q] * l[j * d + q];
            if (i == j) {
                if (s <= 0.0) return false;
                l[i * d + i] = std::sqrt(s);
            } else {
                l[i * d + j] = s / l[j * d + j];
            }
        }
    for (int i = 0; i < d; ++i) {
        double s = 1.0;
        for (int q = 0; q < i; ++q) s -= l[i * d + q] * t[q];
        t[i] = s / l[i * d + i];
    }
    for (int i = d - 1; i >= 0; --i) {
        double s = t[i];
        for (int q = i + 1; q < d; ++q) s -= l[q * d + i] * a[q];
        a[i] = s / l[i * d + i];
    }
    double sa = 0.0;
    for (int i = 0; i < d; ++i) sa += a[i];
    if (std::abs(sa) < 1e-15) return false;
    double* ha = h + pairs;
    for (int i = 0; i < d; ++i) ha[i] = a[i] / sa;
    MPX_HIP(hipMemcpyAsync(aa_alpha_, ha, d * sizeof(double), hipMemcpyHostToDevice, st));
    kAaResid<<<nbaa_, kBlock, 0, st>>>(dim, d, first, aa_w_, aa_r_, aa_alpha_, acc_part_);
    kFinalize<<<1, kBlock, 0, st>>>(acc_part_, nbaa_, 1, 0u, acc_out_);
    MPX_HIP(hipGetLastError());
    MPX_HIP(hipMemcpyAsync(h, acc_out_, sizeof(double), hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
    if (!(h[0] < newest)) return false;
    kAaApply<T><<<ceilDiv(dim, kBlock), kBlock, 0, st>>>(n_, m_, d, first, aa_w_, aa_z_, aa_alpha_,
                                                         v64_.l, v64_.u, v.x, v.x0, v.y, v.y0);
    MPX_HIP(hipGetLastError());
    return true;
}

// Evaluations without halving the KKT ratio after which the plain FP32 phase
// hands over (kRefineStall), and evaluations without a 1% gain after which
// the correction phase falls back to FP64 (ten times that: on the seeded
// families the correction phase converges in FP32 with the iteration count
// of FP64 PDHG, so only a real precision stall of the FP32 matrix should
// trigger it).
constexpr int kRefineStall = 20;
constexpr double kRefineGain = 0.99;

void SingleSolver::solve(const std::chrono::steady_clock::time_point t0, double* x_out,
                         double* y_out, mpx_pdlp_result* res) {
    const mpx_pdlp_options& o = o_;
    const int N = o.termination_eval_frequency;
    auto timeUp = [&]() {
        if (o.max_solve_seconds < 0.0) return false;
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >=
               o.max_solve_seconds;
    };
    auto stopRequested = [&]() { return o.stop_fn != nullptr && o.stop_fn(o.stop_user) != 0; };
    const double work_per_iter =
        4.0 * static_cast<double>(nnz_) + 2.0 * static_cast<double>(m_ + n_);
    const double step_lb = step_ * o.adaptive_step_min, step_ub = step_ * o.adaptive_step_max;
    have_prev_ = false;
    aa_count_ = 0;

    int total = 0, restarts = 0, fp32_iters = 0, refine_iters = 0;
    double fpe0 = -1.0, last_fpe = std::numeric_limits<double>::infinity();
    double pw_err_sum = 0.0, pw_last_err = 0.0, best_pw = pw_, best_score =
                                                                 std::numeric_limits<double>::infinity();
    double precond_pw = o.primal_weight;
    bool converged = false;
    EvalOut e{};
    inner_ = 0;
    int status = MPX_STATUS_ITER_LIMIT;
    bool done = false;
    Graph g64, g32, g32r;
    // Mixed precision: FP32 blocks until the FP64-evaluated KKT reaches the
    // handover ratio or stops improving; then FP32 blocks on the correction
    // around an FP64 base that moves at every restart (iterative refinement,
    // pdlp_refine.h), or -- with mixed_refine off, adaptive steps or Anderson
    // acceleration -- FP64 blocks from the FP32 iterate.  A correction phase
    // that stops improving hands over to FP64 as well.
    bool fp32_phase = mixed_;
    refine_ = false;
    const bool can_refine = mixed_ && o.mixed_refine != 0 && !adaptive_ && !aa_w_;
    int stall = 0, rstall = 0;
    double best_kkt = std::numeric_limits<double>::infinity(), rbest = best_kkt;
    if (fp32_phase) {
        convertProblem<float>(stream_.s, v32_, v64_, m_, n_, nnz_);
        convertChunks();
        convertIterates64to32();
    }
    // Mixed precision: the FP64 phase runs to refine_tol (default: the solve's
    // own tolerances), so the answer is an FP64-converged point.
    const double rfac = 1.0;
    const double rtol_p = o.refine_tol > 0.0 ? std::min(o.refine_tol, o.primal_tol) : rfac * o.primal_tol;
    const double rtol_d = o.refine_tol > 0.0 ? std::min(o.refine_tol, o.dual_tol) : rfac * o.dual_tol;
    const double rtol_g =
        o.refine_tol > 0.0 ? std::min(o.refine_tol, o.optimality_tol) : rfac * o.optimality_tol;
    auto kktOf = [&](const EvalOut& ev) {
        const double tp = o.optimality_norm == 1 ? ev.rel_p_inf : ev.rel_p;
        const double td = o.optimality_norm == 1 ? ev.rel_d_inf : ev.rel_d;
        return std::max(std::max(tp / o.primal_tol, td / o.dual_tol), ev.gap / o.optimality_tol);
    };
    // A restart of the main loop may extrapolate (Anderson acceleration);
    // the handover and the polish phase restart plainly.
    auto restartNow = [&](bool accelerate) {
        bool extrapolated = false;
        if (accelerate && aa_w_)
            extrapolated = fp32_phase ? andersonRestart<float>() : andersonRestart<double>();
        if (!extrapolated) {
            if (refine_) recenter();
            else if (fp32_phase) restartCopy<float>();
            else restartCopy<double>();
        }
        have_prev_ = false;
        inner_ = 0;
        fpe0 = -1.0;
        last_fpe = std::numeric_limits<double>::infinity();
    };
    auto updatePw = [&](const EvalOut& ev, bool polish) {
        const double pd = std::sqrt(ev.dx2), dd = std::sqrt(ev.dy2);
        const double ratio = (ev.rel_p > 1e-15) ? ev.rel_d / ev.rel_p : 0.0;
        if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
            const double err = std::log(dd) - std::log(pd) - std::log(pw_);
            pw_err_sum = o.pid_i_smooth * pw_err_sum + err;
            const double derr = err - pw_last_err;
            pw_ *= std::exp(o.pid_kp * err + o.pid_ki * pw_err_sum + o.pid_kd * derr);
            pw_ = clampRef(pw_, 1e-4, 1e8);
            pw_last_err = err;
            if (!polish) {
                const double score = std::abs(std::log10(std::max(ev.rel_d, 1e-15)) -
                                              std::log10(std::max(ev.rel_p, 1e-15)));
                if (score < best_score) {
                    best_score = score;
                    best_pw = pw_;
                }
            }
        } else if (!polish) {
            pw_ = best_pw;
            pw_err_sum = 0.0;
            pw_last_err = 0.0;
        }
    };

    while (total < o.max_iter) {
        if (timeUp()) {
            status = MPX_STATUS_TIME_LIMIT;
            done = true;
            break;
        }
        if (stopRequested()) {
            status = MPX_STATUS_ITER_LIMIT;
            done = true;
            break;
        }
        const int batch = std::min(N, o.max_iter - total);
        if (refine_) runBlock<float>(batch, g32r, true);
        else if (fp32_phase) runBlock<float>(batch, g32, true);
        else runBlock<double>(batch, g64, true);
        inner_ += batch;
        total += batch;
        if (fp32_phase || refine_) fp32_iters += batch;
        if (refine_) refine_iters += batch;
        if (refine_) {
            materialize();
            e = evaluate<double>();
        } else {
            e = fp32_phase ? evaluate<float>() : evaluate<double>();
        }
        if (fpe0 < 0.0) fpe0 = e.fpe;
        const double tp = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
        const double td = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
        if (o.verbose && (total <= N || total % (10 * N) == 0))
            std::printf("PDLP %7d  pobj=% .8e  pinf=% .2e  dinf=% .2e  gap=% .2e  fpe=% .2e  "
                        "pw=% .2e  step=% .2e  [MI355X%s]\n",
                        total, e.pobj, tp, td, e.gap, e.fpe, pw_, step_,
                        refine_ ? " fp32 refine" : (fp32_phase ? " fp32" : ""));
        if (fp32_phase) {
            // FP32 iterations judged in FP64: hand over to FP64 (restarted at
            // the FP32 point) once the KKT ratio is within mixed_handover, or when FP32
            // stops making progress.
            const double k = kktOf(e);
            bool handover = k <= o.mixed_handover;
            if (k < 0.5 * best_kkt) {
                best_kkt = k;
                stall = 0;
            } else if (++stall >= kRefineStall) {
                handover = true;
            }
            if (handover) {
                convertIterates32to64();
                fp32_phase = false;
                refine_ = can_refine;  // the base is the FP32 point (v64_.xh / yh)
                restartNow(false);
                rbest = kktOf(e);
                rstall = 0;
                ++restarts;
                continue;
            }
        } else if (tp <= (mixed_ ? rtol_p : o.primal_tol) && td <= (mixed_ ? rtol_d : o.dual_tol) &&
                   e.gap <= (mixed_ ? rtol_g : o.optimality_tol)) {
            converged = true;
            break;
        } else if (refine_) {
            // A correction phase that stops improving continues in FP64
            // from its evaluated point.  Re-centring keeps the FP32
            // correction small, so only a KKT ratio that no longer improves
            // counts: not 1% below its best for 10 kRefineStall evaluations
            // and for half of the phase so far.  The second bound follows
            // PDHG's own pace: on C1 the gap sits near 1e-5 for ~100,000
            // iterations while FP64 PDHG does the same from the same point.
            const double k = kktOf(e);
            const int window = std::max(10 * kRefineStall, refine_iters / (2 * N));
            if (k < kRefineGain * rbest) {
                rbest = k;
                rstall = 0;
            } else if (++rstall >= window) {
                refine_ = false;
                restartNow(false);
                ++restarts;
                continue;
            }
        }
        if (adaptive_) {
            if (fp32_phase) adaptStep<float>(step_lb, step_ub);
            else adaptStep<double>(step_lb, step_ub);
        }
        // Certificates every 10 N (pdlp.cpp:2309-2341), on the host.
        if (total >= 10 * N && total % (10 * N) == 0) {
            std::vector<double> xh, yh;
            if (fp32_phase) downloadHats<float>(xh, yh);
            else downloadHats<double>(xh, yh);
            std::vector<double> rsc(static_cast<size_t>(m_)), csc(static_cast<size_t>(n_));
            MPX_HIP(hipMemcpy(rsc.data(), row_scale_, rsc.size() * 8, hipMemcpyDeviceToHost));
            MPX_HIP(hipMemcpy(csc.data(), col_scale_, csc.size() * 8, hipMemcpyDeviceToHost));
            std::vector<double> yo(static_cast<size_t>(m_)), xo(static_cast<size_t>(n_));
            const double inv_os = 1.0 / os_;
            for (int i = 0; i < m_; ++i) yo[i] = sgn_ * yh[i] * rsc[i] * inv_os;
            if (farkasHolds(orig_, yo, 1e-6)) {
                status = MPX_STATUS_INFEASIBLE;
                if (y_out) std::memcpy(y_out, yo.data(), yo.size() * 8);
                if (x_out) std::memset(x_out, 0, static_cast<size_t>(n_) * 8);
                done = true;
                break;
            }
            const double inv_cs = 1.0 / cs_host_;
            for (int j = 0; j < n_; ++j) xo[j] = xh[j] * csc[j] * inv_cs;
            if (rayHolds(orig_, xo, 1e-6)) {
                status = MPX_STATUS_UNBOUNDED;
                if (x_out) std::memcpy(x_out, xo.data(), xo.size() * 8);
                if (y_out) std::memset(y_out, 0, static_cast<size_t>(m_) * 8);
                done = true;
                break;
            }
        }
        bool restart;
        if (total == N) restart = true;
        else if (fpe0 > 0.0 && e.fpe <= o.restart_sufficient_decay * fpe0) restart = true;
        else if (fpe0 > 0.0 && e.fpe <= o.restart_necessary_decay * fpe0 && e.fpe > last_fpe)
            restart = true;
        else
            restart = inner_ >= static_cast<int>(o.restart_artificial_fraction *
                                                  static_cast<double>(total));
        last_fpe = e.fpe;
        if (restart) {
            ++restarts;
            if (o.update_primal_weight) updatePw(e, false);
            if (o.preconditioner_refresh && precond_pw > 0.0) {
                // The refresh recomputes sigma/tau from the unchanged scaled
                // matrix (pdlp.cpp:713-742): only the recorded weight moves.
                const double drift = pw_ / precond_pw;
                if ((drift > o.preconditioner_refresh_ratio ||
                     drift < 1.0 / o.preconditioner_refresh_ratio) &&
                    o.do_pock_chambolle_scaling)
                    precond_pw = pw_;
            }
            restartNow(true);
        }
    }

    // The polish phase and recovery run in FP64 on the evaluated point.
    refine_ = false;
    if (!done && converged && o.enable_polishing && !timeUp()) {
        const double rp = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
        const double rd = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
        if (rp > o.polish_primal_tol || rd > o.polish_dual_tol) {
            restartNow(false);
            int pcount = 0;
            while (pcount < o.polish_max_iter) {
                if (timeUp() || stopRequested()) break;
                const int batch = std::min(N, o.polish_max_iter - pcount);
                runBlock<double>(batch, g64, false);
                inner_ += batch;
                pcount += batch;
                total += batch;
                e = evaluate<double>();
                if (fpe0 < 0.0) fpe0 = e.fpe;
                const double rpn = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
                const double rdn = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
                if (o.verbose && pcount % (10 * N) == 0)
                    std::printf("PDLP polish %7d  pinf=% .2e  dinf=% .2e  fpe=% .2e  [MI355X]\n",
                                pcount, rpn, rdn, e.fpe);
                if (rpn <= o.polish_primal_tol && rdn <= o.polish_dual_tol) break;
                bool restart;
                if (pcount == batch) restart = true;
                else if (fpe0 > 0.0 && e.fpe <= o.restart_sufficient_decay * fpe0) restart = true;
                else if (fpe0 > 0.0 && e.fpe <= o.restart_necessary_decay * fpe0 &&
                         e.fpe > last_fpe)
                    restart = true;
                else
                    restart = inner_ >= static_cast<int>(o.restart_artificial_fraction *
                                                          static_cast<double>(pcount));
                last_fpe = e.fpe;
                if (restart) {
                    ++restarts;
                    if (o.update_primal_weight) updatePw(e, true);
                    restartNow(false);
                }
            }
        }
    }

    double objv = 0.0;
    if (!done) {
        if (converged) {
            recover<double>(x_out, y_out, &objv);
            status = MPX_STATUS_OPTIMAL;
        } else {
            status = timeUp() ? MPX_STATUS_TIME_LIMIT : MPX_STATUS_ITER_LIMIT;
        }
    }
    res->status = status;
    res->iterations = total;
    res->objective = status == MPX_STATUS_OPTIMAL ? objv : 0.0;
    res->work_units = static_cast<double>(total) * work_per_iter;
    res->rel_primal = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
    res->rel_dual = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
    res->rel_gap = e.gap;
    res->primal_obj = e.pobj;
    res->dual_obj = e.dobj;
    res->primal_weight = pw_;
    res->step = step_;
    res->restarts = restarts;
    res->iterations_fp32 = fp32_iters;
    res->iterations_refine = refine_iters;
}

void SingleSolver::iterateFrom(const mpx_scaled_problem* sp, int iters, int inner, double step,
                               double pw, int precision, double* cx, double* ix, double* px,
                               double* cy, double* iy, double* py) {
    hipStream_t st = stream_.s;
    const size_t m = static_cast<size_t>(m_), n = static_cast<size_t>(n_),
                 nnz = static_cast<size_t>(nnz_);
    // Overwrite the scaled data with the caller's.
    MPX_HIP(hipMemcpyAsync(v64_.val, sp->values, nnz * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.valt, sp->at_values, nnz * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.c, sp->cscaled, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.l, sp->col_lower, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.u, sp->col_upper, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.rl, sp->row_lower, m * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.ru, sp->row_upper, m * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.sig, sp->sigma_base, m * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.tau, sp->tau_base, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.x, cx, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.x0, ix, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.xh, px, n * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.y, cy, m * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.y0, iy, m * 8, hipMemcpyHostToDevice, st));
    MPX_HIP(hipMemcpyAsync(v64_.yh, py, m * 8, hipMemcpyHostToDevice, st));
    // The chunk-major copy of A used by the chunked dual step follows the
    // caller's values too (setup gathered it from its own scaling).
    if (nnz_ && nch_)
        kGather<<<ceilDiv(nnz_, kBlock), kBlock, 0, st>>>(nnz_, permch_, v64_.val, valch64_);
    step_ = step;
    pw_ = pw;
    inner_ = inner;
    Graph none;
    if (precision == MPX_PREC_MIXED) {
        if (!mixed_) {
            setError("mpx_pdhg_iterate: fp32 buffers not allocated");
            throw DeviceError();
        }
        convertProblem<float>(st, v32_, v64_, m_, n_, nnz_);
        convertChunks();
        auto cv = [&](float* d, const double* s, int count) {
            if (count > 0)
                kConvert<float, double><<<ceilDiv(count, kBlock), kBlock, 0, st>>>(count, s, d);
        };
        cv(v32_.x, v64_.x, n_);
        cv(v32_.x0, v64_.x0, n_);
        cv(v32_.xh, v64_.xh, n_);
        cv(v32_.y, v64_.y, m_);
        cv(v32_.y0, v64_.y0, m_);
        cv(v32_.yh, v64_.yh, m_);
        runBlock<float>(iters, none, false);
        auto cb = [&](double* d, const float* s, int count) {
            if (count > 0)
                kConvert<double, float><<<ceilDiv(count, kBlock), kBlock, 0, st>>>(count, s, d);
        };
        cb(v64_.x, v32_.x, n_);
        cb(v64_.xh, v32_.xh, n_);
        cb(v64_.y, v32_.y, m_);
        cb(v64_.yh, v32_.yh, m_);
    } else {
        runBlock<double>(iters, none, false);
    }
    MPX_HIP(hipMemcpyAsync(cx, v64_.x, n * 8, hipMemcpyDeviceToHost, st));
    MPX_HIP(hipMemcpyAsync(px, v64_.xh, n * 8, hipMemcpyDeviceToHost, st));
    MPX_HIP(hipMemcpyAsync(cy, v64_.y, m * 8, hipMemcpyDeviceToHost, st));
    MPX_HIP(hipMemcpyAsync(py, v64_.yh, m * 8, hipMemcpyDeviceToHost, st));
    MPX_HIP(hipStreamSynchronize(st));
}

bool validLp(const mpx_lp* lp) {
    if (!lp) return false;
    if (lp->num_rows < 0 || lp->num_cols < 0) return false;
    if (!lp->row_starts && lp->num_rows > 0) return false;
    if (lp->num_cols > 0 && (!lp->obj || !lp->col_lower || !lp->col_upper)) return false;
    if (lp->num_rows > 0 && (!lp->row_lower || !lp->row_upper)) return false;
    if (lp->num_rows > 0) {
        if (lp->row_starts[0] != 0) return false;
        const int nnz = lp->row_starts[lp->num_rows];
        if (nnz < 0 || (nnz > 0 && (!lp->col_indices || !lp->values))) return false;
        for (int i = 0; i < lp->num_rows; ++i)
            if (lp->row_starts[i + 1] < lp->row_starts[i]) return false;
        for (int k = 0; k < nnz; ++k)
            if (lp->col_indices[k] < 0 || lp->col_indices[k] >= lp->num_cols) return false;
    }
    return true;
}

}  // namespace mpx

using namespace mpx;

extern "C" {

void mpx_pdlp_default_options(mpx_pdlp_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iter = 4000000;
    o->optimality_tol = 5e-5;
    o->primal_tol = 5e-5;
    o->dual_tol = 5e-5;
    o->max_solve_seconds = -1.0;
    o->optimality_norm = 0;
    o->sv_max_iter = 5000;
    o->sv_tol = 1e-4;
    o->adaptive_step = 0;
    o->adaptive_step_down = 0.7;
    o->adaptive_step_up = 1.0;
    o->adaptive_step_min = 1e-6;
    o->adaptive_step_max = 1e3;
    o->primal_weight = 1.0;
    o->update_primal_weight = 1;
    o->pid_kp = 0.99;
    o->pid_ki = 0.01;
    o->pid_kd = 0.0;
    o->pid_i_smooth = 0.3;
    o->restart_sufficient_decay = 0.2;
    o->restart_necessary_decay = 0.5;
    o->restart_artificial_fraction = 0.36;
    o->termination_eval_frequency = 200;
    o->matrix_zero_tol = 0.0;
    o->enable_polishing = 1;
    o->polish_primal_tol = 1e-6;
    o->polish_dual_tol = 1e-6;
    o->polish_max_iter = 500000;
    o->do_ruiz_scaling = 1;
    o->ruiz_iterations = 10;
    o->do_pock_chambolle_scaling = 1;
    o->do_bound_obj_rescaling = 1;
    o->preconditioner_refresh = 0;
    o->preconditioner_refresh_ratio = 10.0;
    o->anderson_acceleration = 0;
    o->anderson_window_size = 5;
    o->anderson_regularization = 1e-10;
    o->use_gpu = 1;
    o->gpu_min_rows = 512;
    o->gpu_min_nnz = 10000;
    o->verbose = 1;
    o->stop_fn = nullptr;
    o->stop_user = nullptr;
    o->precision = MPX_PREC_FP64;
    o->refine_tol = -1.0;
    o->spmv_group = 0;
    o->device = 0;
    o->mixed_handover = 3.0;
    o->spmv_chunk_cols = 0;
    o->packed_preset = -1;
    o->mixed_refine = 1;
}

const char* mpx_last_error(void) { return mpx::lastError(); }
const char* mpx_version(void) { return "mipx_amd 0.1 (gfx950)"; }

int mpx_device_count(int32_t* count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
    return MPX_OK;
}

// Streaming copy: four 16-byte loads in flight per thread, non-temporal
// (the data is touched once); the tail is copied one element at a time.
using V4i = int __attribute__((ext_vector_type(4)));
__global__ void kCopyBench(long long n4, const int4* __restrict__ src4, int4* __restrict__ dst4) {
    const V4i* src = reinterpret_cast<const V4i*>(src4);
    V4i* dst = reinterpret_cast<V4i*>(dst4);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    for (; t + 3 * stride < n4; t += 4 * stride) {
        const V4i a = __builtin_nontemporal_load(src + t);
        const V4i b = __builtin_nontemporal_load(src + t + stride);
        const V4i c = __builtin_nontemporal_load(src + t + 2 * stride);
        const V4i d = __builtin_nontemporal_load(src + t + 3 * stride);
        __builtin_nontemporal_store(a, dst + t);
        __builtin_nontemporal_store(b, dst + t + stride);
        __builtin_nontemporal_store(c, dst + t + 2 * stride);
        __builtin_nontemporal_store(d, dst + t + 3 * stride);
    }
    for (; t < n4; t += stride) dst[t] = src[t];
}

int mpx_copy_bench(int64_t bytes, double* ms) {
    if (bytes <= 0 || bytes % 16 != 0) return MPX_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPX_OK;
    try {
        MPX_HIP(hipMalloc(&a, bytes));
        MPX_HIP(hipMalloc(&b, bytes));
        MPX_HIP(hipMemset(a, 1, bytes));
        MPX_HIP(hipEventCreate(&e0));
        MPX_HIP(hipEventCreate(&e1));
        MPX_HIP(hipEventRecord(e0, nullptr));
        kCopyBench<<<4096, kBlock>>>(bytes / 16, static_cast<const int4*>(a), static_cast<int4*>(b));
        MPX_HIP(hipEventRecord(e1, nullptr));
        MPX_HIP(hipEventSynchronize(e1));
        float t = 0.0f;
        MPX_HIP(hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = t;
    } catch (const DeviceError&) {
        rc = MPX_ERR_DEVICE;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int mpx_device_synchronize(int32_t device) {
    if (hipSetDevice(device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        setError("mpx_device_synchronize failed");
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_pdlp_solve(const mpx_lp* lp, const mpx_pdlp_options* opts, const double* warm_x,
                   const double* warm_y, double* x_out, double* y_out, mpx_pdlp_result* res) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!validLp(lp) || !opts || !res) {
        setError("mpx_pdlp_solve: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (opts->termination_eval_frequency <= 0) {
        setError("mpx_pdlp_solve: termination_eval_frequency must be positive");
        return MPX_ERR_INVALID;
    }
    if (opts->anderson_acceleration && opts->anderson_window_size > kMaxAndersonWindow) {
        setError("mpx_pdlp_solve: anderson_window_size above 64");
        return MPX_ERR_INVALID;
    }
    std::memset(res, 0, sizeof(*res));
    const int m = lp->num_rows, n = lp->num_cols;
    const double sgn = lp->sense ? -1.0 : 1.0;
    auto finish = [&]() {
        res->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return MPX_OK;
    };
    // Closed-form trivial problems (pdlp.cpp:812-860): no device work needed.
    if (n == 0) {
        res->status = MPX_STATUS_OPTIMAL;
        res->objective = lp->obj_offset;
        if (y_out) std::memset(y_out, 0, sizeof(double) * static_cast<size_t>(m));
        return finish();
    }
    if (m == 0) {
        // Without rows every scale is 1 (Ruiz and rescaling need m > 0).
        if (x_out) std::memset(x_out, 0, sizeof(double) * static_cast<size_t>(n));
        double objv = lp->obj_offset;
        for (int j = 0; j < n; ++j) {
            const double c = sgn * lp->obj[j], lb = lp->col_lower[j], ub = lp->col_upper[j];
            double xj;
            if (c > 0.0) {
                if (!isFin(lb)) {
                    res->status = MPX_STATUS_UNBOUNDED;
                    return finish();
                }
                xj = lb;
            } else if (c < 0.0) {
                if (!isFin(ub)) {
                    res->status = MPX_STATUS_UNBOUNDED;
                    return finish();
                }
                xj = ub;
            } else {
                xj = isFin(lb) ? lb : (isFin(ub) ? ub : 0.0);
            }
            if (x_out) x_out[j] = xj;
            objv += lp->obj[j] * xj;
        }
        res->status = MPX_STATUS_OPTIMAL;
        res->objective = objv;
        return finish();
    }
    try {
        MPX_HIP(hipSetDevice(opts->device));
        std::unique_ptr<FilteredLp> filtered;
        if (opts->matrix_zero_tol > 0.0) filtered = std::make_unique<FilteredLp>(*lp, opts->matrix_zero_tol);
        SingleSolver s(filtered ? filtered->lp : *lp, *opts, lp);
        s.setup(warm_x, warm_y);
        res->setup_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s.solve(t0, x_out, y_out, res);
    } catch (const DeviceError& e) {
        res->status = MPX_STATUS_ERROR;
        res->objective = 0.0;
        if (e.code == hipErrorOutOfMemory) {
            finish();
            return MPX_ERR_OUT_OF_MEMORY;
        }
    } catch (const std::bad_alloc&) {
        setError("host out of memory");
        res->status = MPX_STATUS_ERROR;
    }
    return finish();
}

int mpx_pdlp_scale_problem(const mpx_lp* lp, const mpx_pdlp_options* opts,
                           mpx_scaled_problem* out) {
    if (!validLp(lp) || !opts || !out || lp->num_rows == 0 || lp->num_cols == 0) {
        setError("mpx_pdlp_scale_problem: invalid arguments");
        return MPX_ERR_INVALID;
    }
    if (opts->matrix_zero_tol > 0.0) {
        setError("mpx_pdlp_scale_problem: matrix_zero_tol changes the structure; not supported here");
        return MPX_ERR_UNSUPPORTED;
    }
    try {
        MPX_HIP(hipSetDevice(opts->device));
        SingleSolver s(*lp, *opts);
        s.setup(nullptr, nullptr);
        s.download(out);
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

int mpx_pdhg_iterate(const mpx_lp* lp, const mpx_scaled_problem* sp, int32_t iters,
                     int32_t inner_count, double step, double primal_weight, int32_t precision,
                     int32_t group, double* current_x, double* initial_x, double* pdhg_x,
                     double* current_y, double* initial_y, double* pdhg_y) {
    if (!validLp(lp) || !sp || iters < 0 || lp->num_rows == 0 || lp->num_cols == 0) {
        setError("mpx_pdhg_iterate: invalid arguments");
        return MPX_ERR_INVALID;
    }
    try {
        mpx_pdlp_options o;
        mpx_pdlp_default_options(&o);
        o.spmv_group = group;
        o.precision = precision;
        o.do_ruiz_scaling = 0;
        o.do_pock_chambolle_scaling = 0;
        o.do_bound_obj_rescaling = 0;
        o.sv_max_iter = 1;
        o.termination_eval_frequency = std::max(1, iters);
        MPX_HIP(hipSetDevice(0));
        SingleSolver s(*lp, o);
        s.setup(nullptr, nullptr);
        s.iterateFrom(sp, iters, inner_count, step, primal_weight, precision, current_x, initial_x,
                      pdhg_x, current_y, initial_y, pdhg_y);
    } catch (const DeviceError&) {
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

}  // extern "C"

template <typename T>
static int spmvImpl(int32_t rows, const int32_t* rs, const int32_t* ci, const T* val, const T* x,
                    T* y, int32_t group, void* stream) {
    if (rows < 0 || !rs || !y) return MPX_ERR_INVALID;
    if (rows == 0) return MPX_OK;
    const int g = group > 0 ? pickGroup(0, rows, group) : 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int grid = std::max(1, ceilDiv(static_cast<long long>(rows) * g, kBlock));
    withGroup(g, [&](auto GI) {
        constexpr int G = decltype(GI)::value;
        kSpmv<G, T><<<grid, kBlock, 0, st>>>(rows, rs, ci, val, x, y);
    });
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        setError(std::string("spmv launch: ") + hipGetErrorString(e));
        return MPX_ERR_DEVICE;
    }
    return MPX_OK;
}

extern "C" {

int mpx_spmv_csr_f64(int32_t rows, int32_t cols, const int32_t* rs, const int32_t* ci,
                     const double* val, const double* x, double* y, int32_t group, void* stream) {
    (void)cols;
    return spmvImpl<double>(rows, rs, ci, val, x, y, group, stream);
}

int mpx_spmv_csr_f32(int32_t rows, int32_t cols, const int32_t* rs, const int32_t* ci,
                     const float* val, const float* x, float* y, int32_t group, void* stream) {
    (void)cols;
    return spmvImpl<float>(rows, rs, ci, val, x, y, group, stream);
}

}  // extern "C"
