// pdlp_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement, in FP64, of mipx's CPU PDLP path.  It is the checker the
// parity tests compare the HIP library against; nothing in the product links
// it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it.
//
// It follows the reference CPU code operation by operation (ISO C++, built
// with -ffp-contract=off like the reference's CMAKE_CXX_EXTENSIONS OFF), so
// that per-element results are bit-identical to the reference's:
//   SparseMatrix::multiply / multiplyTranspose   src/lp/sparse_matrix.cpp:181-208
//   computeTerminationNorms / timeLimitReached   src/lp/pdlp.cpp:53-90
//   checkInfeasibilityCertificate                src/lp/pdlp.cpp:119-211
//   checkUnboundednessCertificate                src/lp/pdlp.cpp:220-284
//   buildScaledProblem                           src/lp/pdlp.cpp:469-669
//   buildTransposeCSR                            src/lp/pdlp.cpp:671-707
//   refreshPreconditioner                        src/lp/pdlp.cpp:713-742
//   estimateSpectralNorm                         src/lp/pdlp.cpp:748-795
//   solve (trivial cases, loop, restart/PI,
//          polish, recovery, work units)         src/lp/pdlp.cpp:801-1720
//   getReducedCosts                              src/lp/pdlp.cpp:2808-2819
//   AndersonAccelerator + its restarts           src/lp/pdlp.cpp:293-437, 1339-1378
//   adaptive step, CPU and GPU loops             src/lp/pdlp.cpp:1004-1040, 2271-2307
// orc_pdlp_solve follows the CPU loop; orc_pdlp_solve_gpu differs from it only
// in the adaptive step, which the GPU loop takes from x^ and A x^ at each
// evaluation instead of from x_r and A x_r at each iteration.
//
// Parity pins for this file live in tests/test_oracle.py: the reference's
// closed-form LPs (tests/test_pdlp.cpp), the SpMV known answers
// (tests/test_sparse_matrix.cpp:112-302), dense brute force, adjoint and
// scaling invariants, an SVD check of the power iteration, scipy-HiGHS optima
// of seeded LPs, and, where oracle/_ref was built, the reference itself.
//
// The long-double KKT checker orc_kkt_check() is an independent certificate:
// it evaluates relative primal residual, dual residual and duality gap of an
// original-space (x, y) pair without using any of the solver's internals.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "../include/mipx_pdlp.h"

namespace {

using Real = double;
using Index = int32_t;
constexpr Real kInf = std::numeric_limits<Real>::infinity();

inline bool isFin(Real v) { return std::isfinite(v); }

// --- CSR matrix with the reference's two SpMV loops -----------------------
struct Csr {
    Index rows = 0, cols = 0;
    std::vector<Index> starts, idx;
    std::vector<Real> val;

    // sparse_matrix.cpp:181-193: one sequential sum per row, storage order.
    void mul(const Real* x, Real* y) const {
        for (Index i = 0; i < rows; ++i) {
            Real acc = 0.0;
            for (Index k = starts[i]; k < starts[i + 1]; ++k) acc += val[k] * x[idx[k]];
            y[i] = acc;
        }
    }
    // sparse_matrix.cpp:195-208: row scatter in increasing row order, rows
    // with a zero multiplier skipped.
    void mulT(const Real* x, Real* y) const {
        std::fill(y, y + cols, 0.0);
        for (Index i = 0; i < rows; ++i) {
            const Real xi = x[i];
            if (xi == 0.0) continue;
            for (Index k = starts[i]; k < starts[i + 1]; ++k) y[idx[k]] += val[k] * xi;
        }
    }
};

Real sumsq(const std::vector<Real>& v) {
    Real s = 0.0;
    for (Real e : v) s += e * e;
    return s;
}
Real norm2(const std::vector<Real>& v) { return std::sqrt(sumsq(v)); }
Real normInf(const std::vector<Real>& v) {
    Real m = 0.0;
    for (Real e : v) m = std::max(m, std::abs(e));
    return m;
}

// --- original problem copy -------------------------------------------------
struct Problem {
    Index m = 0, n = 0;
    Csr a;
    std::vector<Real> obj, cl, cu, rl, ru;
    bool maximize = false;
    Real offset = 0.0;
};

Problem copyProblem(const mpx_lp& lp) {
    Problem p;
    p.m = lp.num_rows;
    p.n = lp.num_cols;
    p.a.rows = p.m;
    p.a.cols = p.n;
    p.a.starts.assign(lp.row_starts, lp.row_starts + p.m + 1);
    const Index nnz = p.m > 0 ? lp.row_starts[p.m] : 0;
    p.a.idx.assign(lp.col_indices, lp.col_indices + nnz);
    p.a.val.assign(lp.values, lp.values + nnz);
    p.obj.assign(lp.obj, lp.obj + p.n);
    p.cl.assign(lp.col_lower, lp.col_lower + p.n);
    p.cu.assign(lp.col_upper, lp.col_upper + p.n);
    p.rl.assign(lp.row_lower, lp.row_lower + p.m);
    p.ru.assign(lp.row_upper, lp.row_upper + p.m);
    p.maximize = lp.sense != 0;
    p.offset = lp.obj_offset;
    return p;
}

// pdlp.cpp:53-81.  ||c|| and the norm of the finite row bounds (rl counted
// only when it differs from ru).
void terminationNorms(const Problem& p, int norm_mode, Real* cnorm, Real* bnorm) {
    if (norm_mode == 1) {
        *cnorm = normInf(p.obj);
        Real b = 0.0;
        for (Index i = 0; i < p.m; ++i) {
            if (isFin(p.ru[i])) b = std::max(b, std::abs(p.ru[i]));
            if (isFin(p.rl[i]) && p.rl[i] != p.ru[i]) b = std::max(b, std::abs(p.rl[i]));
        }
        *bnorm = b;
        return;
    }
    *cnorm = norm2(p.obj);
    Real sq = 0.0;
    for (Index i = 0; i < p.m; ++i) {
        if (isFin(p.ru[i])) sq += p.ru[i] * p.ru[i];
        if (isFin(p.rl[i]) && p.rl[i] != p.ru[i]) sq += p.rl[i] * p.rl[i];
    }
    *bnorm = std::sqrt(sq);
}

// pdlp.cpp:119-211.  Farkas test on an original-space dual y.
bool farkasCertificate(const std::vector<Real>& y, const Problem& p, Real tol) {
    if (p.m == 0) return false;
    const Real ny = norm2(y);
    if (ny < 1e-15) return false;
    std::vector<Real> r(y);
    for (Real& v : r) v /= ny;
    for (Index i = 0; i < p.m; ++i) {
        const bool lo = isFin(p.rl[i]), hi = isFin(p.ru[i]);
        if (hi && !lo && r[i] < -tol) return false;
        if (lo && !hi && r[i] > tol) return false;
    }
    std::vector<Real> atr(static_cast<size_t>(p.n), 0.0);
    p.a.mulT(r.data(), atr.data());
    for (Index j = 0; j < p.n; ++j) {
        if (isFin(p.cl[j]) && atr[j] < -tol) return false;
        if (isFin(p.cu[j]) && atr[j] > tol) return false;
    }
    Real rowpart = 0.0;
    for (Index i = 0; i < p.m; ++i) {
        if (p.rl[i] == p.ru[i]) rowpart += r[i] * p.rl[i];
        else if (r[i] > 0.0 && isFin(p.ru[i])) rowpart += r[i] * p.ru[i];
        else if (r[i] < 0.0 && isFin(p.rl[i])) rowpart += r[i] * p.rl[i];
    }
    Real colpart = 0.0;
    for (Index j = 0; j < p.n; ++j) {
        if (atr[j] > 0.0 && isFin(p.cl[j])) colpart += atr[j] * p.cl[j];
        else if (atr[j] < 0.0 && isFin(p.cu[j])) colpart += atr[j] * p.cu[j];
    }
    return rowpart - colpart < -tol;
}

// pdlp.cpp:220-284.  Primal-ray test on an original-space x.
bool rayCertificate(const std::vector<Real>& x, const Problem& p, Real tol) {
    if (p.n == 0) return false;
    const Real nx = norm2(x);
    if (nx < 1e-15) return false;
    std::vector<Real> r(x);
    for (Real& v : r) v /= nx;
    for (Index j = 0; j < p.n; ++j) {
        if (isFin(p.cl[j]) && r[j] < -tol) return false;
        if (isFin(p.cu[j]) && r[j] > tol) return false;
    }
    std::vector<Real> ar(static_cast<size_t>(p.m), 0.0);
    p.a.mul(r.data(), ar.data());
    for (Index i = 0; i < p.m; ++i) {
        const bool lo = isFin(p.rl[i]), hi = isFin(p.ru[i]);
        if (lo && hi) {
            if (std::abs(ar[i]) > tol) return false;
        } else if (hi) {
            if (ar[i] > tol) return false;
        } else if (lo) {
            if (ar[i] < -tol) return false;
        }
    }
    const Real sgn = p.maximize ? -1.0 : 1.0;
    Real cr = 0.0;
    for (Index j = 0; j < p.n; ++j) cr += p.obj[j] * r[j];
    return sgn * cr < -tol;
}

// --- scaled problem (pdlp.cpp:469-707) -----------------------------------
struct Scaled {
    Real sgn = 1.0;
    Csr a, at;  // at: explicit A' in CSR (n rows), rows of A increasing per column
    std::vector<Real> c, cl, cu, rl, ru, rs, cs, sig, tau;
    Real os = 1.0, cscale = 1.0;
};

void pockChambolle(Scaled& s) {
    const Index m = s.a.rows, n = s.a.cols;
    std::vector<Real> rsq(static_cast<size_t>(m), 0.0), csq(static_cast<size_t>(n), 0.0);
    for (Index i = 0; i < m; ++i)
        for (Index k = s.a.starts[i]; k < s.a.starts[i + 1]; ++k) {
            const Real a = s.a.val[k];
            rsq[i] += a * a;
            csq[s.a.idx[k]] += a * a;
        }
    for (Index i = 0; i < m; ++i) s.sig[i] = 1.0 / std::max(std::sqrt(rsq[i]), 1e-8);
    for (Index j = 0; j < n; ++j) s.tau[j] = 1.0 / std::max(std::sqrt(csq[j]), 1e-8);
}

Scaled buildScaled(const Problem& p, const mpx_pdlp_options& o) {
    Scaled s;
    const Index m = p.m, n = p.n;
    s.sgn = p.maximize ? -1.0 : 1.0;
    s.c.resize(static_cast<size_t>(n));
    for (Index j = 0; j < n; ++j) s.c[j] = s.sgn * p.obj[j];
    s.cl = p.cl;
    s.cu = p.cu;
    s.rl = p.rl;
    s.ru = p.ru;
    s.rs.assign(static_cast<size_t>(m), 1.0);
    s.cs.assign(static_cast<size_t>(n), 1.0);
    Csr a = p.a;

    // pdlp.cpp:499-518: optional drop of |a| < matrix_zero_tol.
    if (o.matrix_zero_tol > 0.0 && !a.val.empty()) {
        Index w = 0;
        for (Index i = 0; i < m; ++i) {
            const Index b = a.starts[i], e = a.starts[i + 1];
            a.starts[i] = w;
            for (Index k = b; k < e; ++k)
                if (std::abs(a.val[k]) >= o.matrix_zero_tol) {
                    a.val[w] = a.val[k];
                    a.idx[w] = a.idx[k];
                    ++w;
                }
        }
        a.starts[m] = w;
        a.val.resize(static_cast<size_t>(w));
        a.idx.resize(static_cast<size_t>(w));
    }

    // pdlp.cpp:520-558: Ruiz, rows then columns, each by 1/sqrt(inf-norm).
    if (o.do_ruiz_scaling && m > 0 && n > 0 && !a.val.empty()) {
        std::vector<Real> cmax(static_cast<size_t>(n), 0.0);
        for (int it = 0; it < o.ruiz_iterations; ++it) {
            for (Index i = 0; i < m; ++i) {
                Real rmax = 0.0;
                for (Index k = a.starts[i]; k < a.starts[i + 1]; ++k)
                    rmax = std::max(rmax, std::abs(a.val[k]));
                if (rmax <= 1e-12) continue;
                const Real f = 1.0 / std::sqrt(rmax);
                s.rs[i] *= f;
                for (Index k = a.starts[i]; k < a.starts[i + 1]; ++k) a.val[k] *= f;
            }
            std::fill(cmax.begin(), cmax.end(), 0.0);
            for (size_t k = 0; k < a.val.size(); ++k)
                cmax[a.idx[k]] = std::max(cmax[a.idx[k]], std::abs(a.val[k]));
            for (Index j = 0; j < n; ++j)
                if (cmax[j] > 1e-12) s.cs[j] *= 1.0 / std::sqrt(cmax[j]);
            for (size_t k = 0; k < a.val.size(); ++k) {
                const Index j = a.idx[k];
                if (cmax[j] > 1e-12) a.val[k] *= 1.0 / std::sqrt(cmax[j]);
            }
        }
    }

    // pdlp.cpp:560-573.
    for (Index j = 0; j < n; ++j) {
        s.c[j] *= s.cs[j];
        if (isFin(s.cl[j])) s.cl[j] /= s.cs[j];
        if (isFin(s.cu[j])) s.cu[j] /= s.cs[j];
    }
    for (Index i = 0; i < m; ++i) {
        if (isFin(s.rl[i])) s.rl[i] *= s.rs[i];
        if (isFin(s.ru[i])) s.ru[i] *= s.rs[i];
    }
    s.a = std::move(a);

    // pdlp.cpp:671-707: counting transpose, rows of A visited in order.
    const Index nnz = static_cast<Index>(s.a.val.size());
    s.at.rows = n;
    s.at.cols = m;
    s.at.starts.assign(static_cast<size_t>(n + 1), 0);
    s.at.idx.resize(static_cast<size_t>(nnz));
    s.at.val.resize(static_cast<size_t>(nnz));
    if (nnz > 0) {
        for (Index k = 0; k < nnz; ++k) ++s.at.starts[s.a.idx[k] + 1];
        for (Index j = 1; j <= n; ++j) s.at.starts[j] += s.at.starts[j - 1];
        std::vector<Index> next(s.at.starts.begin(), s.at.starts.begin() + n);
        for (Index i = 0; i < m; ++i)
            for (Index k = s.a.starts[i]; k < s.a.starts[i + 1]; ++k) {
                const Index dst = next[s.a.idx[k]]++;
                s.at.idx[dst] = i;
                s.at.val[dst] = s.a.val[k];
            }
    }

    // pdlp.cpp:581-604.
    s.sig.assign(static_cast<size_t>(m), 1.0);
    s.tau.assign(static_cast<size_t>(n), 1.0);
    if (o.do_pock_chambolle_scaling && m > 0 && n > 0) pockChambolle(s);

    // pdlp.cpp:606-640: objective and bound rescaling.
    s.os = 1.0;
    s.cscale = 1.0;
    if (o.do_bound_obj_rescaling && m > 0 && n > 0) {
        s.os = 1.0 / (std::sqrt(sumsq(s.c)) + 1.0);
        Real bsq = 0.0;
        for (Index i = 0; i < m; ++i) {
            if (isFin(s.rl[i])) bsq += s.rl[i] * s.rl[i];
            if (isFin(s.ru[i])) bsq += s.ru[i] * s.ru[i];
        }
        s.cscale = 1.0 / (std::sqrt(bsq) + 1.0);
        for (Index j = 0; j < n; ++j) s.c[j] *= s.os;
        for (Index i = 0; i < m; ++i) {
            if (isFin(s.rl[i])) s.rl[i] *= s.cscale;
            if (isFin(s.ru[i])) s.ru[i] *= s.cscale;
        }
        for (Index j = 0; j < n; ++j) {
            if (isFin(s.cl[j])) s.cl[j] *= s.cscale;
            if (isFin(s.cu[j])) s.cu[j] *= s.cscale;
        }
    }
    return s;
}

// pdlp.cpp:748-795.  Power iteration on A'A with the reference's start vector
// (std::mt19937(42) + std::normal_distribution<double>).
Real spectralNorm(const Scaled& s, const mpx_pdlp_options& o, int* iters_out) {
    const Index n = s.a.cols, m = s.a.rows;
    if (iters_out) *iters_out = 0;
    if (n == 0 || m == 0) return 1.0;
    std::mt19937 gen(42);
    std::normal_distribution<Real> nd(0.0, 1.0);
    std::vector<Real> x(static_cast<size_t>(n));
    for (Real& v : x) v = nd(gen);
    Real xn = norm2(x);
    if (xn < 1e-15) xn = 1.0;
    for (Real& v : x) v /= xn;
    std::vector<Real> y(static_cast<size_t>(m)), xn_vec(static_cast<size_t>(n));
    Real s2 = 1.0;
    int it = 0;
    for (; it < o.sv_max_iter; ++it) {
        s.a.mul(x.data(), y.data());
        s.a.mulT(y.data(), xn_vec.data());
        s2 = 0.0;
        for (Index j = 0; j < n; ++j) s2 += x[j] * xn_vec[j];
        Real rsq = 0.0;
        for (Index j = 0; j < n; ++j) {
            const Real d = xn_vec[j] - s2 * x[j];
            rsq += d * d;
        }
        const Real res = std::sqrt(rsq);
        x.swap(xn_vec);
        xn = norm2(x);
        if (xn < 1e-15) break;
        for (Real& v : x) v /= xn;
        if (res < o.sv_tol * std::abs(s2)) break;
    }
    if (iters_out) *iters_out = it + 1 <= o.sv_max_iter ? it + 1 : o.sv_max_iter;
    return std::sqrt(std::max(s2, 1e-12));
}

// One Halpern PDHG iteration, pdlp.cpp:986-1071 (adaptive step excluded; the
// caller handles it between the two half steps).
struct Iterate {
    std::vector<Real> x, x0, xh, xr, y, y0, yh, yr;
    std::vector<Real> aty, axr;
};

inline void primalHalf(const Scaled& s, Iterate& st, Real step, Real pw, Real lambda) {
    const Index n = s.a.cols;
    s.a.mulT(st.y.data(), st.aty.data());
    for (Index j = 0; j < n; ++j) {
        const Real tau = step * s.tau[j] / pw;
        const Real t = st.x[j] - tau * (s.c[j] + st.aty[j]);
        st.xh[j] = std::clamp(t, s.cl[j], s.cu[j]);
        st.xr[j] = 2.0 * st.xh[j] - st.x[j];
        st.x[j] = lambda * st.xr[j] + (1.0 - lambda) * st.x0[j];
    }
    s.a.mul(st.xr.data(), st.axr.data());
}

inline void dualHalf(const Scaled& s, Iterate& st, Real step, Real pw, Real lambda) {
    const Index m = s.a.rows;
    for (Index i = 0; i < m; ++i) {
        const Real sig = step * s.sig[i] * pw;
        const Real v = st.y[i] + sig * st.axr[i];
        const Real lo = s.rl[i], hi = s.ru[i];
        Real yh;
        if (lo == hi) {
            yh = v - sig * lo;
        } else if (!isFin(lo)) {
            yh = std::max(0.0, v - sig * hi);
        } else if (!isFin(hi)) {
            yh = std::min(0.0, v - sig * lo);
        } else if (v >= sig * hi) {
            yh = v - sig * hi;
        } else if (v <= sig * lo) {
            yh = v - sig * lo;
        } else {
            yh = 0.0;
        }
        st.yh[i] = yh;
        st.yr[i] = 2.0 * yh - st.y[i];
        st.y[i] = lambda * st.yr[i] + (1.0 - lambda) * st.y0[i];
    }
}

// Evaluation at (xh, yh), pdlp.cpp:1076-1202.
struct Eval {
    Real fpe = 0, rel_p = 0, rel_d = 0, rel_p_inf = 0, rel_d_inf = 0, gap = 0, pobj = 0, dobj = 0;
    Real dx2 = 0, dy2 = 0;
};

Eval evaluate(const Problem& p, const Scaled& s, Iterate& st, Real step, Real pw, Real cnorm,
              Real bnorm, std::vector<Real>& ax, std::vector<Real>& dy, std::vector<Real>& atdy,
              std::vector<Real>& atyh) {
    const Index m = s.a.rows, n = s.a.cols;
    Eval e;
    Real dx2 = 0.0, dy2 = 0.0;
    for (Index j = 0; j < n; ++j) {
        const Real d = st.xh[j] - st.x0[j];
        dx2 += d * d;
    }
    for (Index i = 0; i < m; ++i) {
        const Real d = st.yh[i] - st.y0[i];
        dy2 += d * d;
    }
    e.dx2 = dx2;
    e.dy2 = dy2;
    const Real movement = dx2 * pw + dy2 / pw;
    for (Index i = 0; i < m; ++i) dy[i] = st.yh[i] - st.y0[i];
    s.a.mulT(dy.data(), atdy.data());
    Real inter = 0.0;
    for (Index j = 0; j < n; ++j) inter += atdy[j] * (st.xh[j] - st.x0[j]);
    inter *= 2.0 * step;
    e.fpe = std::sqrt(std::max(movement + inter, 0.0));

    s.a.mul(st.xh.data(), ax.data());
    Real psq = 0.0, pmax = 0.0;
    for (Index i = 0; i < m; ++i) {
        const Real vs = ax[i] - std::clamp(ax[i], s.rl[i], s.ru[i]);
        const Real v = vs / (s.cscale * s.rs[i]);
        psq += v * v;
        pmax = std::max(pmax, std::abs(v));
    }
    e.rel_p = std::sqrt(psq) / (1.0 + bnorm);
    e.rel_p_inf = pmax / (1.0 + bnorm);

    s.a.mulT(st.yh.data(), atyh.data());
    Real dsq = 0.0, dmax = 0.0;
    for (Index j = 0; j < n; ++j) {
        const Real g = s.c[j] + atyh[j];
        const Real xj = st.xh[j], lb = s.cl[j], ub = s.cu[j];
        Real dr;
        if (xj <= lb + 1e-12 && isFin(lb)) dr = std::min(g, 0.0);
        else if (xj >= ub - 1e-12 && isFin(ub)) dr = std::max(g, 0.0);
        else dr = g;
        const Real d = dr / (s.os * s.cs[j]);
        dsq += d * d;
        dmax = std::max(dmax, std::abs(d));
    }
    e.rel_d = std::sqrt(dsq) / (1.0 + cnorm);
    e.rel_d_inf = dmax / (1.0 + cnorm);

    Real ps = 0.0;
    for (Index j = 0; j < n; ++j) ps += s.c[j] * st.xh[j];
    const Real unscale = 1.0 / (s.os * s.cscale);
    e.pobj = p.offset + s.sgn * ps * unscale;
    Real dcol = 0.0;
    for (Index j = 0; j < n; ++j) {
        Real g = s.c[j] + atyh[j];
        const Real lb = s.cl[j], ub = s.cu[j];
        if (!isFin(lb)) g = std::min(g, 0.0);
        if (!isFin(ub)) g = std::max(g, 0.0);
        if (g > 0.0 && isFin(lb)) dcol += g * lb;
        else if (g < 0.0 && isFin(ub)) dcol += g * ub;
    }
    Real drow = 0.0;
    for (Index i = 0; i < m; ++i) {
        Real yi = st.yh[i];
        const Real lo = s.rl[i], hi = s.ru[i];
        if (!isFin(hi)) yi = std::min(yi, 0.0);
        if (!isFin(lo)) yi = std::max(yi, 0.0);
        if (yi > 0.0 && isFin(hi)) drow -= yi * hi;
        else if (yi < 0.0 && isFin(lo)) drow -= yi * lo;
    }
    e.dobj = p.offset + s.sgn * (dcol + drow) * unscale;
    e.gap = std::abs(e.pobj - e.dobj) / (1.0 + std::abs(e.pobj) + std::abs(e.dobj));
    return e;
}

// Restart test, pdlp.cpp:1264-1277 (main loop) / 1607-1618 (polish).
bool restartDue(bool first_batch, Real fpe, Real fpe0, Real last_fpe, int inner, int count,
                const mpx_pdlp_options& o) {
    if (first_batch) return true;
    if (fpe0 > 0.0 && fpe <= o.restart_sufficient_decay * fpe0) return true;
    if (fpe0 > 0.0 && fpe <= o.restart_necessary_decay * fpe0 && fpe > last_fpe) return true;
    return inner >= static_cast<int>(o.restart_artificial_fraction * static_cast<Real>(count));
}

struct PiState {
    Real err_sum = 0.0, last_err = 0.0, best_pw = 1.0, best_score = kInf;
};

// PI primal-weight update, pdlp.cpp:1281-1323 (main) and 1626-1651 (polish:
// no best-weight tracking and no fallback branch).
void updatePrimalWeight(Real& pw, PiState& pi, Real pdist, Real ddist, Real rel_p, Real rel_d,
                        bool polish, const mpx_pdlp_options& o) {
    const Real ratio = (rel_p > 1e-15) ? rel_d / rel_p : 0.0;
    if (pdist > 1e-16 && ddist > 1e-16 && pdist < 1e12 && ddist < 1e12 && ratio > 1e-8 &&
        ratio < 1e8) {
        const Real err = std::log(ddist) - std::log(pdist) - std::log(pw);
        pi.err_sum = o.pid_i_smooth * pi.err_sum + err;
        const Real derr = err - pi.last_err;
        pw *= std::exp(o.pid_kp * err + o.pid_ki * pi.err_sum + o.pid_kd * derr);
        pw = std::clamp(pw, 1e-4, 1e8);
        pi.last_err = err;
        if (!polish) {
            const Real score = std::abs(std::log10(std::max(rel_d, 1e-15)) -
                                        std::log10(std::max(rel_p, 1e-15)));
            if (score < pi.best_score) {
                pi.best_score = score;
                pi.best_pw = pw;
            }
        }
    } else if (!polish) {
        pw = pi.best_pw;
        pi.err_sum = 0.0;
        pi.last_err = 0.0;
    }
}

// Type-I Anderson acceleration over restart points, pdlp.cpp:293-437 (the
// accelerator) as applied at main-loop restarts, pdlp.cpp:1339-1378 (CPU) and
// 2418-2466 (GPU, same arithmetic).  The window holds the last `w` restart
// candidates z = [x^; y^] and residuals r = z - [x0; y0]; extrapolation
// solves (R'R + mu I) a = 1 by Cholesky (mu = reg * trace / d), normalises
// sum(a) = 1 and accepts sum(a_i z_i) only if its residual sum(a_i r_i) is
// shorter than the newest one.  Every sum runs in index order.
struct AaWindow {
    int w = 0, count = 0;
    size_t dim = 0;
    Real reg = 0.0;
    std::vector<std::vector<Real>> z, r;

    void init(int window, size_t d, Real regularization) {
        w = window;
        dim = d;
        reg = regularization;
        count = 0;
        z.assign(static_cast<size_t>(w), std::vector<Real>(dim, 0.0));
        r.assign(static_cast<size_t>(w), std::vector<Real>(dim, 0.0));
    }
    void push(const Iterate& st) {
        const size_t n = st.xh.size(), m = st.yh.size();
        std::vector<Real>& zs = z[static_cast<size_t>(count % w)];
        std::vector<Real>& rs = r[static_cast<size_t>(count % w)];
        for (size_t j = 0; j < n; ++j) {
            zs[j] = st.xh[j];
            rs[j] = st.xh[j] - st.x0[j];
        }
        for (size_t i = 0; i < m; ++i) {
            zs[n + i] = st.yh[i];
            rs[n + i] = st.yh[i] - st.y0[i];
        }
        ++count;
    }
    bool extrapolate(std::vector<Real>& out) const {
        const int d = std::min(count, w);
        if (d < 2) return false;
        auto slot = [&](int i) { return static_cast<size_t>((count - d + i) % w); };
        std::vector<Real> g(static_cast<size_t>(d * d), 0.0);
        for (int i = 0; i < d; ++i)
            for (int j = i; j < d; ++j) {
                const std::vector<Real>&a = r[slot(i)], &b = r[slot(j)];
                Real acc = 0.0;
                for (size_t k = 0; k < dim; ++k) acc += a[k] * b[k];
                g[i * d + j] = g[j * d + i] = acc;
            }
        Real tr = 0.0;
        for (int i = 0; i < d; ++i) tr += g[i * d + i];
        const Real mu = reg * tr / static_cast<Real>(d);
        for (int i = 0; i < d; ++i) g[i * d + i] += mu;
        std::vector<Real> l(static_cast<size_t>(d * d), 0.0);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j <= i; ++j) {
                Real v = g[i * d + j];
                for (int q = 0; q < j; ++q) v -= l[i * d + q] * l[j * d + q];
                if (i == j) {
                    if (v <= 0.0) return false;
                    l[i * d + i] = std::sqrt(v);
                } else {
                    l[i * d + j] = v / l[j * d + j];
                }
            }
        std::vector<Real> t(static_cast<size_t>(d)), a(static_cast<size_t>(d));
        for (int i = 0; i < d; ++i) {
            Real v = 1.0;
            for (int q = 0; q < i; ++q) v -= l[i * d + q] * t[q];
            t[i] = v / l[i * d + i];
        }
        for (int i = d - 1; i >= 0; --i) {
            Real v = t[i];
            for (int q = i + 1; q < d; ++q) v -= l[q * d + i] * a[q];
            a[i] = v / l[i * d + i];
        }
        Real sa = 0.0;
        for (int i = 0; i < d; ++i) sa += a[i];
        if (std::abs(sa) < 1e-15) return false;
        for (int i = 0; i < d; ++i) a[i] /= sa;
        out.assign(dim, 0.0);
        for (int i = 0; i < d; ++i) {
            const std::vector<Real>& zi = z[slot(i)];
            for (size_t k = 0; k < dim; ++k) out[k] += a[i] * zi[k];
        }
        Real ext = 0.0;
        for (size_t k = 0; k < dim; ++k) {
            Real v = 0.0;
            for (int i = 0; i < d; ++i) v += a[i] * r[slot(i)][k];
            ext += v * v;
        }
        const std::vector<Real>& rl = r[slot(d - 1)];
        Real last = 0.0;
        for (size_t k = 0; k < dim; ++k) last += rl[k] * rl[k];
        return ext < last;
    }
};

bool timeUp(std::chrono::steady_clock::time_point t0, Real limit) {
    if (limit < 0.0) return false;
    return std::chrono::duration<Real>(std::chrono::steady_clock::now() - t0).count() >= limit;
}

void fillResult(mpx_pdlp_result* r, int status, int iters, Real obj, Real work) {
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->status = status;
    r->iterations = iters;
    r->objective = obj;
    r->work_units = work;
}

}  // namespace

extern "C" {

// y = A x and y = A' x with the reference's CPU loops.
void orc_spmv(int32_t m, int32_t n, const int32_t* rs, const int32_t* ci, const double* v,
              const double* x, double* y) {
    Csr a;
    a.rows = m;
    a.cols = n;
    a.starts.assign(rs, rs + m + 1);
    a.idx.assign(ci, ci + rs[m]);
    a.val.assign(v, v + rs[m]);
    a.mul(x, y);
}

void orc_spmv_t(int32_t m, int32_t n, const int32_t* rs, const int32_t* ci, const double* v,
                const double* x, double* y) {
    Csr a;
    a.rows = m;
    a.cols = n;
    a.starts.assign(rs, rs + m + 1);
    a.idx.assign(ci, ci + rs[m]);
    a.val.assign(v, v + rs[m]);
    a.mulT(x, y);
}

// Scaled problem + power iteration, same output layout as the library's
// mpx_pdlp_scale_problem.
int orc_scale_problem(const mpx_lp* lp, const mpx_pdlp_options* o, mpx_scaled_problem* out) {
    if (!lp || !o || !out) return MPX_ERR_INVALID;
    const Problem p = copyProblem(*lp);
    const Scaled s = buildScaled(p, *o);
    auto put = [](double* dst, const std::vector<Real>& src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    auto puti = [](int32_t* dst, const std::vector<Index>& src) {
        if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(out->values, s.a.val);
    put(out->at_values, s.at.val);
    puti(out->at_row_starts, s.at.starts);
    puti(out->at_col_indices, s.at.idx);
    put(out->cscaled, s.c);
    put(out->col_lower, s.cl);
    put(out->col_upper, s.cu);
    put(out->row_lower, s.rl);
    put(out->row_upper, s.ru);
    put(out->row_scale, s.rs);
    put(out->col_scale, s.cs);
    put(out->sigma_base, s.sig);
    put(out->tau_base, s.tau);
    out->objective_scale = s.os;
    out->constraint_scale = s.cscale;
    int pit = 0;
    out->spectral_norm = spectralNorm(s, *o, &pit);
    out->power_iterations = pit;
    return MPX_OK;
}

// `iters` Halpern iterations from the given scaled state (same contract as
// the library's mpx_pdhg_iterate; the scaled data comes from `sp`).
int orc_pdhg_iterate(const mpx_lp* lp, const mpx_scaled_problem* sp, int32_t iters,
                     int32_t inner_count, double step, double pw, double* cx, double* ix,
                     double* px, double* cy, double* iy, double* py) {
    const Index m = lp->num_rows, n = lp->num_cols, nnz = lp->row_starts[m];
    Scaled s;
    s.a.rows = m;
    s.a.cols = n;
    s.a.starts.assign(lp->row_starts, lp->row_starts + m + 1);
    s.a.idx.assign(lp->col_indices, lp->col_indices + nnz);
    s.a.val.assign(sp->values, sp->values + nnz);
    s.c.assign(sp->cscaled, sp->cscaled + n);
    s.cl.assign(sp->col_lower, sp->col_lower + n);
    s.cu.assign(sp->col_upper, sp->col_upper + n);
    s.rl.assign(sp->row_lower, sp->row_lower + m);
    s.ru.assign(sp->row_upper, sp->row_upper + m);
    s.sig.assign(sp->sigma_base, sp->sigma_base + m);
    s.tau.assign(sp->tau_base, sp->tau_base + n);
    Iterate st;
    st.x.assign(cx, cx + n);
    st.x0.assign(ix, ix + n);
    st.xh.assign(px, px + n);
    st.xr.assign(static_cast<size_t>(n), 0.0);
    st.y.assign(cy, cy + m);
    st.y0.assign(iy, iy + m);
    st.yh.assign(py, py + m);
    st.yr.assign(static_cast<size_t>(m), 0.0);
    st.aty.assign(static_cast<size_t>(n), 0.0);
    st.axr.assign(static_cast<size_t>(m), 0.0);
    for (int k = 1; k <= iters; ++k) {
        const Real lambda =
            static_cast<Real>(inner_count + k) / static_cast<Real>(inner_count + k + 1);
        primalHalf(s, st, step, pw, lambda);
        dualHalf(s, st, step, pw, lambda);
    }
    std::copy(st.x.begin(), st.x.end(), cx);
    std::copy(st.xh.begin(), st.xh.end(), px);
    std::copy(st.y.begin(), st.y.end(), cy);
    std::copy(st.yh.begin(), st.yh.end(), py);
    (void)ix;
    (void)iy;
    return MPX_OK;
}

// The whole CPU PDLP solve, pdlp.cpp:801-1720.  With gpu_path the adaptive
// step follows the GPU host loop instead (pdlp.cpp:2271-2307: x^ and A x^ at
// evaluation boundaries rather than x_r and A x_r every iteration); nothing
// else differs between the two paths' results.
static int solveImpl(const mpx_lp* lp, const mpx_pdlp_options* opts, const double* warm_x,
                     const double* warm_y, double* x_out, double* y_out, mpx_pdlp_result* res,
                     bool gpu_path) {
    if (!lp || !opts) return MPX_ERR_INVALID;
    const mpx_pdlp_options& o = *opts;
    const auto t0 = std::chrono::steady_clock::now();
    const Problem p = copyProblem(*lp);
    Scaled s = buildScaled(p, o);
    const Index m = p.m, n = p.n;
    std::vector<Real> xo(static_cast<size_t>(n), 0.0), yo(static_cast<size_t>(m), 0.0);
    auto emit = [&](int status, int iters, Real obj, Real work) {
        if (x_out) std::copy(xo.begin(), xo.end(), x_out);
        if (y_out) std::copy(yo.begin(), yo.end(), y_out);
        fillResult(res, status, iters, obj, work);
        if (res) res->seconds =
            std::chrono::duration<Real>(std::chrono::steady_clock::now() - t0).count();
        return MPX_OK;
    };

    // pdlp.cpp:812-860: trivial problems.
    if (n == 0) return emit(MPX_STATUS_OPTIMAL, 0, p.offset, 0.0);
    if (m == 0) {
        Real objv = p.offset;
        for (Index j = 0; j < n; ++j) {
            const Real c = s.c[j], lb = s.cl[j], ub = s.cu[j];
            Real xj;
            if (c > 0.0) {
                if (!isFin(lb)) return emit(MPX_STATUS_UNBOUNDED, 0, 0.0, 0.0);
                xj = lb;
            } else if (c < 0.0) {
                if (!isFin(ub)) return emit(MPX_STATUS_UNBOUNDED, 0, 0.0, 0.0);
                xj = ub;
            } else {
                xj = isFin(lb) ? lb : (isFin(ub) ? ub : 0.0);
            }
            xo[j] = xj * s.cs[j] / s.cscale;
            objv += p.obj[j] * xo[j];
        }
        return emit(MPX_STATUS_OPTIMAL, 0, objv, 0.0);
    }

    const Real nnzd = static_cast<Real>(s.a.val.size());
    auto work = [&](int it) {
        return static_cast<double>(it) * (4.0 * nnzd + 2.0 * static_cast<double>(m + n));
    };

    int pit = 0;
    Real step = 0.998 / spectralNorm(s, o, &pit);
    const bool adaptive = o.adaptive_step != 0;
    const Real step_lb = step * o.adaptive_step_min, step_ub = step * o.adaptive_step_max;
    Real pw = o.primal_weight;
    const int N = o.termination_eval_frequency;

    Iterate st;
    st.x.assign(static_cast<size_t>(n), 0.0);
    st.x0 = st.x;
    st.xh = st.x;
    st.xr = st.x;
    st.aty = st.x;
    st.y.assign(static_cast<size_t>(m), 0.0);
    st.y0 = st.y;
    st.yh = st.y;
    st.yr = st.y;
    st.axr = st.y;
    std::vector<Real> ax(static_cast<size_t>(m)), dyv(static_cast<size_t>(m)),
        atdy(static_cast<size_t>(n)), atyh(static_cast<size_t>(n));
    bool have_prev = false;
    AaWindow aa;
    std::vector<Real> aa_out;
    const bool use_aa = o.anderson_acceleration && o.anderson_window_size >= 2;
    if (use_aa)
        aa.init(o.anderson_window_size, static_cast<size_t>(n) + static_cast<size_t>(m),
                o.anderson_regularization);

    // pdlp.cpp:915-930: warm start mapped into the scaled space.
    if (warm_x && warm_y) {
        for (Index j = 0; j < n; ++j) st.x[j] = st.x0[j] = warm_x[j] / s.cs[j] * s.cscale;
        for (Index i = 0; i < m; ++i) st.y[i] = st.y0[i] = warm_y[i] * s.os / (s.sgn * s.rs[i]);
    }

    int inner = 0, total = 0, restarts = 0;
    Real fpe0 = -1.0, last_fpe = kInf;
    PiState pi;
    pi.best_pw = pw;
    Real cnorm = 0.0, bnorm = 0.0;
    terminationNorms(p, o.optimality_norm, &cnorm, &bnorm);
    bool converged = false;
    Real precond_pw = o.primal_weight;  // pdlp.cpp:668
    Eval e;
    auto stopRequested = [&]() { return o.stop_fn != nullptr && o.stop_fn(o.stop_user) != 0; };

    // Malitsky-Tam style step adaptation from two consecutive points (z, A z):
    // pdlp.cpp:1004-1040 (CPU: z = x_r every iteration) and 2271-2307 (GPU:
    // z = x^ at every evaluation).
    std::vector<Real> prev_az, prev_z;
    if (adaptive) {
        prev_az.assign(static_cast<size_t>(m), 0.0);
        prev_z.assign(static_cast<size_t>(n), 0.0);
    }
    auto adaptStep = [&](const std::vector<Real>& az, const std::vector<Real>& z) {
        if (have_prev) {
            Real adiff = 0.0;
            for (Index i = 0; i < m; ++i) {
                const Real d = az[i] - prev_az[i];
                adiff += s.sig[i] * d * d;
            }
            Real xdiff = 0.0;
            for (Index j = 0; j < n; ++j) {
                const Real d = z[j] - prev_z[j];
                xdiff += d * d / s.tau[j];
            }
            if (xdiff > 1e-30) {
                const Real lip2 = adiff / xdiff;
                const Real local = 0.998 / std::sqrt(std::max(lip2, 1e-30));
                const Real ns = step > local ? step * o.adaptive_step_down : step * o.adaptive_step_up;
                step = std::clamp(ns, step_lb, step_ub);
            }
        }
        prev_az = az;
        prev_z = z;
        have_prev = true;
    };

    auto oneIteration = [&](Real lambda) {
        primalHalf(s, st, step, pw, lambda);
        if (adaptive && !gpu_path) adaptStep(st.axr, st.xr);
        dualHalf(s, st, step, pw, lambda);
    };

    while (total < o.max_iter) {
        if (timeUp(t0, o.max_solve_seconds)) return emit(MPX_STATUS_TIME_LIMIT, total, 0.0, work(total));
        if (stopRequested()) return emit(MPX_STATUS_ITER_LIMIT, total, 0.0, work(total));
        const int batch = std::min(N, o.max_iter - total);
        for (int k = 1; k <= batch; ++k) {
            if (timeUp(t0, o.max_solve_seconds))
                return emit(MPX_STATUS_TIME_LIMIT, total + k - 1, 0.0, work(total + k - 1));
            oneIteration(static_cast<Real>(inner + k) / static_cast<Real>(inner + k + 1));
        }
        inner += batch;
        total += batch;
        e = evaluate(p, s, st, step, pw, cnorm, bnorm, ax, dyv, atdy, atyh);
        if (fpe0 < 0.0) fpe0 = e.fpe;
        const Real tp = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
        const Real td = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
        if (o.verbose && (total <= N || total % (10 * N) == 0))
            std::printf("ORACLE %7d  pobj=% .8e  pinf=% .2e  dinf=% .2e  gap=% .2e  fpe=% .2e  pw=% .2e\n",
                        total, e.pobj, tp, td, e.gap, e.fpe, pw);
        if (tp <= o.primal_tol && td <= o.dual_tol && e.gap <= o.optimality_tol) {
            converged = true;
            break;
        }
        if (adaptive && gpu_path) adaptStep(ax, st.xh);
        // pdlp.cpp:1220-1262: certificates every 10 N iterations.  The
        // reference adds elapsed seconds * 1e-6 to the work units here.
        if (total >= 10 * N && total % (10 * N) == 0) {
            std::vector<Real> yorig(static_cast<size_t>(m));
            const Real inv_os = 1.0 / s.os;
            for (Index i = 0; i < m; ++i) yorig[i] = s.sgn * st.yh[i] * s.rs[i] * inv_os;
            if (farkasCertificate(yorig, p, 1e-6)) {
                yo = yorig;
                const double secs =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                return emit(MPX_STATUS_INFEASIBLE, total, 0.0, work(total) + secs * 1e-6);
            }
            std::vector<Real> xorig(static_cast<size_t>(n));
            const Real inv_cs = 1.0 / s.cscale;
            for (Index j = 0; j < n; ++j) xorig[j] = st.xh[j] * s.cs[j] * inv_cs;
            if (rayCertificate(xorig, p, 1e-6)) {
                xo = xorig;
                const double secs =
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                return emit(MPX_STATUS_UNBOUNDED, total, 0.0, work(total) + secs * 1e-6);
            }
        }
        const bool restart = restartDue(total == N, e.fpe, fpe0, last_fpe, inner, total, o);
        last_fpe = e.fpe;
        if (restart) {
            ++restarts;
            if (o.update_primal_weight)
                updatePrimalWeight(pw, pi, std::sqrt(e.dx2), std::sqrt(e.dy2), e.rel_p, e.rel_d,
                                   false, o);
            // pdlp.cpp:1326-1340 / 713-742: preconditioner refresh on
            // primal-weight drift.  The refresh recomputes sigma/tau from the
            // unchanged scaled matrix, so only the recorded weight moves.
            if (o.preconditioner_refresh && precond_pw > 0.0) {
                const Real drift = pw / precond_pw;
                if (drift > o.preconditioner_refresh_ratio ||
                    drift < 1.0 / o.preconditioner_refresh_ratio) {
                    if (o.do_pock_chambolle_scaling) {
                        pockChambolle(s);
                        precond_pw = pw;
                    }
                }
            }
            bool extrapolated = false;
            if (use_aa) {
                aa.push(st);
                extrapolated = aa.extrapolate(aa_out);
            }
            if (extrapolated) {
                for (Index j = 0; j < n; ++j)
                    st.x0[j] = st.x[j] = std::clamp(aa_out[static_cast<size_t>(j)], s.cl[j], s.cu[j]);
                for (Index i = 0; i < m; ++i) st.y0[i] = st.y[i] = aa_out[static_cast<size_t>(n + i)];
            } else {
                st.x0 = st.xh;
                st.x = st.xh;
                st.y0 = st.yh;
                st.y = st.yh;
            }
            inner = 0;
            fpe0 = -1.0;
            last_fpe = kInf;
            have_prev = false;
        }
    }

    // pdlp.cpp:1407-1675: polish.
    if (converged && o.enable_polishing) {
        const Real rp = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
        const Real rd = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
        if ((rp > o.polish_primal_tol || rd > o.polish_dual_tol) && !timeUp(t0, o.max_solve_seconds)) {
            st.x0 = st.xh;
            st.x = st.xh;
            st.y0 = st.yh;
            st.y = st.yh;
            inner = 0;
            fpe0 = -1.0;
            last_fpe = kInf;
            int pcount = 0;
            while (pcount < o.polish_max_iter) {
                if (timeUp(t0, o.max_solve_seconds)) break;
                if (stopRequested()) break;
                const int batch = std::min(N, o.polish_max_iter - pcount);
                for (int k = 1; k <= batch; ++k) {
                    if (timeUp(t0, o.max_solve_seconds)) break;
                    const Real lambda =
                        static_cast<Real>(inner + k) / static_cast<Real>(inner + k + 1);
                    primalHalf(s, st, step, pw, lambda);
                    dualHalf(s, st, step, pw, lambda);
                }
                inner += batch;
                pcount += batch;
                total += batch;
                e = evaluate(p, s, st, step, pw, cnorm, bnorm, ax, dyv, atdy, atyh);
                if (fpe0 < 0.0) fpe0 = e.fpe;
                const Real rpn = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
                const Real rdn = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
                if (rpn <= o.polish_primal_tol && rdn <= o.polish_dual_tol) break;
                const bool restart =
                    restartDue(pcount == batch, e.fpe, fpe0, last_fpe, inner, pcount, o);
                last_fpe = e.fpe;
                if (restart) {
                    ++restarts;
                    if (o.update_primal_weight)
                        updatePrimalWeight(pw, pi, std::sqrt(e.dx2), std::sqrt(e.dy2), e.rel_p,
                                           e.rel_d, true, o);
                    st.x0 = st.xh;
                    st.x = st.xh;
                    st.y0 = st.yh;
                    st.y = st.yh;
                    inner = 0;
                    fpe0 = -1.0;
                    last_fpe = kInf;
                }
            }
        }
    }

    Real objv = 0.0;
    int status;
    if (converged) {
        // pdlp.cpp:1677-1704: recovery.
        const Real inv_cs = 1.0 / s.cscale;
        for (Index j = 0; j < n; ++j) {
            Real xj = st.xh[j] * s.cs[j] * inv_cs;
            if (isFin(p.cl[j])) xj = std::max(xj, p.cl[j]);
            if (isFin(p.cu[j])) xj = std::min(xj, p.cu[j]);
            xo[j] = xj;
        }
        const Real inv_os = 1.0 / s.os;
        for (Index i = 0; i < m; ++i) yo[i] = s.sgn * st.yh[i] * s.rs[i] * inv_os;
        objv = p.offset;
        for (Index j = 0; j < n; ++j) objv += p.obj[j] * xo[j];
        status = MPX_STATUS_OPTIMAL;
    } else {
        status = timeUp(t0, o.max_solve_seconds) ? MPX_STATUS_TIME_LIMIT : MPX_STATUS_ITER_LIMIT;
    }
    emit(status, total, objv, work(total));
    if (res) {
        res->rel_primal = o.optimality_norm == 1 ? e.rel_p_inf : e.rel_p;
        res->rel_dual = o.optimality_norm == 1 ? e.rel_d_inf : e.rel_d;
        res->rel_gap = e.gap;
        res->primal_obj = e.pobj;
        res->dual_obj = e.dobj;
        res->step = step;
        res->primal_weight = pw;
        res->restarts = restarts;
    }
    return MPX_OK;
}

int orc_pdlp_solve(const mpx_lp* lp, const mpx_pdlp_options* opts, const double* warm_x,
                   const double* warm_y, double* x_out, double* y_out, mpx_pdlp_result* res) {
    return solveImpl(lp, opts, warm_x, warm_y, x_out, y_out, res, false);
}

// The same solve with the GPU host loop's adaptive step (pdlp.cpp:2271-2307).
int orc_pdlp_solve_gpu(const mpx_lp* lp, const mpx_pdlp_options* opts, const double* warm_x,
                       const double* warm_y, double* x_out, double* y_out, mpx_pdlp_result* res) {
    return solveImpl(lp, opts, warm_x, warm_y, x_out, y_out, res, true);
}

// Reduced costs c + A'y in original space, pdlp.cpp:2808-2819.
void orc_reduced_costs(const mpx_lp* lp, const double* y, double* rc) {
    const Index m = lp->num_rows, n = lp->num_cols;
    Csr a;
    a.rows = m;
    a.cols = n;
    a.starts.assign(lp->row_starts, lp->row_starts + m + 1);
    a.idx.assign(lp->col_indices, lp->col_indices + lp->row_starts[m]);
    a.val.assign(lp->values, lp->values + lp->row_starts[m]);
    std::vector<Real> aty(static_cast<size_t>(n), 0.0);
    a.mulT(y, aty.data());
    for (Index j = 0; j < n; ++j) rc[j] = lp->obj[j] + aty[j];
}

// Independent long-double KKT certificate of an original-space pair (x, y)
// in the reference's sign convention (reduced costs c + A'y >= 0 at lower
// bounds for minimisation).  out[0..7] = rel_primal (L2), rel_dual (L2, cone
// form: the part of the normalised reduced cost that no bound can absorb),
// rel_gap, primal objective, dual objective, max bound violation of x,
// max primal row violation (absolute), max dual-sign violation of y.
void orc_kkt_check(const mpx_lp* lp, const double* x, const double* y, double* out) {
    using LD = long double;
    const Index m = lp->num_rows, n = lp->num_cols;
    const LD sgn = lp->sense ? -1.0L : 1.0L;
    LD bsq = 0, csq = 0;
    for (Index i = 0; i < m; ++i) {
        const LD rl = lp->row_lower[i], ru = lp->row_upper[i];
        if (std::isfinite(lp->row_upper[i])) bsq += ru * ru;
        if (std::isfinite(lp->row_lower[i]) && lp->row_lower[i] != lp->row_upper[i]) bsq += rl * rl;
    }
    for (Index j = 0; j < n; ++j) csq += static_cast<LD>(lp->obj[j]) * lp->obj[j];
    const LD bnorm = std::sqrt(bsq), cnorm = std::sqrt(csq);
    std::vector<LD> ax(static_cast<size_t>(m), 0.0L), aty(static_cast<size_t>(n), 0.0L);
    for (Index i = 0; i < m; ++i)
        for (Index k = lp->row_starts[i]; k < lp->row_starts[i + 1]; ++k) {
            ax[i] += static_cast<LD>(lp->values[k]) * x[lp->col_indices[k]];
            aty[lp->col_indices[k]] += static_cast<LD>(lp->values[k]) * y[i];
        }
    LD psq = 0, pmax = 0, ysignmax = 0;
    LD drow = 0;
    for (Index i = 0; i < m; ++i) {
        const LD lo = lp->row_lower[i], hi = lp->row_upper[i];
        LD v = 0;
        if (ax[i] < lo) v = ax[i] - lo;
        else if (ax[i] > hi) v = ax[i] - hi;
        psq += v * v;
        pmax = std::max(pmax, std::abs(v));
        // Normalised dual: ybar = sgn * y must be <= 0 on rows without an
        // upper bound and >= 0 on rows without a lower bound (the scaled
        // projections of pdlp.cpp:1185-1194 in original units).
        LD yb = sgn * static_cast<LD>(y[i]);
        const bool hasLo = std::isfinite(lp->row_lower[i]), hasHi = std::isfinite(lp->row_upper[i]);
        if (!hasHi && yb > 0) { ysignmax = std::max(ysignmax, yb); yb = 0; }
        if (!hasLo && yb < 0) { ysignmax = std::max(ysignmax, -yb); yb = 0; }
        if (yb > 0 && hasHi) drow -= yb * hi;
        else if (yb < 0 && hasLo) drow -= yb * lo;
    }
    LD dsq = 0, dcol = 0, bviol = 0, pobj = 0;
    for (Index j = 0; j < n; ++j) {
        const LD lb = lp->col_lower[j], ub = lp->col_upper[j];
        const bool hasL = std::isfinite(lp->col_lower[j]), hasU = std::isfinite(lp->col_upper[j]);
        if (hasL && x[j] < lb) bviol = std::max(bviol, lb - x[j]);
        if (hasU && x[j] > ub) bviol = std::max(bviol, static_cast<LD>(x[j]) - ub);
        pobj += static_cast<LD>(lp->obj[j]) * x[j];
        const LD g = sgn * (static_cast<LD>(lp->obj[j]) + aty[j]);
        LD gf = g;
        if (!hasL) gf = std::min(gf, 0.0L);
        if (!hasU) gf = std::max(gf, 0.0L);
        const LD dres = g - gf;
        dsq += dres * dres;
        if (gf > 0 && hasL) dcol += gf * lb;
        else if (gf < 0 && hasU) dcol += gf * ub;
    }
    const LD off = lp->obj_offset;
    const LD pob = off + pobj;
    const LD dob = off + sgn * (dcol + drow);
    out[0] = static_cast<double>(std::sqrt(psq) / (1 + bnorm));
    out[1] = static_cast<double>(std::sqrt(dsq) / (1 + cnorm));
    out[2] = static_cast<double>(std::abs(pob - dob) / (1 + std::abs(pob) + std::abs(dob)));
    out[3] = static_cast<double>(pob);
    out[4] = static_cast<double>(dob);
    out[5] = static_cast<double>(bviol);
    out[6] = static_cast<double>(pmax);
    out[7] = static_cast<double>(ysignmax);
}

}  // extern "C"
