// lpgen.cpp -- seeded synthetic LP generator for the PDLP workloads.
//
// Portable and deterministic: SplitMix64 seeds a xoshiro256** stream,
// uniforms take the top 53 bits, normals use Box-Muller.  No
// std::*_distribution is used (their output is implementation-defined).
//
// Construction (feasible primal and dual by design):
//   * per column j: k_j distinct rows drawn uniformly; k_j fixed, uniform on
//     [lo, hi], or a truncated power law P(k) ~ k^-alpha on [lo, hi] with
//     alpha chosen by bisection so that E[k] = mean;
//   * values: sign * U[vlo, vhi], sign * 10^U[vlo, vhi], or an integer
//     uniform on [-K, K] \ {0};
//   * 0 <= x <= xhi; the first round(eq_fraction * m) rows are equalities,
//     the rest are <= rows;
//   * x0 ~ U(0, xhi), b = A x0; <= rows get ru = b + U(0, slack), rl = -inf;
//     equality rows rl = ru = b;
//   * y0 ~ N(0, 1) with y0 <= 0 on <= rows, s0 = |N(0, 1)|, c = A' y0 + s0.
//   * optional column kinds (fractions; 0 keeps every column boxed and the
//     LP bit-identical to a generator without them): free columns
//     (-inf < x < inf, c = A' y0), lower-only (0 <= x, c = A' y0 + s0) and
//     upper-only (x <= xhi, c = A' y0 - s0), drawn from a second stream, so
//     that x0 stays feasible and y0 dual feasible.
// CSR output has sorted column indices in every row.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& z) {
        uint64_t r = (z += 0x9E3779B97F4A7C15ull);
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        return r ^ (r >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) w = splitmix(z);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint32_t below(uint32_t n) { return static_cast<uint32_t>((next() >> 32) * n >> 32); }
    double normal() {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" {

typedef struct mpxgen_params {
    int32_t m, n;
    int32_t colnnz_kind;  // 0 fixed (lo), 1 uniform on [lo, hi], 2 power law on [lo, hi]
    int32_t colnnz_lo, colnnz_hi;
    double colnnz_mean;   // power law target mean
    int32_t value_kind;   // 0 sign*U[vlo,vhi], 1 sign*10^U[vlo,vhi], 2 integer [-K,K]\{0}, K=vhi
    double vlo, vhi;
    double xhi;           // column upper bound (lower is 0)
    double slack;         // <= row slack ~ U(0, slack)
    double eq_fraction;   // leading fraction of equality rows
    uint64_t seed;
    double free_fraction;        // columns with no bounds
    double lower_only_fraction;  // columns 0 <= x
    double upper_only_fraction;  // columns x <= xhi
} mpxgen_params;

struct MpxGen {
    int32_t m = 0, n = 0;
    std::vector<int32_t> rs, ci;
    std::vector<double> val, obj, cl, cu, rl, ru, x0;
};

static double powerLawMean(double alpha, int lo, int hi) {
    double num = 0.0, den = 0.0;
    for (int k = lo; k <= hi; ++k) {
        const double w = std::pow(static_cast<double>(k), -alpha);
        num += k * w;
        den += w;
    }
    return num / den;
}

void* mpxgen_create(const mpxgen_params* p) {
    if (!p || p->m <= 0 || p->n <= 0) return nullptr;
    auto* g = new MpxGen();
    const int32_t m = p->m, n = p->n;
    g->m = m;
    g->n = n;
    Rng rng(p->seed);

    // Column counts.
    std::vector<int32_t> kcol(static_cast<size_t>(n));
    std::vector<double> cdf;
    if (p->colnnz_kind == 2) {
        double a = 0.0, b = 6.0;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (a + b);
            if (powerLawMean(mid, p->colnnz_lo, p->colnnz_hi) > p->colnnz_mean) a = mid;
            else b = mid;
        }
        const double alpha = 0.5 * (a + b);
        double acc = 0.0;
        for (int k = p->colnnz_lo; k <= p->colnnz_hi; ++k) {
            acc += std::pow(static_cast<double>(k), -alpha);
            cdf.push_back(acc);
        }
        for (double& v : cdf) v /= acc;
    }
    for (int32_t j = 0; j < n; ++j) {
        int32_t k = p->colnnz_lo;
        if (p->colnnz_kind == 1) {
            k = p->colnnz_lo + static_cast<int32_t>(rng.below(
                                   static_cast<uint32_t>(p->colnnz_hi - p->colnnz_lo + 1)));
        } else if (p->colnnz_kind == 2) {
            const double u = rng.uniform();
            const auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
            k = p->colnnz_lo + static_cast<int32_t>(std::min<size_t>(
                                   static_cast<size_t>(it - cdf.begin()), cdf.size() - 1));
        }
        kcol[j] = std::min(k, m);
    }

    // Entries, column by column (rows distinct within a column).
    std::vector<int32_t> crow;
    std::vector<double> cval;
    std::vector<int64_t> cstart(static_cast<size_t>(n) + 1, 0);
    for (int32_t j = 0; j < n; ++j) cstart[j + 1] = cstart[j] + kcol[j];
    crow.resize(static_cast<size_t>(cstart[n]));
    cval.resize(static_cast<size_t>(cstart[n]));
    std::vector<int32_t> pick;
    for (int32_t j = 0; j < n; ++j) {
        pick.clear();
        while (static_cast<int32_t>(pick.size()) < kcol[j]) {
            const int32_t r = static_cast<int32_t>(rng.below(static_cast<uint32_t>(m)));
            if (std::find(pick.begin(), pick.end(), r) == pick.end()) pick.push_back(r);
        }
        std::sort(pick.begin(), pick.end());
        for (int32_t t = 0; t < kcol[j]; ++t) {
            double v;
            if (p->value_kind == 2) {
                const int32_t kk = static_cast<int32_t>(p->vhi);
                int32_t q = static_cast<int32_t>(rng.below(static_cast<uint32_t>(2 * kk))) - kk;
                if (q >= 0) ++q;  // skip zero
                v = static_cast<double>(q);
            } else {
                const double sign = (rng.next() >> 63) ? -1.0 : 1.0;
                const double u = rng.uniform(p->vlo, p->vhi);
                v = sign * (p->value_kind == 1 ? std::pow(10.0, u) : u);
            }
            crow[cstart[j] + t] = pick[t];
            cval[cstart[j] + t] = v;
        }
    }

    // CSC -> CSR (columns visited in order, so rows come out sorted).
    const int64_t nnz = cstart[n];
    g->rs.assign(static_cast<size_t>(m) + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++g->rs[crow[k] + 1];
    for (int32_t i = 0; i < m; ++i) g->rs[i + 1] += g->rs[i];
    g->ci.resize(static_cast<size_t>(nnz));
    g->val.resize(static_cast<size_t>(nnz));
    std::vector<int32_t> next(g->rs.begin(), g->rs.end() - 1);
    for (int32_t j = 0; j < n; ++j)
        for (int64_t k = cstart[j]; k < cstart[j + 1]; ++k) {
            const int32_t dst = next[crow[k]]++;
            g->ci[dst] = j;
            g->val[dst] = cval[k];
        }

    // Bounds, feasible point, row bounds.
    g->cl.assign(static_cast<size_t>(n), 0.0);
    g->cu.assign(static_cast<size_t>(n), p->xhi);
    g->x0.resize(static_cast<size_t>(n));
    for (auto& v : g->x0) v = rng.uniform(0.0, p->xhi);
    const int32_t meq = static_cast<int32_t>(std::lround(p->eq_fraction * m));
    g->rl.resize(static_cast<size_t>(m));
    g->ru.resize(static_cast<size_t>(m));
    const double inf = std::numeric_limits<double>::infinity();
    for (int32_t i = 0; i < m; ++i) {
        double b = 0.0;
        for (int32_t k = g->rs[i]; k < g->rs[i + 1]; ++k) b += g->val[k] * g->x0[g->ci[k]];
        if (i < meq) {
            g->rl[i] = g->ru[i] = b;
        } else {
            g->rl[i] = -inf;
            g->ru[i] = b + rng.uniform(0.0, p->slack);
        }
    }
    // Dual construction: c = A' y0 + s0.
    std::vector<double> y0(static_cast<size_t>(m));
    for (int32_t i = 0; i < m; ++i) {
        const double z = rng.normal();
        y0[i] = i < meq ? z : -std::abs(z);
    }
    g->obj.assign(static_cast<size_t>(n), 0.0);
    for (int32_t i = 0; i < m; ++i)
        for (int32_t k = g->rs[i]; k < g->rs[i + 1]; ++k) g->obj[g->ci[k]] += g->val[k] * y0[i];
    std::vector<double> s0(static_cast<size_t>(n));
    for (int32_t j = 0; j < n; ++j) s0[j] = std::abs(rng.normal());
    const double f_free = p->free_fraction, f_lo = p->lower_only_fraction,
                 f_up = p->upper_only_fraction;
    if (f_free > 0.0 || f_lo > 0.0 || f_up > 0.0) {
        Rng kinds(p->seed ^ 0x0B0DDEDC01ull);
        for (int32_t j = 0; j < n; ++j) {
            const double u = kinds.uniform();
            if (u < f_free) {
                g->cl[j] = -inf;
                g->cu[j] = inf;
                s0[j] = 0.0;
            } else if (u < f_free + f_lo) {
                g->cu[j] = inf;
            } else if (u < f_free + f_lo + f_up) {
                g->cl[j] = -inf;
                s0[j] = -s0[j];
            }
        }
    }
    for (int32_t j = 0; j < n; ++j) g->obj[j] += s0[j];
    return g;
}

void mpxgen_dims(void* h, int32_t* m, int32_t* n, int64_t* nnz) {
    auto* g = static_cast<MpxGen*>(h);
    *m = g->m;
    *n = g->n;
    *nnz = static_cast<int64_t>(g->val.size());
}

void mpxgen_get(void* h, int32_t* rs, int32_t* ci, double* val, double* obj, double* cl,
                double* cu, double* rl, double* ru, double* x0) {
    auto* g = static_cast<MpxGen*>(h);
    auto put = [](auto* dst, const auto& src) {
        if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(rs, g->rs);
    put(ci, g->ci);
    put(val, g->val);
    put(obj, g->obj);
    put(cl, g->cl);
    put(cu, g->cu);
    put(rl, g->rl);
    put(ru, g->ru);
    put(x0, g->x0);
}

// Branch-and-bound style bound tightening of one instance: the first
// round(int_fraction * n) columns of a seeded permutation are integer; a
// seeded count, uniform on [frac_lo, frac_hi] of them, is branched to
// x_j <= floor(x0_j) or x_j >= ceil(x0_j) (coin flip).  cl_out / cu_out
// start as the base bounds.
void mpxgen_branch_bounds(void* h, uint64_t seed, double int_fraction, double frac_lo,
                          double frac_hi, double* cl_out, double* cu_out) {
    auto* g = static_cast<MpxGen*>(h);
    const int32_t n = g->n;
    std::memcpy(cl_out, g->cl.data(), sizeof(double) * static_cast<size_t>(n));
    std::memcpy(cu_out, g->cu.data(), sizeof(double) * static_cast<size_t>(n));
    // The integer set is a property of the base model: seed it from column
    // count so that every instance sees the same set.
    Rng base(0x1A7E6E5ull ^ static_cast<uint64_t>(n));
    std::vector<int32_t> perm(static_cast<size_t>(n));
    for (int32_t j = 0; j < n; ++j) perm[j] = j;
    for (int32_t j = n - 1; j > 0; --j)
        std::swap(perm[j], perm[base.below(static_cast<uint32_t>(j + 1))]);
    const int32_t nint = static_cast<int32_t>(std::lround(int_fraction * n));
    Rng rng(seed);
    const double frac = rng.uniform(frac_lo, frac_hi);
    const int32_t nb = std::min(nint, static_cast<int32_t>(std::lround(frac * nint)));
    // Branchable integer columns: those whose branch does not fix the
    // variable (floor(x0) > lower and ceil(x0) < upper).  A fixed column
    // (lower == upper) with a negative reduced cost never passes the
    // reference's dual-residual test, which checks the lower bound first
    // (src/lp/pdlp.cpp:1135-1142), so such a node would only exercise the
    // iteration limit.
    std::vector<int32_t> ints;
    for (int32_t t = 0; t < nint; ++t) {
        const int32_t j = perm[t];
        if (std::floor(g->x0[j]) > g->cl[j] && std::ceil(g->x0[j]) < g->cu[j]) ints.push_back(j);
    }
    const int32_t ne = static_cast<int32_t>(ints.size());
    const int32_t take = std::min(nb, ne);
    // Partial Fisher-Yates over the branchable integer columns.
    for (int32_t t = 0; t < take; ++t) {
        const int32_t r = t + static_cast<int32_t>(rng.below(static_cast<uint32_t>(ne - t)));
        std::swap(ints[t], ints[r]);
        const int32_t j = ints[t];
        if (rng.next() >> 63) cu_out[j] = std::floor(g->x0[j]);
        else cl_out[j] = std::ceil(g->x0[j]);
    }
}

void mpxgen_destroy(void* h) { delete static_cast<MpxGen*>(h); }

}  // extern "C"
