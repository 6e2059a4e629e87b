// mipx_solve.cpp -- command-line PDLP solve on the MI355X.
//
// Mirrors the LP branch of the reference CLI for `--pdlp` (src/cli/main.cpp:
// 491-558, 572-609): builds a PdlpSolver with verbose / time limit /
// optimality norm, prints "PDLP backend: GPU" and the stable summary lines
// (Status / Objective / Iterations / Work units / Time) that the reference's
// tests/perf scripts parse.  Input is an MPS file (the reference CLI's
// positional argument; .mps or .mps.gz, read as its LP relaxation) or a
// seeded synthetic LP from the library's generator; --write-mps stores the
// problem that was solved.
//
//   mipx_solve [FILE.mps[.gz] | --gen c0|c1|c2|c3 [--seed S] [--rows M] [--cols N]]
//              [--pdlp] [--time-limit S] [--pdlp-opt-norm l2|linf]
//              [--precision fp64|mixed] [--tol T] [--quiet] [--print-backend]
//              [--write-mps OUT.mps] [--max-iter N]
//              [--gpu|--no-gpu] [--gpu-min-rows N] [--gpu-min-nnz N]
//              [--no-presolve]
//
// The gate flags keep the reference CLI's defaults (512 rows, 10,000
// nonzeros); an LP below them reports "PDLP backend: CPU" and status Error,
// since the reference's CPU loop is not part of this library.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mipx/io.h"
#include "mipx/pdlp.h"
#include "mipx/presolve.h"

extern "C" {
struct mpxgen_params {
    int32_t m, n, colnnz_kind, colnnz_lo, colnnz_hi;
    double colnnz_mean;
    int32_t value_kind;
    double vlo, vhi, xhi, slack, eq_fraction;
    uint64_t seed;
};
void* mpxgen_create(const mpxgen_params* p);
void mpxgen_dims(void* h, int32_t* m, int32_t* n, int64_t* nnz);
void mpxgen_get(void* h, int32_t* rs, int32_t* ci, double* val, double* obj, double* cl,
                double* cu, double* rl, double* ru, double* x0);
void mpxgen_destroy(void* h);
}

static mipx::LpProblem generate(const std::string& kind, uint64_t seed, int m, int n) {
    mpxgen_params p{};
    p.xhi = 10.0;
    p.slack = 1.0;
    p.eq_fraction = 0.5;
    p.seed = seed;
    if (kind == "c1") {
        p.m = m ? m : 1000000;
        p.n = n ? n : 2 * p.m;
        p.colnnz_kind = 2;
        p.colnnz_lo = 1;
        p.colnnz_hi = 200;
        p.colnnz_mean = 6.0;
        p.value_kind = 1;
        p.vlo = -2.0;
        p.vhi = 2.0;
    } else if (kind == "c2") {
        p.m = m ? m : 3000;
        p.n = n ? n : 8000;
        p.colnnz_lo = p.colnnz_hi = 6;
        p.value_kind = 2;
        p.vhi = 5.0;
    } else if (kind == "c3") {
        p.m = m ? m : 2000;
        p.n = n ? n : (5 * p.m) / 2;
        p.colnnz_kind = 1;
        p.colnnz_lo = 4;
        p.colnnz_hi = 8;
        p.value_kind = 1;
        p.vlo = -3.0;
        p.vhi = 3.0;
    } else {
        p.m = m ? m : 2000;
        p.n = n ? n : 5000;
        p.colnnz_lo = p.colnnz_hi = 5;
        p.value_kind = 0;
        p.vlo = 0.1;
        p.vhi = 1.0;
    }
    void* h = mpxgen_create(&p);
    if (!h) {
        std::fprintf(stderr, "Error: bad generator parameters\n");
        std::exit(1);
    }
    int32_t mm, nn;
    int64_t nnz;
    mpxgen_dims(h, &mm, &nn, &nnz);
    std::vector<int32_t> rs(mm + 1), ci(nnz);
    std::vector<double> val(nnz), obj(nn), cl(nn), cu(nn), rl(mm), ru(mm);
    mpxgen_get(h, rs.data(), ci.data(), val.data(), obj.data(), cl.data(), cu.data(), rl.data(),
               ru.data(), nullptr);
    mpxgen_destroy(h);
    mipx::LpProblem lp;
    lp.name = kind;
    lp.num_rows = mm;
    lp.num_cols = nn;
    lp.obj = std::move(obj);
    lp.col_lower = std::move(cl);
    lp.col_upper = std::move(cu);
    lp.row_lower = std::move(rl);
    lp.row_upper = std::move(ru);
    lp.matrix = mipx::SparseMatrix(mm, nn, std::move(val), std::vector<int>(ci.begin(), ci.end()),
                                   std::vector<int>(rs.begin(), rs.end()));
    return lp;
}

int main(int argc, char** argv) {
    std::string kind = "c0", precision = "fp64", mps, write_mps;
    uint64_t seed = 0xC0;
    int m = 0, n = 0;
    bool verbose = true, print_backend = false;
    double time_limit = -1.0, tol = 1e-6;
    int max_iter = 0;
    // The reference CLI's GPU gate flags and defaults (main.cpp:75-80, 284-293).
    bool use_gpu = true;
    bool presolve = true;  // the reference CLI presolves unless --no-presolve
    int gpu_min_rows = 512, gpu_min_nnz = 10000;
    mipx::PdlpOptimalityNorm norm = mipx::PdlpOptimalityNorm::L2;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "Error: %s needs a value\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--gen") kind = next();
        else if (a == "--seed") seed = std::strtoull(next().c_str(), nullptr, 0);
        else if (a == "--rows") m = std::atoi(next().c_str());
        else if (a == "--cols") n = std::atoi(next().c_str());
        else if (a == "--pdlp") {
        } else if (a == "--time-limit") time_limit = std::atof(next().c_str());
        else if (a == "--tol") tol = std::atof(next().c_str());
        else if (a == "--precision") precision = next();
        else if (a == "--pdlp-opt-norm") {
            const std::string v = next();
            if (v == "linf") norm = mipx::PdlpOptimalityNorm::LInf;
            else if (v != "l2") {
                std::fprintf(stderr, "Error: --pdlp-opt-norm must be l2 or linf\n");
                return 1;
            }
        } else if (a == "--quiet") verbose = false;
        else if (a == "--print-backend") print_backend = true;
        else if (a == "--write-mps") write_mps = next();
        else if (a == "--max-iter") max_iter = std::atoi(next().c_str());
        else if (a == "--no-presolve") presolve = false;
        else if (a == "--presolve") presolve = true;
        else if (a == "--gpu") use_gpu = true;
        else if (a == "--no-gpu") use_gpu = false;
        else if (a == "--gpu-min-rows") gpu_min_rows = std::max(0, std::atoi(next().c_str()));
        else if (a == "--gpu-min-nnz") gpu_min_nnz = std::max(0, std::atoi(next().c_str()));
        else if (a.size() > 4 && a[0] != '-') mps = a;
        else {
            std::fprintf(stderr, "Error: unknown option %s\n", a.c_str());
            return 1;
        }
    }
    mipx::LpProblem lp;
    try {
        lp = mps.empty() ? generate(kind, seed, m, n) : mipx::readMps(mps);
        if (!write_mps.empty()) mipx::writeMps(write_mps, lp);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    if (verbose)
        std::printf("LP %s: %d rows, %d cols, %d nonzeros\n", lp.name.c_str(), lp.num_rows,
                    lp.num_cols, lp.matrix.numNonzeros());
    const auto t0 = std::chrono::steady_clock::now();
    // Presolve, as the reference CLI does before an LP solve
    // (src/cli/main.cpp:455-467); the reduced LP is used only when something
    // was removed.
    mipx::Presolver pre;
    mipx::LpProblem working;
    bool did_presolve = false;
    if (presolve) {
        working = pre.presolve(lp);
        if (pre.isInfeasible()) {
            std::printf("Presolve: infeasible\n");
            std::printf("PDLP backend: CPU\n");
            std::printf("Status: Infeasible\nObjective: %.10e\nIterations: 0\nWork units: 0.00\n",
                        0.0);
            std::printf("Time: %.2fs\n", std::chrono::duration<double>(
                                             std::chrono::steady_clock::now() - t0).count());
            return 0;
        }
        did_presolve = pre.stats().vars_removed > 0 || pre.stats().rows_removed > 0;
        if (verbose || print_backend)
            std::printf("Presolve: removed %d rows, %d cols\n", pre.stats().rows_removed,
                        pre.stats().vars_removed);
    }
    mipx::PdlpSolver solver;
    mipx::PdlpOptions o;
    o.verbose = verbose;
    o.max_solve_seconds = time_limit;
    o.optimality_norm = norm;
    o.use_gpu = use_gpu;
    o.gpu_min_rows = gpu_min_rows;
    o.gpu_min_nnz = gpu_min_nnz;
    o.primal_tol = o.dual_tol = o.optimality_tol = tol;
    if (max_iter > 0) o.max_iter = max_iter;
    o.precision = precision == "mixed" ? mipx::PdlpPrecision::Mixed : mipx::PdlpPrecision::FP64;
    solver.setOptions(o);
    solver.load(did_presolve ? working : lp);
    mipx::LpResult r = solver.solve();
    // Postsolve and the objective on the original columns (main.cpp:560-568).
    if (did_presolve && r.status == mipx::Status::Optimal) {
        const std::vector<double> x = pre.postsolve(solver.getPrimalValues());
        double obj = lp.obj_offset;
        for (int j = 0; j < lp.num_cols; ++j) obj += lp.obj[j] * x[j];
        r.objective = obj;
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fflush(stdout);
    if (verbose || print_backend) std::printf("PDLP backend: %s\n", solver.usedGpu() ? "GPU" : "CPU");
    const char* text = "Error";
    switch (r.status) {
        case mipx::Status::Optimal: text = "Optimal"; break;
        case mipx::Status::Infeasible: text = "Infeasible"; break;
        case mipx::Status::Unbounded: text = "Unbounded"; break;
        case mipx::Status::IterLimit: text = "Iteration limit"; break;
        case mipx::Status::TimeLimit: text = "Time limit"; break;
        default: text = "Error"; break;
    }
    if (r.status == mipx::Status::Error && !solver.lastError().empty())
        std::printf("Error detail: %s\n", solver.lastError().c_str());
    std::printf("Status: %s\n", text);
    std::printf("Objective: %.10e\n", r.objective);
    std::printf("Iterations: %d\n", r.iterations);
    std::printf("Work units: %.2f\n", r.work_units);
    std::printf("Time: %.2fs\n", secs);
    return 0;
}
