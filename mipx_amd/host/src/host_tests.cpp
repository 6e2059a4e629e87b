// host_tests.cpp -- C++ tests of mipx::PdlpSolver on the device, after the
// reference's tests/test_pdlp.cpp closed-form cases (line numbers cited per
// case).  A tiny runner (no test framework): prints one line per check and
// exits nonzero on any failure.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "mipx/pdlp.h"

using namespace mipx;

static int g_fail = 0, g_pass = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        if (cond) {                                                             \
            ++g_pass;                                                           \
        } else {                                                                \
            ++g_fail;                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
        }                                                                       \
    } while (0)

static LpProblem lp2(std::vector<Real> obj, std::vector<Real> lo, std::vector<Real> hi,
                     std::vector<Real> rl, std::vector<Real> ru, std::vector<Triplet> t) {
    LpProblem p;
    p.num_cols = static_cast<Index>(obj.size());
    p.num_rows = static_cast<Index>(rl.size());
    p.obj = std::move(obj);
    p.col_lower = std::move(lo);
    p.col_upper = std::move(hi);
    p.row_lower = std::move(rl);
    p.row_upper = std::move(ru);
    p.matrix = SparseMatrix(p.num_rows, p.num_cols, std::move(t));
    return p;
}

// The closed-form LPs are far below the reference's GPU gate (512 rows,
// 10,000 nonzeros), so the device tests open it.
static PdlpOptions quiet(Int max_iter = 50000) {
    PdlpOptions o;
    o.verbose = false;
    o.max_iter = max_iter;
    o.gpu_min_rows = 0;
    o.gpu_min_nnz = 0;
    return o;
}

// Reduced costs in original space, c + A'y, by a dense loop over the
// triplets (independent of SparseMatrix::multiplyTranspose).
static std::vector<Real> denseReducedCosts(const std::vector<Real>& c,
                                           const std::vector<Triplet>& t,
                                           const std::vector<Real>& y) {
    std::vector<Real> rc = c;
    for (const auto& e : t) rc[static_cast<size_t>(e.col)] += e.val * y[static_cast<size_t>(e.row)];
    return rc;
}

// The reference's GPU gate (pdlp.cpp:862-869) in the mirror.  No device call:
// every LP here is either rejected by the gate before the library is used,
// or only asked whether it would pass.
static void gateTests(const LpProblem& simple) {
    {  // default options: 1 row < 512, so the reference would run its CPU loop
        PdlpSolver s;
        PdlpOptions o;
        o.verbose = false;
        s.setOptions(o);
        s.load(simple);
        CHECK(!s.passesGpuGate());
        const LpResult r = s.solve();
        CHECK(r.status == Status::Error && r.objective == 0.0 && r.iterations == 0);
        CHECK(!s.usedGpu());
        CHECK(s.lastError().find("GPU gate") != std::string::npos);
        CHECK(s.getPrimalValues().size() == 2 && s.getPrimalValues()[0] == 0.0);
    }
    {  // thresholds open, use_gpu off
        PdlpSolver s;
        PdlpOptions o = quiet();
        o.use_gpu = false;
        s.setOptions(o);
        s.load(simple);
        CHECK(!s.passesGpuGate());
        CHECK(s.solve().status == Status::Error && !s.usedGpu());
        o.use_gpu = true;
        s.setOptions(o);
        CHECK(s.passesGpuGate());
    }
    {  // rows and nonzeros at the thresholds pass; one below fails
        PdlpSolver s;
        PdlpOptions o = quiet();
        o.gpu_min_rows = 1;
        o.gpu_min_nnz = 2;
        s.setOptions(o);
        s.load(simple);
        CHECK(s.passesGpuGate());
        o.gpu_min_nnz = 3;
        s.setOptions(o);
        CHECK(!s.passesGpuGate());
        o.gpu_min_nnz = 2;
        o.gpu_min_rows = 2;
        s.setOptions(o);
        CHECK(!s.passesGpuGate());
    }
    {  // the gate counts entries that survive matrix_zero_tol (pdlp.cpp:499-518)
        PdlpSolver s;
        PdlpOptions o = quiet();
        o.gpu_min_nnz = 2;
        o.matrix_zero_tol = 1e-3;
        s.setOptions(o);
        s.load(lp2({-1, -1}, {0, 0}, {kInf, kInf}, {-kInf}, {4.0}, {{0, 0, 1.0}, {0, 1, 1e-4}}));
        CHECK(!s.passesGpuGate());
        o.matrix_zero_tol = 0.0;
        s.setOptions(o);
        CHECK(s.passesGpuGate());
    }
    {  // trivial LPs (no rows or no columns) are closed forms before the gate
        PdlpSolver s;
        PdlpOptions o;
        o.use_gpu = false;
        s.setOptions(o);
        s.load(lp2({1.0}, {1.0}, {3.0}, {}, {}, {}));
        CHECK(s.passesGpuGate());
        LpProblem empty;
        s.load(empty);
        CHECK(s.passesGpuGate());
    }
}

int main(int argc, char** argv) {
    const LpProblem simple = lp2({-1, -1}, {0, 0}, {kInf, kInf}, {-kInf}, {4.0},
                                 {{0, 0, 1.0}, {0, 1, 1.0}});
    gateTests(simple);
    if (argc > 1 && std::string(argv[1]) == "--gate-only") {
        std::printf("host_tests: %d passed, %d failed\n", g_pass, g_fail);
        return g_fail == 0 ? 0 : 1;
    }
    {  // test_pdlp.cpp:88-107
        PdlpSolver s;
        s.setOptions(quiet(10000));
        s.load(simple);
        const LpResult r = s.solve();
        CHECK(r.status == Status::Optimal);
        CHECK(std::abs(r.objective + 4.0) <= 1e-4);
        CHECK(r.work_units > 0.0);
        const auto x = s.getPrimalValues();
        CHECK(x.size() == 2 && x[0] >= -1e-6 && x[1] >= -1e-6 && x[0] + x[1] <= 4.0 + 1e-3);
        CHECK(s.usedGpu());
        // getReducedCosts = c + A'y in original space (pdlp.cpp:2808-2819)
        const auto rc = s.getReducedCosts();
        const auto want = denseReducedCosts({-1, -1}, {{0, 0, 1.0}, {0, 1, 1.0}}, s.getDualValues());
        CHECK(rc.size() == 2 && rc == want);
        // at the optimum of min -x0 - x1 s.t. x0 + x1 <= 4 the row is tight
        // and both reduced costs vanish with y = 1 (mipx's sign convention)
        CHECK(std::abs(rc[0]) <= 1e-4 && std::abs(rc[1]) <= 1e-4);
        CHECK(std::abs(s.getDualValues()[0] - 1.0) <= 1e-4);
    }
    {  // bounds-only, test_pdlp.cpp:109-124
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(lp2({1.0}, {1.0}, {3.0}, {}, {}, {}));
        const LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective - 1.0) <= 1e-6);
        CHECK(std::abs(s.getPrimalValues()[0] - 1.0) <= 1e-5);
    }
    {  // empty, offset 42, test_pdlp.cpp:303-322
        LpProblem p;
        p.obj_offset = 42.0;
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(p);
        const LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective - 42.0) <= 1e-9 &&
              r.iterations == 0);
    }
    {  // infeasible, test_pdlp.cpp:347-376 + Farkas ray 445-496
        PdlpSolver s;
        s.setOptions(quiet(100000));
        s.load(lp2({1.0}, {1.0}, {kInf}, {-kInf}, {-1.0}, {{0, 0, 1.0}}));
        CHECK(s.solve().status == Status::Infeasible);
        const auto y = s.getDualValues();
        CHECK(y.size() == 1 && std::abs(y[0]) > 1e-10);
        const Real ray = y[0] / std::abs(y[0]);
        CHECK(ray >= -1e-6 && ray * (-1.0) - ray * 1.0 < -1e-6);
    }
    {  // unbounded constrained, test_pdlp.cpp:414-443
        PdlpSolver s;
        s.setOptions(quiet(100000));
        s.load(lp2({-1.0, 0.0}, {0, 0}, {kInf, kInf}, {-kInf}, {0.0}, {{0, 0, 1.0}, {0, 1, -1.0}}));
        CHECK(s.solve().status == Status::Unbounded);
    }
    {  // binding column bounds, -13 (test_pdlp.cpp:639-663); then setColBounds
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(lp2({-1, -2}, {0, 0}, {3, 5}, {-kInf}, {100.0}, {{0, 0, 1.0}, {0, 1, 1.0}}));
        LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective + 13.0) <= 1e-3);
        // both columns sit at their upper bounds with a slack row: y = 0 and
        // the reduced costs are the objective (-1, -2)
        const auto rc = s.getReducedCosts();
        CHECK(rc == denseReducedCosts({-1, -2}, {{0, 0, 1.0}, {0, 1, 1.0}}, s.getDualValues()));
        CHECK(std::abs(rc[0] + 1.0) <= 1e-4 && std::abs(rc[1] + 2.0) <= 1e-4);
        s.setColBounds(1, 0.0, 4.0);  // optimum moves to -11
        r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective + 11.0) <= 1e-3);
        s.setColBounds(7, 0.0, 1.0);  // invalid column: ignored
        CHECK(s.solve().status == Status::Optimal);
    }
    {  // equality, 10 (test_pdlp.cpp:692-716); addRows / removeRows round trip
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(lp2({1, 2}, {0, 0}, {kInf, kInf}, {10.0}, {10.0}, {{0, 0, 1.0}, {0, 1, 1.0}}));
        CHECK(std::abs(s.solve().objective - 10.0) <= 0.1);
        const std::vector<Index> st = {0, 1}, idx = {0};
        const std::vector<Real> val = {1.0}, lo = {-kInf}, hi = {4.0};
        s.addRows(st, idx, val, lo, hi);  // x0 <= 4 forces x1 = 6: objective 16
        LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective - 16.0) <= 0.1);
        const std::vector<Index> rm = {1};
        s.removeRows(rm);
        r = s.solve();
        CHECK(std::abs(r.objective - 10.0) <= 0.1);
    }
    {  // time limit 0 (test_pdlp.cpp:141-154)
        PdlpSolver s;
        PdlpOptions o = quiet();
        o.max_solve_seconds = 0.0;
        s.setOptions(o);
        s.load(simple);
        CHECK(s.solve().status == Status::TimeLimit);
    }
    {  // stop flag: cancellation reports IterLimit (pdlp.cpp:972-975)
        std::atomic<bool> stop{true};
        PdlpSolver s;
        PdlpOptions o = quiet();
        o.stop_flag = &stop;
        s.setOptions(o);
        s.load(simple);
        CHECK(s.solve().status == Status::IterLimit);
    }
    {  // solve before load is an error (pdlp.cpp:802-805)
        PdlpSolver s;
        CHECK(s.solve().status == Status::Error);
    }
    {  // warm start with wrong sizes is ignored (test_pdlp.cpp:1182-1201)
        PdlpSolver s;
        s.setOptions(quiet(10000));
        s.load(simple);
        const std::vector<Real> wx = {1.0, 2.0, 3.0}, wy = {0.5};
        s.setWarmStart(wx, wy);
        const LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective + 4.0) <= 1e-4);
    }
    {  // warm start after a bound change (test_pdlp.cpp:1116-1153): the row
       // bound 4 -> 3 moves the optimum to -3; cold and warm solves agree,
       // and the warm one starts from the old solution
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(simple);
        const LpResult r1 = s.solve();
        CHECK(r1.status == Status::Optimal);
        const std::vector<Real> px = s.getPrimalValues(), py = s.getDualValues();
        LpProblem moved = simple;
        moved.row_upper[0] = 3.0;
        PdlpSolver cold, warm;
        cold.setOptions(quiet());
        cold.load(moved);
        const LpResult rc = cold.solve();
        warm.setOptions(quiet());
        warm.load(moved);
        warm.setWarmStart(px, py);
        const LpResult rw = warm.solve();
        CHECK(rc.status == Status::Optimal && std::abs(rc.objective + 3.0) <= 1e-3);
        CHECK(rw.status == Status::Optimal && std::abs(rw.objective + 3.0) <= 1e-3);
    }
    {  // clearWarmStart disables the warm start (test_pdlp.cpp:1155-1180):
       // the re-solve is the cold solve, bit for bit
        PdlpSolver s;
        s.setOptions(quiet());
        s.load(simple);
        const LpResult r1 = s.solve();
        CHECK(r1.status == Status::Optimal);
        const std::vector<Real> px = s.getPrimalValues(), py = s.getDualValues();
        s.setWarmStart(px, py);
        s.clearWarmStart();
        s.load(simple);
        const LpResult r2 = s.solve();
        CHECK(r2.status == Status::Optimal && std::abs(r2.objective + 4.0) <= 1e-4);
        CHECK(r2.iterations == r1.iterations && r2.objective == r1.objective);
        CHECK(s.getPrimalValues() == px && s.getDualValues() == py);
    }
    {  // mixed precision on the simple LP
        PdlpSolver s;
        PdlpOptions o = quiet(10000);
        o.precision = PdlpPrecision::Mixed;
        s.setOptions(o);
        s.load(simple);
        const LpResult r = s.solve();
        CHECK(r.status == Status::Optimal && std::abs(r.objective + 4.0) <= 1e-4);
    }
    std::printf("host_tests: %d passed, %d failed\n", g_pass, g_fail);
    return g_fail == 0 ? 0 : 1;
}
