// io_tests.cpp -- checks of the MPS reader / writer and the .solu reader
// (no GPU).  Usage: io_tests GOLDEN_DIR SCRATCH_DIR
//
// tiny.mps and tiny.solu are the reference's test data files
// (tests/data/tiny.mps, tests/data/tiny.solu); the expected fields are those
// of its MPS reader test (tests/test_lp_problem.cpp:50-108).
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "mipx/io.h"

using namespace mipx;

static int g_pass = 0, g_fail = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        if (cond) ++g_pass;                                                    \
        else {                                                                 \
            ++g_fail;                                                          \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
        }                                                                      \
    } while (0)

static Real coeff(const LpProblem& p, Index i, Index j) {
    const auto rs = p.matrix.csr_row_starts();
    const auto ci = p.matrix.csr_col_indices();
    const auto v = p.matrix.csr_values();
    for (Index k = rs[i]; k < rs[i + 1]; ++k)
        if (ci[k] == j) return v[k];
    return 0.0;
}

static bool sameLp(const LpProblem& a, const LpProblem& b) {
    if (a.num_rows != b.num_rows || a.num_cols != b.num_cols) return false;
    if (a.obj != b.obj || a.col_lower != b.col_lower || a.col_upper != b.col_upper) return false;
    if (a.row_lower != b.row_lower || a.row_upper != b.row_upper) return false;
    if (a.obj_offset != b.obj_offset) return false;
    const auto ra = a.matrix.csr_row_starts(), rb = b.matrix.csr_row_starts();
    const auto ca = a.matrix.csr_col_indices(), cb = b.matrix.csr_col_indices();
    const auto va = a.matrix.csr_values(), vb = b.matrix.csr_values();
    if (ra.size() != rb.size() || ca.size() != cb.size()) return false;
    for (size_t k = 0; k < ra.size(); ++k)
        if (ra[k] != rb[k]) return false;
    for (size_t k = 0; k < ca.size(); ++k)
        if (ca[k] != cb[k] || va[k] != vb[k]) return false;
    return true;
}

static void checkTiny(const LpProblem& p) {
    CHECK(p.name == "tiny");
    CHECK(p.num_cols == 3 && p.num_rows == 3);
    CHECK(p.col_names.size() == 3 && p.col_names[0] == "x1" && p.col_names[1] == "x2" &&
          p.col_names[2] == "x3");
    CHECK(p.obj[0] == 1.0 && p.obj[1] == 3.0 && p.obj[2] == 2.0);
    CHECK(p.col_lower[0] == 0.0 && p.col_upper[0] == 8.0);
    CHECK(p.col_lower[1] == 0.0 && p.col_upper[1] == 1.0);
    CHECK(p.col_lower[2] == 1.0 && p.col_upper[2] == 5.0);
    CHECK(p.col_type[0] == VarType::Continuous);
    CHECK(p.col_type[1] == VarType::Binary);  // BV after INTORG
    CHECK(p.col_type[2] == VarType::Continuous);
    CHECK(p.row_names.size() == 3 && p.row_names[0] == "c1" && p.row_names[1] == "c2" &&
          p.row_names[2] == "c3");
    CHECK(p.row_lower[0] == -kInf && p.row_upper[0] == 10.0);  // L
    CHECK(p.row_lower[1] == 5.0 && p.row_upper[1] == kInf);    // G
    CHECK(p.row_lower[2] == 7.0 && p.row_upper[2] == 7.0);     // E
    CHECK(coeff(p, 0, 0) == 2.0 && coeff(p, 0, 1) == 1.0 && coeff(p, 0, 2) == 0.0);
    CHECK(coeff(p, 1, 0) == 1.0 && coeff(p, 1, 2) == 1.0);
    CHECK(coeff(p, 2, 1) == 1.0 && coeff(p, 2, 2) == 1.0);
    CHECK(p.matrix.numNonzeros() == 6);
}

static void writeText(const std::string& path, const std::string& text) {
    std::ofstream out(path, std::ios::binary);
    out << text;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: io_tests GOLDEN_DIR SCRATCH_DIR\n");
        return 2;
    }
    const std::string gold = argv[1], tmp = argv[2];
    const LpProblem tiny = readMps(gold + "/tiny.mps");
    checkTiny(tiny);

    // Round trip through the writer, and through a gzip-compressed copy.
    writeMps(tmp + "/tiny_rt.mps", tiny);
    const LpProblem back = readMps(tmp + "/tiny_rt.mps");
    CHECK(sameLp(tiny, back));
    CHECK(back.col_type[1] == VarType::Integer || back.col_type[1] == VarType::Binary);
    {
        std::ifstream in(gold + "/tiny.mps", std::ios::binary);
        std::ostringstream ss;
        ss << in.rdbuf();
        const std::string text = ss.str();
        gzFile z = gzopen((tmp + "/tiny.mps.gz").c_str(), "wb");
        CHECK(z != nullptr);
        if (z) {
            gzwrite(z, text.data(), static_cast<unsigned>(text.size()));
            gzclose(z);
            const LpProblem gz = readMps(tmp + "/tiny.mps.gz");
            CHECK(sameLp(tiny, gz));
        }
    }

    // RANGES on L, G and E rows, objective offset, free and negative bounds,
    // duplicate entries summed, an end-of-line comment.
    writeText(tmp + "/ranges.mps",
              "NAME rng\n"
              "* comment line\n"
              "ROWS\n N cost\n L a\n G b\n E c\n E d\n"
              "COLUMNS\n"
              "    x cost 1.5 a 1.0\n"
              "    x b 2.0 c 1.0 $ trailing comment\n"
              "    x a 0.5\n"
              "    y cost -1.0 d 3.0\n"
              "RHS\n    rhs cost 4.0 a 10.0\n    rhs b 2.0 c 3.0\n    rhs d 6.0\n"
              "RANGES\n    r a 4.0 b -3.0\n    r c 2.0 d -1.5\n"
              "BOUNDS\n FR bnd x\n MI bnd y\n UP bnd y -2.0\n"
              "ENDATA\n");
    const LpProblem r = readMps(tmp + "/ranges.mps");
    CHECK(r.num_rows == 4 && r.num_cols == 2);
    CHECK(r.obj[0] == 1.5 && r.obj[1] == -1.0 && r.obj_offset == -4.0);
    CHECK(r.row_lower[0] == 6.0 && r.row_upper[0] == 10.0);  // L: [rhs - |R|, rhs]
    CHECK(r.row_lower[1] == 2.0 && r.row_upper[1] == 5.0);   // G: [rhs, rhs + |R|]
    CHECK(r.row_lower[2] == 3.0 && r.row_upper[2] == 5.0);   // E, R > 0
    CHECK(r.row_lower[3] == 4.5 && r.row_upper[3] == 6.0);   // E, R < 0
    CHECK(coeff(r, 0, 0) == 1.5);                            // 1.0 + 0.5
    CHECK(r.col_lower[0] == -kInf && r.col_upper[0] == kInf);
    CHECK(r.col_lower[1] == -kInf && r.col_upper[1] == -2.0);

    // An entry of an undeclared row is an error.
    writeText(tmp + "/bad.mps", "NAME bad\nROWS\n N obj\n L r1\nCOLUMNS\n    x r2 1.0\nENDATA\n");
    bool threw = false;
    try {
        (void)readMps(tmp + "/bad.mps");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);

    const auto solu = readSolu(gold + "/tiny.solu");
    CHECK(solu.size() == 3);
    CHECK(solu.size() == 3 && solu[0].name == "tiny" && solu[0].value == 42.0 &&
          !solu[0].is_infeasible);
    CHECK(solu.size() == 3 && solu[1].is_infeasible && solu[2].value == 123.456);

    std::printf("io_tests: %d passed, %d failed\n", g_pass, g_fail);
    return g_fail == 0 ? 0 : 1;
}
