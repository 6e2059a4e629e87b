// mps_io.cpp -- MPS reader / writer and .solu reader.
//
// Section and field meaning follow the MPS format as the reference reads it
// (src/io/mps_reader.cpp, src/io/mps_writer.cpp, src/io/solu_reader.cpp):
//   * a line whose first character is not blank starts a section (NAME,
//     ROWS, COLUMNS, RHS, RANGES, BOUNDS, ENDATA; any other header, such as
//     OBJSENSE, opens a section whose lines are ignored); '*' starts a
//     comment line and " $" an end-of-line comment;
//   * every N row is the objective; entries of unknown rows are an error;
//   * columns default to [0, +inf); 'MARKER' 'INTORG' / 'INTEND' mark
//     integer columns; BOUNDS types LO UP FX FR MI PL BV LI UI (SC and SI
//     mark the type and keep [0, upper]);
//   * the RHS of the objective row is the negated objective offset;
//   * RANGES: L rows get [rhs - |R|, rhs], G rows [rhs, rhs + |R|], E rows
//     [rhs, rhs + R] for R > 0 and [rhs + R, rhs] for R < 0;
//   * duplicate matrix entries are summed and zero sums dropped.
#include "mipx/io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string_view>
#include <unordered_map>

namespace mipx {

namespace {

std::string readWhole(const std::string& path) {
    const bool gz = path.size() >= 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    std::string text;
    if (gz) {
        gzFile f = gzopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("MPS: cannot open " + path);
        char buf[1 << 16];
        int got;
        while ((got = gzread(f, buf, sizeof(buf))) > 0) text.append(buf, static_cast<size_t>(got));
        const bool bad = got < 0;
        gzclose(f);
        if (bad) throw std::runtime_error("MPS: cannot inflate " + path);
        return text;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("MPS: cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Blank-separated fields of one line (at most 8 are kept; MPS lines have 6).
struct Fields {
    std::string_view f[8];
    int n = 0;
};

Fields split(std::string_view line) {
    Fields out;
    size_t i = 0;
    while (i < line.size() && out.n < 8) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r')) ++i;
        if (i >= line.size()) break;
        size_t j = i;
        while (j < line.size() && line[j] != ' ' && line[j] != '\t' && line[j] != '\r') ++j;
        out.f[out.n++] = line.substr(i, j - i);
        i = j;
    }
    return out;
}

Real number(std::string_view s) {
    const std::string t(s);
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str()) throw std::runtime_error("MPS: bad number '" + t + "'");
    return v;
}

enum class Part { Other, Rows, Columns, Rhs, Ranges, Bounds };

}  // namespace

LpProblem readMps(const std::string& filename) {
    const std::string text = readWhole(filename);
    LpProblem p;
    std::unordered_map<std::string, Index> rowOf;  // -1: an objective (N) row
    std::unordered_map<std::string, Index> colOf;
    std::vector<char> kind;
    std::vector<Real> rhs, rng;
    std::vector<Triplet> entries;
    Part part = Part::Other;
    bool integerRun = false;

    auto column = [&](std::string_view name) {
        const std::string key(name);
        auto it = colOf.find(key);
        if (it != colOf.end()) return it->second;
        const Index j = p.num_cols++;
        colOf.emplace(key, j);
        p.col_names.push_back(key);
        p.obj.push_back(0.0);
        p.col_lower.push_back(0.0);
        p.col_upper.push_back(kInf);
        p.col_type.push_back(VarType::Continuous);
        return j;
    };
    auto rowValues = [&](const Fields& t, std::vector<Real>& dst, bool objectiveOffset) {
        // [set name] (row value)+ : the set name is present when the count is odd.
        for (int i = (t.n % 2 == 1) ? 1 : 0; i + 1 < t.n; i += 2) {
            auto it = rowOf.find(std::string(t.f[i]));
            if (it == rowOf.end()) continue;
            const Real v = number(t.f[i + 1]);
            if (it->second < 0) {
                if (objectiveOffset) p.obj_offset = -v;
                continue;
            }
            const size_t i2 = static_cast<size_t>(it->second);
            if (i2 >= dst.size()) dst.resize(static_cast<size_t>(p.num_rows), 0.0);
            dst[i2] = v;
        }
    };

    size_t pos = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string_view line(text.data() + pos, eol - pos);
        pos = eol + 1;
        if (line.empty() || line[0] == '*') continue;
        for (size_t k = 1; k < line.size(); ++k)
            if (line[k] == '$' && (line[k - 1] == ' ' || line[k - 1] == '\t')) {
                line = line.substr(0, k);
                break;
            }
        const Fields t = split(line);
        if (t.n == 0) continue;
        if (line[0] != ' ' && line[0] != '\t') {
            const std::string_view h = t.f[0];
            part = Part::Other;
            if (h == "NAME") {
                if (t.n >= 2) p.name = std::string(t.f[1]);
            } else if (h == "ROWS" || h == "LAZYCONS" || h == "USERCUTS") {
                part = Part::Rows;
            } else if (h == "COLUMNS") {
                part = Part::Columns;
                rhs.assign(static_cast<size_t>(p.num_rows), 0.0);
                rng.assign(static_cast<size_t>(p.num_rows), 0.0);
            } else if (h == "RHS") {
                part = Part::Rhs;
            } else if (h == "RANGES") {
                part = Part::Ranges;
            } else if (h == "BOUNDS") {
                part = Part::Bounds;
            } else if (h == "ENDATA") {
                break;
            }
            continue;
        }
        switch (part) {
            case Part::Rows: {
                if (t.n < 2) break;
                const char k = t.f[0][0];
                const std::string name(t.f[1]);
                if (k == 'N') {
                    rowOf.emplace(name, -1);
                } else {
                    rowOf.emplace(name, p.num_rows++);
                    p.row_names.push_back(name);
                    kind.push_back(k);
                }
                break;
            }
            case Part::Columns: {
                if (t.n >= 3 && t.f[1] == "'MARKER'") {
                    if (t.f[2] == "'INTORG'") integerRun = true;
                    else if (t.f[2] == "'INTEND'") integerRun = false;
                    break;
                }
                if (t.n < 3) break;
                const Index j = column(t.f[0]);
                if (integerRun) p.col_type[static_cast<size_t>(j)] = VarType::Integer;
                for (int i = 1; i + 1 < t.n; i += 2) {
                    const std::string rname(t.f[i]);
                    auto it = rowOf.find(rname);
                    if (it == rowOf.end()) throw std::runtime_error("MPS: unknown row '" + rname + "'");
                    const Real v = number(t.f[i + 1]);
                    if (it->second < 0) p.obj[static_cast<size_t>(j)] = v;
                    else entries.push_back({it->second, j, v});
                }
                break;
            }
            case Part::Rhs:
                rowValues(t, rhs, true);
                break;
            case Part::Ranges:
                rowValues(t, rng, false);
                break;
            case Part::Bounds: {
                if (t.n < 3) break;
                const std::string_view b = t.f[0];
                const size_t j = static_cast<size_t>(column(t.f[2]));
                const bool hasValue = t.n >= 4;
                const Real v = hasValue ? number(t.f[3]) : 0.0;
                if (b == "LO" && hasValue) {
                    p.col_lower[j] = v;
                } else if (b == "UP" && hasValue) {
                    p.col_upper[j] = v;
                } else if (b == "FX" && hasValue) {
                    p.col_lower[j] = p.col_upper[j] = v;
                } else if (b == "FR") {
                    p.col_lower[j] = -kInf;
                    p.col_upper[j] = kInf;
                } else if (b == "MI") {
                    p.col_lower[j] = -kInf;
                } else if (b == "PL") {
                    p.col_upper[j] = kInf;
                } else if (b == "BV") {
                    p.col_lower[j] = 0.0;
                    p.col_upper[j] = 1.0;
                    p.col_type[j] = VarType::Binary;
                } else if (b == "LI" && hasValue) {
                    p.col_lower[j] = v;
                    p.col_type[j] = VarType::Integer;
                } else if (b == "UI" && hasValue) {
                    p.col_upper[j] = v;
                    p.col_type[j] = VarType::Integer;
                } else if (b == "SC" || b == "SI") {
                    p.col_lower[j] = 0.0;
                    p.col_type[j] = b == "SC" ? VarType::SemiContinuous : VarType::SemiInteger;
                }
                break;
            }
            case Part::Other:
                break;
        }
    }

    const size_t m = static_cast<size_t>(p.num_rows);
    rhs.resize(m, 0.0);
    rng.resize(m, 0.0);
    p.row_lower.resize(m);
    p.row_upper.resize(m);
    for (size_t i = 0; i < m; ++i) {
        const Real b = rhs[i], r = rng[i];
        switch (kind[i]) {
            case 'L':
                p.row_upper[i] = b;
                p.row_lower[i] = r != 0.0 ? b - std::abs(r) : -kInf;
                break;
            case 'G':
                p.row_lower[i] = b;
                p.row_upper[i] = r != 0.0 ? b + std::abs(r) : kInf;
                break;
            case 'E':
                p.row_lower[i] = r < 0.0 ? b + r : b;
                p.row_upper[i] = r > 0.0 ? b + r : b;
                break;
            default:
                p.row_lower[i] = -kInf;
                p.row_upper[i] = kInf;
                break;
        }
    }
    p.matrix = SparseMatrix(p.num_rows, p.num_cols, std::move(entries));
    return p;
}

void writeMps(const std::string& filename, const LpProblem& lp) {
    std::FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) throw std::runtime_error("MPS: cannot write " + filename);
    const Index m = lp.num_rows, n = lp.num_cols;
    auto rowName = [&](Index i) {
        return i < static_cast<Index>(lp.row_names.size()) && !lp.row_names[i].empty()
                   ? lp.row_names[i]
                   : "R" + std::to_string(i);
    };
    auto colName = [&](Index j) {
        return j < static_cast<Index>(lp.col_names.size()) && !lp.col_names[j].empty()
                   ? lp.col_names[j]
                   : "C" + std::to_string(j);
    };
    const double sgn = lp.sense == Sense::Maximize ? -1.0 : 1.0;  // MPS minimises
    std::fprintf(f, "NAME %s\nROWS\n N obj\n", lp.name.empty() ? "mipx" : lp.name.c_str());
    // Row kinds: E for lo == up, G for [lo, +inf) and for free rows (written
    // as [-1e30, +inf): an N row would be read as a second objective), L for
    // (-inf, up] and for ranged rows (rhs = up, range = up - lo, so lo comes
    // back as up - (up - lo), exact only when that subtraction is).
    auto rowKind = [&](Index i) {
        const Real lo = lp.row_lower[i], up = lp.row_upper[i];
        if (lo == up) return 'E';
        if (!std::isfinite(up)) return 'G';
        return 'L';
    };
    for (Index i = 0; i < m; ++i) std::fprintf(f, " %c %s\n", rowKind(i), rowName(i).c_str());
    // Column-major listing of the CSR matrix.
    std::vector<std::vector<std::pair<Index, Real>>> cols(static_cast<size_t>(n));
    const auto rs = lp.matrix.csr_row_starts();
    const auto ci = lp.matrix.csr_col_indices();
    const auto va = lp.matrix.csr_values();
    for (Index i = 0; i < m; ++i)
        for (Index k = rs[i]; k < rs[i + 1]; ++k) cols[static_cast<size_t>(ci[k])].push_back({i, va[k]});
    std::fprintf(f, "COLUMNS\n");
    bool inInt = false;
    for (Index j = 0; j < n; ++j) {
        const bool isInt = j < static_cast<Index>(lp.col_type.size()) &&
                           lp.col_type[j] != VarType::Continuous &&
                           lp.col_type[j] != VarType::SemiContinuous;
        if (isInt != inInt) {
            std::fprintf(f, " M%d 'MARKER' '%s'\n", j, isInt ? "INTORG" : "INTEND");
            inInt = isInt;
        }
        const std::string cn = colName(j);
        std::fprintf(f, " %s obj %.17g\n", cn.c_str(), sgn * lp.obj[j]);
        for (const auto& [i, v] : cols[static_cast<size_t>(j)])
            std::fprintf(f, " %s %s %.17g\n", cn.c_str(), rowName(i).c_str(), v);
    }
    if (inInt) std::fprintf(f, " M%d 'MARKER' 'INTEND'\n", n);
    std::fprintf(f, "RHS\n");
    if (lp.obj_offset != 0.0) std::fprintf(f, " rhs obj %.17g\n", -sgn * lp.obj_offset);
    std::vector<std::pair<Index, Real>> ranges;
    for (Index i = 0; i < m; ++i) {
        const Real lo = lp.row_lower[i], up = lp.row_upper[i];
        const char k = rowKind(i);
        const Real b = k == 'E' ? lo : k == 'G' ? (std::isfinite(lo) ? lo : -1e30) : up;
        if (b != 0.0) std::fprintf(f, " rhs %s %.17g\n", rowName(i).c_str(), b);
        if (k == 'L' && std::isfinite(lo)) ranges.push_back({i, up - lo});
    }
    if (!ranges.empty()) {
        std::fprintf(f, "RANGES\n");
        for (const auto& [i, r] : ranges) std::fprintf(f, " rng %s %.17g\n", rowName(i).c_str(), r);
    }
    std::fprintf(f, "BOUNDS\n");
    for (Index j = 0; j < n; ++j) {
        const Real lo = lp.col_lower[j], up = lp.col_upper[j];
        const std::string cn = colName(j);
        if (lo == up) {
            std::fprintf(f, " FX bnd %s %.17g\n", cn.c_str(), lo);
            continue;
        }
        if (!std::isfinite(lo) && !std::isfinite(up)) {
            std::fprintf(f, " FR bnd %s\n", cn.c_str());
            continue;
        }
        if (!std::isfinite(lo)) std::fprintf(f, " MI bnd %s\n", cn.c_str());
        else if (lo != 0.0) std::fprintf(f, " LO bnd %s %.17g\n", cn.c_str(), lo);
        if (std::isfinite(up)) std::fprintf(f, " UP bnd %s %.17g\n", cn.c_str(), up);
    }
    std::fprintf(f, "ENDATA\n");
    if (std::fclose(f) != 0) throw std::runtime_error("MPS: cannot write " + filename);
}

std::vector<SoluEntry> readSolu(const std::string& filename) {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error("solu: cannot open " + filename);
    std::vector<SoluEntry> out;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string tag, name, value;
        ss >> tag >> name;
        if (tag == "=inf=") {
            out.push_back({name, 0.0, true});
        } else if (tag == "=opt=") {
            ss >> value;
            out.push_back({name, value.empty() ? 0.0 : std::strtod(value.c_str(), nullptr), false});
        }
    }
    return out;
}

}  // namespace mipx
