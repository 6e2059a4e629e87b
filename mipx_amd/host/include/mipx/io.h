// io.h -- MPS input / output and .solu reference values, the file interface
// of the reference's include/mipx/io.h:11-28 (readMps, writeMps, SoluEntry,
// readSolu) for the PDLP command line and real-instance runs.
#pragma once

#include <string>
#include <vector>

#include "mipx/lp_problem.h"

namespace mipx {

// Reads an MPS file (free or fixed columns separated by blanks); a name
// ending in ".gz" is inflated with zlib.  Throws std::runtime_error on I/O
// errors and on a COLUMNS entry that names an unknown row.
LpProblem readMps(const std::string& filename);

// Writes the LP as free-format MPS with round-trip precision.
void writeMps(const std::string& filename, const LpProblem& problem);

struct SoluEntry {
    std::string name;
    Real value;
    bool is_infeasible = false;
};

// "=opt= name value" and "=inf= name" lines of a .solu file.
std::vector<SoluEntry> readSolu(const std::string& filename);

}  // namespace mipx
