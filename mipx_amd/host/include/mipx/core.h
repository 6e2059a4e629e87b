// core.h -- scalar types and enums of the mipx PDLP boundary.
// Same names and meaning as the reference's include/mipx/core.h:7-24 so that
// code written against mipx's PdlpSolver compiles against this library.
#pragma once

namespace mipx {

using Real = double;  // FP64 values (reference core.h:7)
using Int = int;      // counters (core.h:8)
using Index = int;    // 32-bit indices (core.h:9)

// Same order as the C-ABI's MPX_STATUS_* codes.
enum class Status { Optimal, Infeasible, Unbounded, IterLimit, TimeLimit, NodeLimit, Error };

enum class Sense { Minimize, Maximize };

enum class VarType { Continuous, Integer, Binary, SemiContinuous, SemiInteger };

}  // namespace mipx
