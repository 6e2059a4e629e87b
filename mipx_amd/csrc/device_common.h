// device_common.h -- shared HIP helpers for the MI355X PDLP kernels.
//
// All device code is built with -ffp-contract=off so that a*b+c is a rounded
// multiply followed by a rounded add, as in the reference's ISO C++ build; the
// per-element arithmetic of the kernels then matches the reference's CPU loops
// bit for bit, and only reduction order can differ.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

namespace mpx {

constexpr int kBlock = 256;  // 4 wave64 per workgroup
constexpr int kWave = 64;
// XCDs of an MI355X; workgroups of a launch are dealt to them round robin.
constexpr int kNumXcd = 8;

// Thread-local last error text of the C-ABI.
void setError(const std::string& msg);
const char* lastError();

#define MPX_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::mpx::setError(std::string(#expr) + ": " + hipGetErrorString(e_));            \
            throw ::mpx::DeviceError{e_};                                                  \
        }                                                                                  \
    } while (0)

// A failed HIP call (code: its error) or an internal consistency failure
// (code hipSuccess).
struct DeviceError {
    hipError_t code = hipSuccess;
};

inline int ceilDiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// MPX_GRAPHS=0 in the environment: every captured launch sequence (PDHG
// blocks, power-iteration groups, batch and server rounds) is launched
// directly instead -- same kernels, same results -- for tools that cannot
// follow graph launches.
inline bool graphsDisabled() {
    static const bool off = [] {
        const char* v = std::getenv("MPX_GRAPHS");
        return v != nullptr && v[0] == '0';
    }();
    return off;
}

// std::clamp(v, lo, hi) of libstdc++: (v < lo) ? lo : (hi < v) ? hi : v.
template <typename T>
__host__ __device__ inline T clampRef(T v, T lo, T hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}
// std::max(a, b) = (a < b) ? b : a ; std::min(a, b) = (b < a) ? b : a.
template <typename T>
__host__ __device__ inline T maxRef(T a, T b) {
    return (a < b) ? b : a;
}
template <typename T>
__host__ __device__ inline T minRef(T a, T b) {
    return (b < a) ? b : a;
}

// Sum across the G lanes of an aligned lane group (butterfly: every lane ends
// with the bitwise-identical value because each step adds the same two
// operands, only commuted).
template <int G, typename T>
__device__ inline T groupSum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// Deterministic block reduction of NF fields (sums, or maxima where isMax).
// Every thread passes its NF values; thread 0 of the block gets the result in
// out[].  Order: butterfly inside each wave, then waves 0..3 sequentially.
template <int NF>
__device__ inline void blockReduce(double (&v)[NF], const bool (&isMax)[NF], double (&out)[NF]) {
    __shared__ double red[NF][kBlock / kWave];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        double x = v[f];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(x, off, kWave);
            x = isMax[f] ? maxRef(x, o) : x + o;
        }
        v[f] = x;
    }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane == 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) red[f][wave] = v[f];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            double a = red[f][0];
            for (int w = 1; w < kBlock / kWave; ++w) a = isMax[f] ? maxRef(a, red[f][w]) : a + red[f][w];
            out[f] = a;
        }
    }
    __syncthreads();
}

// A pointer loaded from memory (a descriptor field) is generic to the
// compiler, which then emits flat loads; every such pointer here refers to
// device global memory, and this cast says so (global loads and stores).
template <typename T>
__device__ inline __attribute__((address_space(1))) T* asGlobal(T* p) {
    return (__attribute__((address_space(1))) T*)(p);
}

}  // namespace mpx
