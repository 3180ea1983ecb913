// Minimal host definitions under which the library's device FFT header
// (shtools_amd/csrc/sht_device.hpp) compiles as ordinary C++: one "thread"
// with blockDim.x = 1 runs every butterfly of a stage in turn, and the
// barrier is a no-op.  Used only by the CPU check of the FFT plans.
#pragma once
#include <algorithm>
#include <cmath>

#define __device__
#define __host__
#define __constant__
#define __forceinline__ inline

struct double2 {
    double x, y;
};
inline double2 make_double2(double a, double b) { return {a, b}; }
struct double4 {
    double x, y, z, w;
};
inline double4 make_double4(double a, double b, double c, double d) { return {a, b, c, d}; }

struct HostDim3 {
    unsigned x, y, z;
};
inline HostDim3 threadIdx{0, 0, 0};
inline HostDim3 blockDim{1, 1, 1};
inline void __syncthreads() {}

using std::frexp;
using std::ldexp;
using std::fma;
using std::max;
using std::min;
