// Stand-in for <hip/hip_runtime.h> used ONLY by the CPU build of the device
// stages under tests/host_stages: the device qualifiers become plain C++ and the
// math functions resolve to the host C library, so the same stage source runs
// on the CPU under sanitizers and against the oracle.  The product library is
// always built by hipcc against the real header.
#pragma once
#include <cmath>
#include <cstddef>
#define __device__
#define __host__
#define __constant__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
using std::copysign;
using std::fma;
using std::fmax;
using std::fmin;
using std::frexp;
using std::ldexp;
using std::rint;
using std::exp;
using std::fabs;
using std::log;
using std::pow;
using std::sqrt;
