// ftwa_device.hpp -- device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace fpsamd {
namespace dev {

// Complex value of precision Real as a two-component vector (8- or 16-byte LDS/global accesses).
template <typename Real>
using V2 = typename std::conditional<std::is_same<Real, float>::value, float2, double2>::type;

// Butterfly sum over a 64-lane wavefront: fixed order, every lane gets the same value.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace dev
}  // namespace fpsamd
