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

// Compensated (Neumaier) FP64 accumulator: the sum is s + c.
struct Neumaier {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x)
    {
        const double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
};

// Packed upper triangle (j >= i) of an L x L matrix, row-major: element (i, j) sits at
// packed_row(L, i) + j.
__device__ __forceinline__ int packed_row(int L, int i) { return i * L - (i * (i + 1)) / 2; }

// Density part of a trajectory record: Re rho_sigma(i, j), j >= i, for both spins (spin-major, each
// packed as above), written row by row: wave `wave` of `nwave` takes rows wave, wave + nwave, ... of
// the 2L spin-rows, its lanes the columns.  re_at(row, j) returns Re rho of spin-row `row` (spin
// row / L, site row % L), column j.
template <typename Real, typename F>
__device__ __forceinline__ void record_density(Real* dst, int L, int lane, int wave, int nwave, F re_at)
{
    const int P = L * (L + 1) / 2;
    for (int row = wave; row < 2 * L; row += nwave) {
        const int sg = row >= L ? 1 : 0, i = row - sg * L;
        Real* d = dst + sg * P + packed_row(L, i);
        for (int jj = i + lane; jj < L; jj += 64) d[jj] = re_at(row, jj);
    }
}

}  // namespace dev
}  // namespace fpsamd
