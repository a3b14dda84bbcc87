// ftwa_device.hpp -- device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter c in, random words out, key (k0, k1).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Complex normal value from one Philox block of counter c and key seed: Box-Muller on two 53-bit
// uniforms, E|w|^2 = 1 with independent N(0, 1/2) real and imaginary parts.
__device__ __forceinline__ void box_muller(uint32_t c[4], uint64_t seed, double& re, double& im)
{
    philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const double two_m53 = 1.1102230246251565404e-16;
    const double u1 = (static_cast<double>(c[0] >> 5) * 67108864.0 + static_cast<double>(c[1] >> 6) + 1.0) * two_m53;
    const double u2 = (static_cast<double>(c[2] >> 5) * 67108864.0 + static_cast<double>(c[3] >> 6)) * two_m53;
    const double r = sqrt(-log(u1));
    double sn, cs;
    sincos(6.283185307179586232 * u2, &sn, &cs);
    re = r * cs;
    im = r * sn;
}

// GPSR diffusion noise xi_k(p) (k = 0, 1) of RK step `step` of trajectory `traj`, unscaled.  Bit 31 of
// the second counter word keeps this stream apart from the initial-state noise (words 0 and 1 there).
__device__ __forceinline__ void gpsr_noise(uint64_t seed, int64_t traj, int64_t step, int k, int p, double& re,
                                           double& im)
{
    uint32_t c[4] = {static_cast<uint32_t>(static_cast<uint64_t>(step)),
                     0x80000000u | (static_cast<uint32_t>(k) << 16) | static_cast<uint32_t>(p),
                     static_cast<uint32_t>(static_cast<uint64_t>(traj)),
                     static_cast<uint32_t>(static_cast<uint64_t>(traj) >> 32)};
    box_muller(c, seed, re, im);
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
