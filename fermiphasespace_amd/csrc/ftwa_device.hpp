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

// Index of the pair (a, c), a < c, among the strictly upper entries of an L x L matrix, row by row.
__device__ __forceinline__ int64_t pair_index(int L, int a, int c)
{
    return static_cast<int64_t>(a) * (2 * L - a - 1) / 2 + (c - a - 1);
}

// Complex noise value for (trajectory, spin, pair) of the initial state: counter (pair, spin, traj lo,
// traj hi), key seed.
__device__ __forceinline__ void initial_noise(uint64_t seed, int64_t traj, int spin, uint32_t pair, double& re,
                                              double& im)
{
    uint32_t c[4] = {pair, static_cast<uint32_t>(spin), static_cast<uint32_t>(static_cast<uint64_t>(traj)),
                     static_cast<uint32_t>(static_cast<uint64_t>(traj) >> 32)};
    box_muller(c, seed, re, im);
}

// Sampled element rho_sigma(a, c), a < c, of the initial state before any rotation
// (src/FHfTWASolver.cpp:342-362): S_ac w with S_ac = sqrt((n_a + n_c - 2 n_a n_c) / 2) and w the complex
// normal (wr, wi); rho_sigma(c, a) is its conjugate.  The one formula every sampling path uses.
__device__ __forceinline__ void sampled_offdiag(double na, double nc, double wr, double wi, double& re, double& im)
{
    const double sij = sqrt((na + nc - 2.0 * na * nc) / 2.0);
    re = sij * wr;
    im = sij * wi;
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

// Element slots of the upper-triangle integrators in class order.  Lane l of element row er holds the p-th
// element (p = 2 er + l / 32) of class c = l % 32, where class c lists the strict-upper (i, j), i < j, with
// (i + j) mod 32 = c, ordered by i + j, then i.  With LDS rows 129 complex long (258 dwords = 2 modulo 64
// banks), both Y_ij and Y_ji (and X_ij, X_ji) of an element sit in bank pair (i + j) mod 32: a pass of 32
// lanes with one element of every class touches each bank pair once.  Row-major packing put several short
// rows of the triangle into one pass and conflicted up to 8-way there (modelled: 22 % extra LDS cycles per
// stage at L = 64).  FP64 elements take four banks and 16 lanes per pass, whose classes are distinct modulo
// 16 as well.
__host__ __device__ inline bool class_slot(int L, int p, int c, int& i, int& j)
{
    for (int s = c; s <= 2 * L - 3; s += 32) {
        if (s < 1) continue;
        const int lo = s - L + 1 > 0 ? s - L + 1 : 0, hi = (s - 1) / 2;
        const int n = hi - lo + 1;
        if (n <= 0) continue;
        if (p < n) {
            i = lo + p;
            j = s - i;
            return true;
        }
        p -= n;
    }
    return false;
}

// Element rows of the class order: half the largest class, rounded up (32 at L = 64).
__host__ __device__ inline int class_rows(int L)
{
    int most = 0;
    for (int c = 0; c < 32; ++c) {
        int n = 0;
        for (int s = c; s <= 2 * L - 3; s += 32) {
            const int lo = s - L + 1 > 0 ? s - L + 1 : 0, hi = (s - 1) / 2;
            if (s >= 1 && hi >= lo) n += hi - lo + 1;
        }
        most = n > most ? n : most;
    }
    return (most + 1) / 2;
}

// Stage input of RK stage ST from the registers: y, y + h p/3, y + h (A - 3p)/6, y + h p.
template <int ST, typename Real>
__device__ __forceinline__ Real stage_input(Real y, Real A, Real p, Real h1, Real h3, Real h6)
{
    if constexpr (ST == 0) return y;
    else if constexpr (ST == 1) return y + h3 * p;
    else if constexpr (ST == 2) return y + h6 * (A - Real(3) * p);
    else return y + h1 * p;
}

// RK registers after stage ST with derivative k (the 3/8 rule in the (y, A, p) form).
template <int ST, typename Real>
__device__ __forceinline__ void rk_update(Real& y, Real& A, Real& p, Real k, Real h8)
{
    if constexpr (ST == 0) {
        A = k;
        p = k;
    } else if constexpr (ST == 1) {
        A += Real(3) * k;
        p -= k;
    } else if constexpr (ST == 2) {
        A += Real(3) * k;
        p += k;
    } else {
        A += k;
        y += h8 * A;
    }
}

}  // namespace dev
}  // namespace fpsamd
