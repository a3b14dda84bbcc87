// ftwa_herm64.hip -- FP64 batched fTWA integrator on the upper triangle of rho.
//
// Same physics, records and RK4 "3/8" register form (y, A, p) as the FP32 upper-triangle kernel of
// ftwa_herm.hip (src/FHfTWASolver.cpp:366-426 records, :568-600 derivative, :629-661 RK4 3/8), for FP64
// trajectories.  In FP64 the [Y | X] rows of both spins (264 KB at L = 64) do not fit the 160 KB of LDS,
// so each RK stage runs the two spins one after the other on the same L rows:
//   M  the spin's stage input Y = f(y, A, p) is written from the RK registers (upper triangle and its
//      mirror); in the spin-up pass every diagonal slot also writes the Hartree field of the other spin
//      and the stage-input diagonal
//   A  X = T Y with lane = row, reading the z neighbour rows
//   B  k_ij = -i [conj(X_ji) - X_ij + (h_j - h_i) Y_ij] (k_ii = -2 Im X_ii) for the spin's slots, RK
//      registers updated in place
// with a barrier after M and after A; B of one spin and M of the next touch disjoint parts of the rows
// (X and h versus Y), so they need none.  Element slots are shared by the two spins (the same (i, j),
// the same LDS addresses), and each thread holds the RK registers of its S slots for both spins: the
// upper triangle halves the register state of the full-matrix stencil kernel, which spills in FP64.
// Null or on-site Coulomb interactions, no pulse, no density records, L <= 64; everything else runs
// on the FP64 stencil or dense kernels.  On the 8x8 workload it runs 1.34e7 trajectory-steps/s against
// 6.62e6 for the full-matrix stencil kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::wave_sum;
using C = double2;

constexpr int kRowB = 129 * 16;     // LDS row: Y (64 complex FP64), X (64 complex FP64), pad
constexpr int kXo = 64 * 16;        // X_ij - Y_ij, bytes
constexpr int kHB = 0;              // h [2 stage parities][2 spins][80] doubles
constexpr int kHSpin = 80 * 8;      // bytes per spin (entries 64..79 take padding lanes)
constexpr int kHPar = 2 * kHSpin;   // bytes per parity
constexpr int kDg = 2 * kHPar;      // stage-input diagonals [2 spins][80] doubles (records)
constexpr int kEW = kDg + kHPar;    // energy partials [16 waves][4] doubles
constexpr int kSite = kEW + 512;    // tM_ii [64], potential of E_pe [64], U_ii [64], pdiag [64] doubles
constexpr int kNbrT = kSite + 2048; // hopping t_n(i) [8][64] doubles
constexpr int kNbrK = kNbrT + 4096; // byte offset of row k_n(i) [8][64] ints
constexpr int kRows = kNbrK + 2048; // first row; L rows, then a spare row
constexpr size_t kLdsLimit = 160 * 1024;

template <int N>
using Ic = std::integral_constant<int, N>;

__device__ __forceinline__ C ld2(const unsigned char* base, int off)
{
    return *reinterpret_cast<const C*>(base + off);
}
__device__ __forceinline__ double ld1(const unsigned char* base, int off)
{
    return *reinterpret_cast<const double*>(base + off);
}

// Stage input of RK stage ST from the registers: y, y + h p/3, y + h (A - 3p)/6, y + h p.
template <int ST>
__device__ __forceinline__ double stage_input(double y, double A, double p, double h1, double h3, double h6)
{
    if constexpr (ST == 0) return y;
    else if constexpr (ST == 1) return y + h3 * p;
    else if constexpr (ST == 2) return y + h6 * (A - 3.0 * p);
    else return y + h1 * p;
}

// RK registers after stage ST with derivative k (the 3/8 rule in the (y, A, p) form).
template <int ST>
__device__ __forceinline__ void rk_update(double& y, double& A, double& p, double k, double h8)
{
    if constexpr (ST == 0) {
        A = k;
        p = k;
    } else if constexpr (ST == 1) {
        A += 3.0 * k;
        p -= k;
    } else if constexpr (ST == 2) {
        A += 3.0 * k;
        p += k;
    } else {
        A += k;
        y += h8 * A;
    }
}

template <int W, int S, int Z>
__global__ void __launch_bounds__(64 * W)
    herm64_kernel(DevSystem<double> s, RunShape r, int t0, int t1, double* __restrict__ state,
                  double* __restrict__ rec)
{
    constexpr int kNT = 64 * W;
    constexpr int NCG = 64 / W;  // columns per lane in phase A
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int NS = L * (L - 1) / 2;        // strict upper elements per spin
    const int n_off = (NS + 63) / 64;      // element rows of one spin's strict upper triangle
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t b = blockIdx.x;
    const int spare = kRows + L * kRowB;
    double* ew = reinterpret_cast<double*>(smem + kEW);

    auto yat = [&](int row, int col) -> C& {
        return *reinterpret_cast<C*>(smem + kRows + row * kRowB + col * 16);
    };

    // ---- clear rows and tables (columns past L stay zero) ----
    {
        const int n16 = (kRows + (L + 1) * kRowB) / 16;
        for (int e = tid; e < n16; e += kNT) reinterpret_cast<C*>(smem)[e] = C{0.0, 0.0};
    }
    __syncthreads();
    for (int i = tid; i < L; i += kNT) {
        *reinterpret_cast<double*>(smem + kSite + i * 8) = s.tdiag[i];
        *reinterpret_cast<double*>(smem + kSite + 512 + i * 8) = s.epot[i];
        *reinterpret_cast<double*>(smem + kSite + 1024 + i * 8) = s.coulomb_type == 1 ? s.u_diag[i] : 0.0;
        *reinterpret_cast<double*>(smem + kSite + 1536 + i * 8) = s.pdiag[i];
    }
    // neighbour tables for lane = row i (lanes past L: row 0 with t = 0)
    for (int e = tid; e < Z * 64; e += kNT) {
        const int n = e >> 6, i = e & 63;
        const bool on = i < L;
        *reinterpret_cast<double*>(smem + kNbrT + e * 8) = on ? s.nbr_t[n * L + i] : 0.0;
        *reinterpret_cast<int*>(smem + kNbrK + e * 4) = kRows + (on ? s.nbr_k[n * L + i] : 0) * kRowB;
    }

    // ---- element slots of the strict upper triangle: LDS addresses of Y_ij (aij) and Y_ji (aji) in
    // the shared rows, h offsets ah = i*8 | (j*8) << 16; padding lanes point at the spare row and at h
    // entry 64.  The diagonal is held apart, site i by thread i (< L) of wave 0: real RK registers only.
    int aij[S], aji[S], ah[S];
    C y0[S], A0[S], p0[S], y1[S], A1[S], p1[S];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int er = q * W + w;
        aij[q] = spare;
        aji[q] = spare;
        ah[q] = 64 * 8 | (64 * 8) << 16;
        y0[q] = A0[q] = p0[q] = y1[q] = A1[q] = p1[q] = C{0.0, 0.0};
        if (er < n_off) {
            const int u = er * 64 + j;
            if (u < NS) {
                // row i of the strict upper triangle: the largest i with i*(2L - i - 1)/2 <= u
                const float c = static_cast<float>(2 * L - 1);
                int i = static_cast<int>(0.5f * (c - sqrtf(fmaxf(c * c - 8.0f * static_cast<float>(u), 0.0f))));
                if (i < 0) i = 0;
                while (i > 0 && i * (2 * L - i - 1) / 2 > u) --i;
                while ((i + 1) * (2 * L - i - 2) / 2 <= u) ++i;
                const int jj = i + 1 + (u - i * (2 * L - i - 1) / 2);
                aij[q] = kRows + i * kRowB + jj * 16;
                aji[q] = kRows + jj * kRowB + i * 16;
                ah[q] = i * 8 | (jj * 8) << 16;
                y0[q] = gs[i * L + jj];
                y1[q] = gs[LL + i * L + jj];
            }
        }
    }
    const bool dslot = tid < L;
    const int da = dslot ? kRows + tid * kRowB + tid * 16 : spare;
    double dy0 = dslot ? gs[tid * L + tid].x : 0.0, dy1 = dslot ? gs[LL + tid * L + tid].x : 0.0;
    double dA0 = 0.0, dp0 = 0.0, dA1 = 0.0, dp1 = 0.0;
    // ---- phase A: lane j is row i = j; wave w covers columns w + W*m.  The neighbour rows and
    // hoppings come from the LDS tables at each use (not held in registers across the time loop);
    // lanes past L read row 0 with t = 0 and write the spare row.
    const int xb = (j < L ? kRows + j * kRowB : spare) + kXo + w * 16;
    auto nbr = [&](int n, int& kb, double& tv) {
        int o = n * 64 + j;
        asm volatile("" : "+v"(o));  // opaque: loaded here, not hoisted out of the time loop
        kb = *reinterpret_cast<const int*>(smem + kNbrK + o * 4) + w * 16;
        tv = *reinterpret_cast<const double*>(smem + kNbrT + o * 8);
    };

    const double h1 = r.h, h3 = r.h / 3.0, h6 = r.h / 6.0, h8 = r.h / 8.0;

    // ---- M: the spin's stage input into the rows; in the spin-up pass, h of both spins ----
    auto materialize = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HP = kHB + (ST & 1) * kHPar;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * W + w >= n_off) continue;
            const C& yq = SG ? y1[q] : y0[q];
            const C& Aq = SG ? A1[q] : A0[q];
            const C& pq = SG ? p1[q] : p0[q];
            const double yx = stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
            const double yy = stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
            *reinterpret_cast<C*>(smem + aij[q]) = C{yx, yy};
            *reinterpret_cast<C*>(smem + aji[q]) = C{yx, -yy};
        }
        if (dslot) {
            const double yx = SG ? stage_input<ST>(dy1, dA1, dp1, h1, h3, h6)
                                 : stage_input<ST>(dy0, dA0, dp0, h1, h3, h6);
            *reinterpret_cast<C*>(smem + da) = C{yx, 0.0};
            if constexpr (SG == 0) {
                // h_up,i = pdiag_i + U_ii n_dn,i and h_dn,i = pdiag_i + U_ii n_up,i, n = Re y + 1/2
                const double y1x = stage_input<ST>(dy1, dA1, dp1, h1, h3, h6);
                int o = tid * 8;
                asm volatile("" : "+v"(o));
                const double u = ld1(smem, kSite + 1024 + o), pd = ld1(smem, kSite + 1536 + o);
                *reinterpret_cast<double*>(smem + HP + o) = u * (y1x + 0.5) + pd;
                *reinterpret_cast<double*>(smem + HP + kHSpin + o) = u * (yx + 0.5) + pd;
                *reinterpret_cast<double*>(smem + kDg + o) = yx;
                *reinterpret_cast<double*>(smem + kDg + kHSpin + o) = y1x;
            }
        }
    };

    // ---- A: X = T Y for the spin in the rows ----
    auto phase_a = [&]() {
        int kb[Z];
        double tv[Z];
#pragma unroll
        for (int n = 0; n < Z; ++n) nbr(n, kb[n], tv[n]);
#pragma unroll
        for (int m = 0; m < NCG; ++m) {
            const int co = m * W * 16;  // immediate column offset
            double ax = 0.0, ay = 0.0;
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C v = ld2(smem, kb[n] + co);
                ax += tv[n] * v.x;
                ay += tv[n] * v.y;
            }
            *reinterpret_cast<C*>(smem + xb + co) = C{ax, ay};
        }
    };

    // ---- B: derivative of the spin's slots, RK registers updated ----
    auto phase_b = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HS = kHB + (ST & 1) * kHPar + SG * kHSpin;
        // one slot at a time (the memory clobber keeps the next slot's loads behind this slot's
        // arithmetic): in FP64 the reads of every slot in flight at once would not fit the registers
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * W + w >= n_off) continue;
            asm volatile("" ::: "memory");
            const C xij = ld2(smem, aij[q] + kXo);
            const C xji = ld2(smem, aji[q] + kXo);
            int ahq = ah[q];  // opaque: decoded here, not kept decoded across the time loop
            asm volatile("" : "+v"(ahq));
            const double dh = ld1(smem, HS + (ahq >> 16)) - ld1(smem, HS + (ahq & 0xffff));
            C& yq = SG ? y1[q] : y0[q];
            C& Aq = SG ? A1[q] : A0[q];
            C& pq = SG ? p1[q] : p0[q];
            const double ycx = stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
            const double ycy = stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
            // c = conj(X_ji) - X_ij + (h_j - h_i) Y_ij;  k = -i c
            const double cx = xji.x - xij.x + dh * ycx;
            const double cy = -xji.y - xij.y + dh * ycy;
            rk_update<ST>(yq.x, Aq.x, pq.x, cy, h8);
            rk_update<ST>(yq.y, Aq.y, pq.y, -cx, h8);
        }
        if (dslot) {  // k_ii = -2 Im X_ii
            const double k = -2.0 * ld2(smem, da + kXo).y;
            if constexpr (SG == 0) rk_update<ST>(dy0, dA0, dp0, k, h8);
            else rk_update<ST>(dy1, dA1, dp1, k, h8);
        }
    };

    // ---- records of the time point, one spin per pass (rows hold that spin's stage-0 input); the
    // per-wave energy partials go through ew, so nothing stays live between the passes ----
    double* rr = nullptr;
    auto record_part = [&](auto sgc) {
        constexpr int SG = decltype(sgc)::value;
        int jt = j, tt = tid;
        asm volatile("" : "+v"(jt), "+v"(tt));
        for (int e = tt; e < L; e += kNT) rr[SG * L + e] = yat(e, e).x;
        if (tid == 0) {
            const C c = yat(r.ci, r.cj);
            rr[R2 + 4 + 2 * SG] = c.x;
            rr[R2 + 5 + 2 * SG] = c.y;
        }
        double eth = 0.0, eu = 0.0, epe = 0.0;
        if (w == 0 && jt < L) {  // hopping terms: Y(k_n(i), i) = conj(Y(i, k_n(i)))
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                int kb;
                double tv;
                nbr(n, kb, tv);
                eth += tv * ld2(smem, kb + jt * 16).x;
            }
        }
        if (SG == 0 && w == 1 && jt < L) {  // on-site terms from the stage-input diagonals
            const double nu = ld1(smem, kDg + jt * 8) + 0.5;
            const double nd = ld1(smem, kDg + kHSpin + jt * 8) + 0.5;
            eth += ld1(smem, kSite + jt * 8) * (nu + nd);
            epe += ld1(smem, kSite + 512 + jt * 8) * (nu + nd);
            eu += ld1(smem, kSite + 1024 + jt * 8) * nu * nd;
        }
        eth = wave_sum(eth);
        eu = wave_sum(eu);
        epe = wave_sum(epe);
        if (j == 0) {  // the same wave wrote this entry in the spin-up pass, two barriers earlier
            const double e0 = SG ? ew[w * 4 + 0] : 0.0, e1 = SG ? ew[w * 4 + 1] : 0.0, e3 = SG ? ew[w * 4 + 3] : 0.0;
            ew[w * 4 + 0] = e0 + eth;
            ew[w * 4 + 1] = e1 + eu;
            ew[w * 4 + 2] = 0.0;
            ew[w * 4 + 3] = e3 + epe;
        }
    };
    auto flush_energies = [&]() {
        if (tid == 0) {
            double e4[4] = {0, 0, 0, 0};
            for (int ww = 0; ww < W; ++ww)
                for (int qq = 0; qq < 4; ++qq) e4[qq] += ew[ww * 4 + qq];
            for (int qq = 0; qq < 4; ++qq) rr[R2 + qq] = e4[qq];
        }
    };

    auto stage = [&](auto stc, bool record) {
        materialize(stc, Ic<0>{});
        __syncthreads();
        if (record) record_part(Ic<0>{});
        phase_a();
        __syncthreads();
        phase_b(stc, Ic<0>{});
        materialize(stc, Ic<1>{});
        __syncthreads();
        if (record) record_part(Ic<1>{});
        phase_a();
        __syncthreads();
        if (record) flush_energies();
        phase_b(stc, Ic<1>{});
    };

    const int64_t Bn = r.B;
    const int R = r.R;
    for (int t = t0; t < t1; ++t) {
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b) * R;
        if (t == r.T - 1) {  // no step after the last record changes an output
            materialize(Ic<0>{}, Ic<0>{});
            __syncthreads();
            record_part(Ic<0>{});
            __syncthreads();
            materialize(Ic<0>{}, Ic<1>{});
            __syncthreads();
            record_part(Ic<1>{});
            __syncthreads();
            flush_energies();
            break;
        }
        for (int sub = 0; sub < r.nsub; ++sub) {
            stage(Ic<0>{}, sub == 0);
            stage(Ic<1>{}, false);
            stage(Ic<2>{}, false);
            stage(Ic<3>{}, false);
        }
    }
    // ---- state back to HBM, from the RK registers ----
    C* go = reinterpret_cast<C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        if (aij[q] == spare) continue;
        const int i = (ah[q] & 0xffff) >> 3, col = (ah[q] >> 16) >> 3;
        go[i * L + col] = y0[q];
        go[col * L + i] = C{y0[q].x, -y0[q].y};
        go[LL + i * L + col] = y1[q];
        go[LL + col * L + i] = C{y1[q].x, -y1[q].y};
    }
    if (dslot) {
        go[tid * L + tid] = C{dy0, 0.0};
        go[LL + tid * L + tid] = C{dy1, 0.0};
    }
}

size_t herm64_lds_bytes(int L) { return kRows + static_cast<size_t>(L + 1) * kRowB; }

template <int W, int S, int Z>
hipError_t run64(const DevSystem<double>& s, const RunShape& r, int t0, int t1, double* state, double* rec,
                 hipStream_t st)
{
    auto kern = herm64_kernel<W, S, Z>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(64 * W), herm64_lds_bytes(s.L), st, s, r, t0, t1,
                       state, rec);
    return hipGetLastError();
}

template <int W, int Z>
hipError_t run64_s(int S, const DevSystem<double>& s, const RunShape& r, int t0, int t1, double* state, double* rec,
                   hipStream_t st)
{
    // L <= 64: at most 32 element rows, so S <= 2 with 16 waves and S <= 4 with 8
    if constexpr (W == 16) {
        switch (S) {
            case 1: return run64<16, 1, Z>(s, r, t0, t1, state, rec, st);
            case 2: return run64<16, 2, Z>(s, r, t0, t1, state, rec, st);
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (S) {
            case 1: return run64<8, 1, Z>(s, r, t0, t1, state, rec, st);
            case 2: return run64<8, 2, Z>(s, r, t0, t1, state, rec, st);
            case 3: return run64<8, 3, Z>(s, r, t0, t1, state, rec, st);
            case 4: return run64<8, 4, Z>(s, r, t0, t1, state, rec, st);
            default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace

bool herm64_fits(const DevSystem<double>& s)
{
    return s.L >= 2 && s.L <= 64 && (s.z == 4 || s.z == 8) && !s.amp && s.coulomb_type != 2 &&
           herm64_lds_bytes(s.L) <= kLdsLimit;
}

hipError_t launch_herm64(const DevSystem<double>& s, const RunShape& r, int t0, int t1, double* state, double* rec,
                         int nt_hint, hipStream_t st)
{
    if (!herm64_fits(s) || r.dens || r.B <= 0) return hipErrorInvalidValue;
    // 8 waves by default: 4 slots per thread at L = 64 fit the 256 registers of a two-wave-per-SIMD
    // workgroup without spilling (1.34e7 trajectory-steps/s on the 8x8 workload, against 1.12e7 with
    // 16 waves, whose 2 slots exceed 128 registers); 16 waves on request (nt_hint 1024)
    const int W = nt_hint == 1024 ? 16 : 8;
    const int NS = s.L * (s.L - 1) / 2;
    const int rows = (NS + 63) / 64;  // element rows of the strict upper triangle
    const int S = std::max(1, (rows + W - 1) / W);
    if (W == 16) return s.z == 4 ? run64_s<16, 4>(S, s, r, t0, t1, state, rec, st)
                                 : run64_s<16, 8>(S, s, r, t0, t1, state, rec, st);
    return s.z == 4 ? run64_s<8, 4>(S, s, r, t0, t1, state, rec, st) : run64_s<8, 8>(S, s, r, t0, t1, state, rec, st);
}

}  // namespace fpsamd
