// ftwa_herm_wide.hip -- batched FP32 fTWA integrator on the upper triangle for 64 < L <= 107, and for the
// balanced L = 64 lattices with one hopping value (the 8x8 periodic square): one spin's rows in LDS at a
// time, the commutator of each element formed directly from them.
//
// Same physics, records and RK4 "3/8" register form (y, A, p) as the spin-sequential kernel of
// ftwa_herm_seq.hip (src/FHfTWASolver.cpp:366-426 records, :568-600 derivative, :629-661 RK4 3/8).  Past
// L = 64 the one-spin [Y | X] rows of that kernel no longer fit LDS (160.8 KB at L = 100), so X = T Y is
// not stored: each element slot sums its own neighbour terms from the rows of Y,
//   M    the spin's stage input Y = f(y, A, p) from the RK registers (upper triangle, mirror, diagonal;
//        in the spin-up pass also the Hartree fields of both spins and the stage-input diagonals)
//   C    c_ij = sum_n t_{j,n} Y_{i,k_n(j)} - sum_n t_{i,n} Y_{k_n(i),j} + (h_j - h_i) Y_ij
//        (the first sum is conj(X_ji), Y being Hermitian; the second is X_ij), k_ij = -i c_ij,
//        k_ii = -2 Im X_ii, and the RK registers
// with a barrier after M and one after C (before the next M rewrites the rows).  Per element that is the
// same 2z complex x real multiply-adds as forming X, with the neighbour tables (row offsets, column
// offsets, hoppings; entry n of every site at the n-th most common shift k - i mod 32) read from LDS per
// slot.  LDS: rows of Y, 1 mod 32 complex long, in a static array.  Past L = 64 one 8-wave workgroup per
// CU (up to 256 VGPRs, S <= 12 slots per thread); at L = 64 two per CU (4 slots, 128 VGPRs).  Element
// slots in class order (dev::class_slot).  Null or on-site Coulomb interactions, no pulse, no density
// records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::wave_sum;
using C = float2;

constexpr int kW = 8;            // wavefronts per workgroup
constexpr int kNT = 64 * kW;
constexpr int kLP = 128;         // site capacity of the tables
constexpr int kCB = 8;           // bytes per complex
constexpr int kRB = 4;           // bytes per real
// LDS layout: h [2 stage parities][144 sites][2 spins], stage-input diagonals [2 spins][144], energy
// partials [8 waves][4], site tables tM_ii / E_pe potential / U_ii / pdiag [4][128], neighbour hoppings
// [8][128], neighbour row offsets [8][128] (ints), the neighbours' row and column offsets, 16 bits each, four
// to an 8-byte entry [2][128] each (entry p of site i holds k_{4p}(i) .. k_{4p+3}(i)), 16 spare bytes, then L
// rows and a spare row.  Every per-site entry the element slots read is 8 bytes, one complex: a slot holds
// 8 i and 8 j, which address its rows' tables and h as they address the columns of Y.
constexpr int kHSite = 2 * kRB;      // h of one site, both spins
constexpr int kHPar = 144 * kHSite;  // sites 128..143 take padding slots (always zero)
constexpr int kHB = 0;
constexpr int kHSpin = 144 * kRB;
static_assert(kHSite == kCB, "h entries are addressed like complex columns");
constexpr int kDg = 2 * kHPar;
constexpr int kEW = kDg + 2 * kHSpin;
constexpr int kSite = kEW + kW * 4 * kRB;
constexpr int kNbrT = kSite + 4 * kLP * kRB;
constexpr int kNbrK = kNbrT + 8 * kLP * kRB;
constexpr int kNbrR = kNbrK + 8 * kLP * 4;
constexpr int kNbrC = kNbrR + 2 * kLP * kCB;
constexpr int kZero = kNbrC + 2 * kLP * kCB;
constexpr int kRows = kZero + 16;
static_assert(kRows % 16 == 0, "rows are 16-byte aligned");
constexpr size_t kLdsLimit = 160 * 1024;

// complex elements per LDS row: L of Y and at least one more (the padding slots' column L), rounded up to
// 1 mod 32.  Y_ij then sits in the bank pair of i + j, the same for Y_ji: the lanes of a half-wave hold
// elements of distinct classes (i + j mod 32), and a neighbour read shifts them all alike.
__host__ __device__ inline int wide_row_len(int L) { return (L + 31) / 32 * 32 + 1; }

template <typename T>
__device__ __forceinline__ T ldl(const unsigned char* base, int off)
{
    return *reinterpret_cast<const T*>(base + off);
}

template <int N>
using Ic = std::integral_constant<int, N>;

// LDS as static arrays whose offsets the compiler folds into the LDS instructions: the tables, then L + 1
// rows of wide_row_len(L) complex.  L <= 64: at most 65 rows of 65 (51.7 KB, two workgroups per CU); past
// L = 64: at most 129 rows of 129 (151 KB, one workgroup per CU).  The kernel names the array directly (a
// pointer variable that the lambdas capture costs the L = 100 instance 3 % more VALU instructions).
constexpr int kStaticLds = kRows + 65 * 65 * kCB;
constexpr int kStaticLdsBig = kRows + (kLP + 1) * 129 * kCB;
__shared__ __align__(16) unsigned char wide_static_lds[kStaticLds];
__shared__ __align__(16) unsigned char wide_static_lds_big[kStaticLdsBig];

// UNI: every neighbour entry holds the same hopping s.hop_value (a regular lattice with one bond type),
// taken from a register instead of the table.
template <int S, int Z, bool UNI>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(S <= 4 ? 4 : 2)))
    herm_wide_kernel(DevSystem<float> s, RunShape r, int t0, int t1, float* __restrict__ state,
                     float* __restrict__ rec)
{
#define WIDE_LDS (S <= 4 ? wide_static_lds : wide_static_lds_big)  // this kernel's LDS (a global array)
    constexpr bool kRowBytes = S <= 4;  // L <= 64: neighbour row offsets in bytes fit 16 bits
    // one hopping value at L <= 64: neighbour terms summed first, then multiplied (2 fewer VALU
    // instructions per slot; at L = 100 the same form measured 2.5 % slower)
    constexpr bool kSumFirst = UNI && S <= 4;
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int RB = wide_row_len(L) * kCB;  // row stride, bytes
    const int n_off = dev::class_rows(L);
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t b = blockIdx.x;
    const int spare = kRows + L * RB;
    const float half = 0.5f;
    float* ew = reinterpret_cast<float*>(WIDE_LDS + kEW);
    // byte offset of a row: rows < 128 and strides < 1,400 B, so the full-rate 24-bit multiply (a 32-bit
    // multiply issues at a quarter of the rate, and the slots form these offsets at every use)
    auto rowb = [&](int row) { return static_cast<int>(__umul24(static_cast<unsigned>(row), static_cast<unsigned>(RB))); };
    auto yat = [&](int row, int col) -> C& { return *reinterpret_cast<C*>(WIDE_LDS + kRows + rowb(row) + col * kCB); };

    // ---- clear everything (h and X padding stay zero), then the tables ----
    {
        const int n16 = (kRows + (L + 1) * RB) / 16;
        for (int e = tid; e < n16; e += kNT) reinterpret_cast<float4*>(WIDE_LDS)[e] = float4{0, 0, 0, 0};
    }
    __syncthreads();
    for (int i = tid; i < L; i += kNT) {
        *reinterpret_cast<float*>(WIDE_LDS + kSite + i * kRB) = s.tdiag[i];
        *reinterpret_cast<float*>(WIDE_LDS + kSite + (kLP + i) * kRB) = s.epot[i];
        *reinterpret_cast<float*>(WIDE_LDS + kSite + (2 * kLP + i) * kRB) = s.coulomb_type == 1 ? s.u_diag[i] : 0.f;
        *reinterpret_cast<float*>(WIDE_LDS + kSite + (3 * kLP + i) * kRB) = s.pdiag[i];
    }
    // neighbour tables for row i (rows past L: row 0 with t = 0): the row byte offset of k_n(i), and k_n(i)
    // packed four to a word
    for (int e = tid; e < Z * kLP; e += kNT) {
        const int n = e / kLP, i = e - n * kLP;
        const bool on = i < L;
        const int k = on ? s.nbr_ks[n * L + i] : 0;
        *reinterpret_cast<float*>(WIDE_LDS + kNbrT + e * kRB) = on ? s.nbr_ts[n * L + i] : 0.f;
        *reinterpret_cast<int*>(WIDE_LDS + kNbrK + e * 4) = kRows + k * RB;
    }
    // 16-bit offsets of Y(k_n(i), 0) from the rows (bytes for L <= 64, else complex elements: at most
    // 127 x 129) and of Y(0, k_n(i)) within a row (bytes), four neighbours to an entry
    for (int e = tid; e < (Z / 4) * kLP; e += kNT) {
        const int p = e / kLP, i = e - p * kLP;
        uint32_t rw[2] = {0, 0}, cw[2] = {0, 0};
        if (i < L)
            for (int m = 0; m < 4; ++m) {
                const int k = s.nbr_ks[(4 * p + m) * L + i];
                rw[m / 2] |= static_cast<uint32_t>(kRowBytes ? k * RB : k * (RB / kCB)) << (16 * (m % 2));
                cw[m / 2] |= static_cast<uint32_t>(k * kCB) << (16 * (m % 2));
            }
        *reinterpret_cast<uint2*>(WIDE_LDS + kNbrR + e * kCB) = uint2{rw[0], rw[1]};
        *reinterpret_cast<uint2*>(WIDE_LDS + kNbrC + e * kCB) = uint2{cw[0], cw[1]};
    }

    // ---- element slots (class order): ah = 8 i | 8 j << 16 of the slot's element; padding lanes hold
    // i = j = L (the spare row; h entry L is zero).  The LDS addresses of Y_ij and Y_ji are formed from ah at
    // each use (registers: the RK state of both spins is 12 per slot).  The diagonal of site i is held by
    // thread i (< L): real RK registers only.
    // The first launch of a run samples the initial state here (src/FHfTWASolver.cpp:342-362), the same
    // FP64 values as launch_sample's rounded to float, one spin at a time through the rows.
    const bool sampled = t0 == 0 && r.sample_occ != nullptr;
    int ah[S];
    C y0[S], A0[S], p0[S], y1[S], A1[S], p1[S];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int er = q * kW + w;
        ah[q] = L * kCB | (L * kCB) << 16;
        y0[q] = A0[q] = p0[q] = y1[q] = A1[q] = p1[q] = C{0, 0};
        int i = 0, jj = 0;
        if (er < n_off && dev::class_slot(L, 2 * er + (j >> 5), j & 31, i, jj)) {
            ah[q] = i * kCB | (jj * kCB) << 16;
            if (!sampled) {
                y0[q] = gs[i * L + jj];
                y1[q] = gs[LL + i * L + jj];
            }
        }
    }
    auto sample_pass = [&](auto sgc) {
        constexpr int SG = decltype(sgc)::value;
        const double* occ = r.sample_occ + SG * L;
        // the strict upper triangle of spin SG into the rows, folded onto rows i0 < ceil(L/2): (i0, c0) with
        // c0 > i0 is itself, with c0 < i0 it stands for (L-1-i0, L-1-c0) (the middle row of an odd L only
        // once), so nearly every lane draws an element
        const int half = (L + 1) / 2;
        for (int e = tid; e < half * L; e += kNT) {
            const int i0 = e / L, c0 = e - i0 * L;
            int i = i0, c = c0;
            if (c0 < i0 && 2 * i0 != L - 1) {
                i = L - 1 - i0;
                c = L - 1 - c0;
            } else if (c0 <= i0) {
                continue;
            }
            double wr, wi, re, im;
            dev::initial_noise(r.seed, r.b0 + b, SG, static_cast<uint32_t>(dev::pair_index(L, i, c)), wr, wi);
            dev::sampled_offdiag(occ[i], occ[c], wr, wi, re, im);
            yat(i, c) = C{static_cast<float>(re), static_cast<float>(im)};
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int i = (ah[q] & 0xffff) / kCB, c = (ah[q] >> 16) / kCB;
            if (i < L) (SG ? y1[q] : y0[q]) = yat(i, c);
        }
        __syncthreads();  // the rows are rewritten next
    };
    if (sampled) {
        sample_pass(Ic<0>{});
        sample_pass(Ic<1>{});
    }
    // the slot's 8 i and 8 j, decoded where used (opaque: not kept across the time loop); 8 i times the row
    // length in complex elements is the byte offset of row i
    auto slot = [&](int q, int& i8, int& j8) {
        int a = ah[q];
        asm volatile("" : "+v"(a));
        i8 = a & 0xffff;
        j8 = a >> 16;
    };
    const unsigned RL = static_cast<unsigned>(wide_row_len(L));
    auto rowb8 = [&](int x8) { return static_cast<int>(__umul24(static_cast<unsigned>(x8), RL)); };
    const bool dslot = tid < L;
    const int da = dslot ? kRows + tid * RB + tid * kCB : spare;
    float dy0 = 0.f, dy1 = 0.f;  // diagonal: n - 1/2
    if (dslot) {
        dy0 = sampled ? static_cast<float>(r.sample_occ[tid] - 0.5) : gs[tid * L + tid].x;
        dy1 = sampled ? static_cast<float>(r.sample_occ[L + tid] - 0.5) : gs[LL + tid * L + tid].x;
    }
    float dA0 = 0, dp0 = 0, dA1 = 0, dp1 = 0;

    const float h1 = static_cast<float>(r.h), h3 = static_cast<float>(r.h / 3.0), h6 = static_cast<float>(r.h / 6.0),
                h8 = static_cast<float>(r.h / 8.0);

    // neighbour n of row i from the LDS tables (read at each use, not held across the time loop): the row
    // offset of k_n(i) (kb) or its column offset (kc), and t_{i,n}
    auto nbr = [&](int n, int i, int& kb, float& tv) {
        int o = n * kLP + i;
        asm volatile("" : "+v"(o));
        kb = *reinterpret_cast<const int*>(WIDE_LDS + kNbrK + o * 4);
        tv = *reinterpret_cast<const float*>(WIDE_LDS + kNbrT + o * kRB);
    };
    const float t_uni = s.hop_value;
    auto hop = [&](int n, int x8) -> float {  // t_{x,n} of site x = x8 / 8: the lattice's one value, or the table's
        if (UNI) return t_uni;
        int o = static_cast<int>(static_cast<unsigned>(x8) >> 1);
        asm volatile("" : "+v"(o));
        return ldl<float>(WIDE_LDS, kNbrT + n * kLP * kRB + o);
    };

    // ---- M: the spin's stage input into the rows; in the spin-up pass, h of both spins ----
    auto materialize = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HP = kHB + (ST & 1) * kHPar;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * kW + w >= n_off) continue;
            const C& yq = SG ? y1[q] : y0[q];
            const C& Aq = SG ? A1[q] : A0[q];
            const C& pq = SG ? p1[q] : p0[q];
            const float yx = dev::stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
            const float yy = dev::stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
            int i8, j8;
            slot(q, i8, j8);
            *reinterpret_cast<C*>(WIDE_LDS + kRows + rowb8(i8) + j8) = C{yx, yy};
            *reinterpret_cast<C*>(WIDE_LDS + kRows + rowb8(j8) + i8) = C{yx, -yy};
        }
        if (dslot) {
            const float yx = SG ? dev::stage_input<ST>(dy1, dA1, dp1, h1, h3, h6)
                                : dev::stage_input<ST>(dy0, dA0, dp0, h1, h3, h6);
            *reinterpret_cast<C*>(WIDE_LDS + da) = C{yx, 0.f};
            if constexpr (SG == 0) {
                // h_up,i = pdiag_i + U_ii n_dn,i and h_dn,i = pdiag_i + U_ii n_up,i, n = Re y + 1/2
                const float y1x = dev::stage_input<ST>(dy1, dA1, dp1, h1, h3, h6);
                int o = tid * kRB;
                asm volatile("" : "+v"(o));
                const float u = ldl<float>(WIDE_LDS, kSite + 2 * kLP * kRB + o);
                const float pd = ldl<float>(WIDE_LDS, kSite + 3 * kLP * kRB + o);
                *reinterpret_cast<float2*>(WIDE_LDS + HP + 2 * o) = float2{u * (y1x + half) + pd, u * (yx + half) + pd};
                *reinterpret_cast<float*>(WIDE_LDS + kDg + o) = yx;
                *reinterpret_cast<float*>(WIDE_LDS + kDg + kHSpin + o) = y1x;
            }
        }
    };

    // ---- C: each slot's commutator from the rows of Y, then its RK registers ----
    auto phase_c = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HS = kHB + (ST & 1) * kHPar + SG * kRB;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * kW + w >= n_off) continue;
            int i8, j8;
            slot(q, i8, j8);
            int ri = kRows + rowb8(i8);
            if constexpr (S <= 4) {
                // L <= 64: one multiply per slot, each neighbour's address one add from it (otherwise the
                // compiler forms every address as its own multiply-add)
                ri = rowb8(i8);
                asm volatile("" : "+v"(ri));
                ri += kRows;
            }
            const int rj = kRows + j8;
            float ax = 0, ay = 0, bx = 0, by = 0;  // conj(X_ji), X_ij
            uint32_t pr[Z / 2], pc[Z / 2];  // row offsets of i's neighbours, column offsets of j's
#pragma unroll
            for (int e = 0; e < Z / 4; ++e) {
                const uint2 a = ldl<uint2>(WIDE_LDS, kNbrR + e * kLP * kCB + i8);
                const uint2 c = ldl<uint2>(WIDE_LDS, kNbrC + e * kLP * kCB + j8);
                pr[2 * e] = a.x;
                pr[2 * e + 1] = a.y;
                pc[2 * e] = c.x;
                pc[2 * e + 1] = c.y;
            }
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const uint32_t ro = (pr[n / 2] >> (16 * (n % 2))) & 0xffff, co = (pc[n / 2] >> (16 * (n % 2))) & 0xffff;
                const float tj = hop(n, j8), ti = hop(n, i8);
                const C u = ldl<C>(WIDE_LDS, ri + co);                          // Y_{i, k_n(j)} = conj(Y_{k_n(j), i})
                const C v = ldl<C>(WIDE_LDS, rj + (kRowBytes ? ro : ro * kCB));  // Y_{k_n(i), j}
                if constexpr (kSumFirst) {  // the one hopping multiplies the sums below
                    ax = n == 0 ? u.x : ax + u.x;
                    ay = n == 0 ? u.y : ay + u.y;
                    bx = n == 0 ? v.x : bx + v.x;
                    by = n == 0 ? v.y : by + v.y;
                } else {
                    ax += tj * u.x;
                    ay += tj * u.y;
                    bx += ti * v.x;
                    by += ti * v.y;
                }
            }
        C& yq = SG ? y1[q] : y0[q];
            C& Aq = SG ? A1[q] : A0[q];
            C& pq = SG ? p1[q] : p0[q];
            const float dh = ldl<float>(WIDE_LDS, HS + j8) - ldl<float>(WIDE_LDS, HS + i8);  // h entries are 8 bytes
            const float ycx = dev::stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
            const float ycy = dev::stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
            // c = conj(X_ji) - X_ij + (h_j - h_i) Y_ij;  k = -i c
            const float cx = kSumFirst ? t_uni * (ax - bx) + dh * ycx : ax - bx + dh * ycx;
            const float cy = kSumFirst ? t_uni * (ay - by) + dh * ycy : ay - by + dh * ycy;
            dev::rk_update<ST>(yq.x, Aq.x, pq.x, cy, h8);
            dev::rk_update<ST>(yq.y, Aq.y, pq.y, -cx, h8);
        }
        if (dslot) {  // k_ii = -2 Im X_ii, X_ii = sum_n t_{i,n} Y_{k_n(i), i}
            float xi = 0;
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                int kb;
                float ti;
                nbr(n, tid, kb, ti);
                xi += ti * ldl<C>(WIDE_LDS, kb + tid * kCB).y;
            }
            const float k = -2.f * xi;
            if constexpr (SG == 0) dev::rk_update<ST>(dy0, dA0, dp0, k, h8);
            else dev::rk_update<ST>(dy1, dA1, dp1, k, h8);
        }
    };

    // ---- records of the time point, one spin per pass (rows hold that spin's stage-0 input) ----
    float* rr = nullptr;
    auto record_part = [&](auto sgc) {
        constexpr int SG = decltype(sgc)::value;
        int tt = tid;
        asm volatile("" : "+v"(tt));
        if (tt < L) rr[SG * L + tt] = yat(tt, tt).x;
        if (tid == 0) {
            const C c = yat(r.ci, r.cj);
            rr[R2 + 4 + 2 * SG] = c.x;
            rr[R2 + 5 + 2 * SG] = c.y;
        }
        float eth = 0, eu = 0, epe = 0;
        if (tt < L) {  // hopping terms of row tt: Y(k_n, tt) = conj(Y(tt, k_n))
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                int kb;
                float tv;
                nbr(n, tt, kb, tv);
                eth += tv * ldl<C>(WIDE_LDS, kb + tt * kCB).x;
            }
        }
        const int st = tt - 2 * 64;  // on-site terms on waves 2 and 3
        if (SG == 0 && st >= 0 && st < L) {
            const float nu = ldl<float>(WIDE_LDS, kDg + st * kRB) + half;
            const float nd = ldl<float>(WIDE_LDS, kDg + kHSpin + st * kRB) + half;
            eth += ldl<float>(WIDE_LDS, kSite + st * kRB) * (nu + nd);
            epe += ldl<float>(WIDE_LDS, kSite + (kLP + st) * kRB) * (nu + nd);
            eu += ldl<float>(WIDE_LDS, kSite + (2 * kLP + st) * kRB) * nu * nd;
        }
        eth = wave_sum(eth);
        eu = wave_sum(eu);
        epe = wave_sum(epe);
        if (j == 0) {  // the same wave wrote this entry in the spin-up pass, barriers earlier
            const float e0 = SG ? ew[w * 4 + 0] : 0.f, e1 = SG ? ew[w * 4 + 1] : 0.f, e3 = SG ? ew[w * 4 + 3] : 0.f;
            ew[w * 4 + 0] = e0 + eth;
            ew[w * 4 + 1] = e1 + eu;
            ew[w * 4 + 2] = 0;
            ew[w * 4 + 3] = e3 + epe;
        }
    };
    auto flush_energies = [&]() {
        if (tid == 0) {
            float e4[4] = {0, 0, 0, 0};
            for (int ww = 0; ww < kW; ++ww)
                for (int qq = 0; qq < 4; ++qq) e4[qq] += ew[ww * 4 + qq];
            for (int qq = 0; qq < 4; ++qq) rr[R2 + qq] = e4[qq];
        }
    };

    auto spin_pass = [&](auto stc, auto sgc, bool record) {
        materialize(stc, sgc);
        __syncthreads();
        if (record) record_part(sgc);
        phase_c(stc, sgc);
        __syncthreads();  // the next M rewrites rows that other slots have read
        if (record && decltype(sgc)::value == 1) flush_energies();
    };
    auto stage = [&](auto stc, bool record) {
        spin_pass(stc, Ic<0>{}, record);
        spin_pass(stc, Ic<1>{}, record);
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
        const int i = (ah[q] & 0xffff) / kCB, col = (ah[q] >> 16) / kCB;
        if (i >= L) continue;
        go[i * L + col] = y0[q];
        go[col * L + i] = C{y0[q].x, -y0[q].y};
        go[LL + i * L + col] = y1[q];
        go[LL + col * L + i] = C{y1[q].x, -y1[q].y};
    }
    if (dslot) {
        go[tid * L + tid] = C{dy0, 0.f};
        go[LL + tid * L + tid] = C{dy1, 0.f};
    }
}
#undef WIDE_LDS

size_t wide_lds_bytes(int L) { return kRows + static_cast<size_t>(L + 1) * wide_row_len(L) * kCB; }

template <int S, int Z, bool UNI>
hipError_t run_wide(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                    hipStream_t st)
{
    auto kern = herm_wide_kernel<S, Z, UNI>;
    if (wide_lds_bytes(s.L) > static_cast<size_t>(S <= 4 ? kStaticLds : kStaticLdsBig)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(kNT), 0, st, s, r, t0, t1, state, rec);
    return hipGetLastError();
}

template <int Z, bool UNI>
hipError_t run_wide_s(int S, const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                      hipStream_t st)
{
    switch (S) {  // L = 64: 32 element rows, 4 slots; 64 < L <= 107: 33 to 91 element rows, 5 to 12 slots
        case 4: return run_wide<4, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 5: return run_wide<5, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 6: return run_wide<6, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 7: return run_wide<7, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 8: return run_wide<8, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 9: return run_wide<9, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 10: return run_wide<10, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 11: return run_wide<11, Z, UNI>(s, r, t0, t1, state, rec, st);
        case 12: return run_wide<12, Z, UNI>(s, r, t0, t1, state, rec, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// L <= 64 where it measured faster than the spin-sequential kernel of ftwa_herm_seq.hip: four whole slot
// rows per wave (L = 64), four neighbours, one hopping value (the 8x8 periodic square: 3.25e7 against
// 3.08e7 trajectory-steps/s when first measured, 3.56e7 now; with unequal slot rows per wave, kagome 48
// and bcc 54, it lost)
bool herm_wide_direct_preferred(const DevSystem<float>& s)
{
    return s.L <= 64 && dev::class_rows(s.L) == 4 * kW && s.z == 4 && s.hop_uniform && !s.amp &&
           s.coulomb_type != 2 && wide_lds_bytes(s.L) <= kLdsLimit;
}

bool herm_wide_fits(const DevSystem<float>& s)
{
    return ((s.L > 64 && s.L <= kLP && (s.z == 4 || s.z == 8) && !s.amp && s.coulomb_type != 2 &&
             wide_lds_bytes(s.L) <= kLdsLimit) ||
            herm_wide_direct_preferred(s));
}

hipError_t launch_herm_wide(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                            hipStream_t st)
{
    if (!herm_wide_fits(s) || r.dens || r.B <= 0 || r.B > 0x7fffffff) return hipErrorInvalidValue;
    const int S = std::max(4, (dev::class_rows(s.L) + kW - 1) / kW);
    if (s.hop_uniform)
        return s.z == 4 ? run_wide_s<4, true>(S, s, r, t0, t1, state, rec, st)
                        : run_wide_s<8, true>(S, s, r, t0, t1, state, rec, st);
    return s.z == 4 ? run_wide_s<4, false>(S, s, r, t0, t1, state, rec, st)
                    : run_wide_s<8, false>(S, s, r, t0, t1, state, rec, st);
}

}  // namespace fpsamd
