// ftwa_herm_seq.hip -- batched fTWA integrator on the upper triangle of rho, one spin's rows in LDS
// at a time (FP64 engine; FP32 variant with two workgroups per CU).
//
// Same physics, records and RK4 "3/8" register form (y, A, p) as the FP32 upper-triangle kernel of
// ftwa_herm.hip (src/FHfTWASolver.cpp:366-426 records, :568-600 derivative, :629-661 RK4 3/8).  The
// [Y | X] rows of both spins do not fit LDS in FP64 (264 KB at L = 64), so each RK stage runs the two
// spins one after the other on the same L rows:
//   M  the spin's stage input Y = f(y, A, p) is written from the RK registers (upper triangle and its
//      mirror); in the spin-up pass the diagonal threads also write the Hartree fields of both spins
//      and the stage-input diagonals
//   A  X = T Y with lane = row, reading the z neighbour rows
//   B  k_ij = -i [conj(X_ji) - X_ij + (h_j - h_i) Y_ij] (k_ii = -2 Im X_ii) for the spin's slots, RK
//      registers updated in place
// with a barrier after M and after A; B of one spin and M of the next touch disjoint parts of the rows
// (X and h versus Y), so they need none.  Element slots are shared by the two spins (the same (i, j),
// the same LDS addresses), and each thread holds the RK registers of its S slots for both spins: the
// upper triangle halves the register state of the full-matrix stencil kernel, which spills in FP64.
// In FP32 the one-spin rows take 67 KB, so two workgroups share a CU and cover each other's barriers.
// For L <= 32 one workgroup holds P = 2 (L <= 16: 4) trajectories: trajectory t's site i is row t L + i
// of a block-diagonal system of P L <= 64 rows, phase A forms only the diagonal blocks (lane = row, the
// lane's column block), and the slots of all P trajectories share the waves.  Every element sees the
// same operands in the same order as with one trajectory per workgroup, and the energy sums of
// trajectory t run on waves 2t and 2t + 1 with its sites on the same lanes, so packing changes no bit.
// Null or on-site Coulomb interactions, no pulse, no density records, L <= 64 (FP32 64 < L <= 107: the
// column-blocked form in ftwa_herm_wide.hip); everything else runs on the stencil, upper-triangle FP32
// or dense kernels.  On the 8x8 workload the FP64 variant runs 1.40e7 trajectory-steps/s against 6.62e6
// for the full-matrix stencil kernel, the FP32 variant 3.12e7.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::V2;
using dev::wave_sum;

// LDS layout for precision Real: h [2 stage parities][2 spins][80], stage-input diagonals [2][80],
// energy partials [16 waves][4], site tables tM_ii / E_pe potential / U_ii / pdiag [4][64], neighbour
// hoppings [8][64] and row offsets [8][64] (ints), then L rows [Y (64 complex) | X (64 complex) | pad]
// and a spare row.
template <typename Real>
struct Lay {
    static constexpr int kC = 2 * static_cast<int>(sizeof(Real));  // bytes per complex
    static constexpr int kRowB = 129 * kC;
    static constexpr int kXo = 64 * kC;     // X_ij - Y_ij, bytes
    static constexpr int kHB = 0;
    static constexpr int kHSpin = 80 * static_cast<int>(sizeof(Real));  // entries 64..79 take padding lanes
    static constexpr int kHPar = 2 * kHSpin;
    static constexpr int kDg = 2 * kHPar;
    static constexpr int kEW = kDg + kHPar;
    static constexpr int kSite = kEW + 64 * static_cast<int>(sizeof(Real));
    static constexpr int kNbrT = kSite + 4 * 64 * static_cast<int>(sizeof(Real));
    static constexpr int kNbrK = kNbrT + 8 * 64 * static_cast<int>(sizeof(Real));
    static constexpr int kRows = kNbrK + 8 * 64 * 4;
    static_assert(kRows % 16 == 0, "rows are 16-byte aligned");
};
constexpr size_t kLdsLimit = 160 * 1024;

// LDS as a static array of the largest size, one per precision, whose offsets the compiler folds into
// the LDS instructions: the tables, then P L + 1 <= 65 rows (FP32: two workgroups per CU still fit; FP64:
// one).  The kernel names the array directly: a pointer variable captured by its lambdas costs VALU
// instructions.
constexpr int kSeqStaticLds = Lay<float>::kRows + 65 * Lay<float>::kRowB;
constexpr int kSeqStaticLds64 = Lay<double>::kRows + 65 * Lay<double>::kRowB;
__shared__ __align__(16) unsigned char seq_static_lds[kSeqStaticLds];
__shared__ __align__(16) unsigned char seq_static_lds64[kSeqStaticLds64];

template <int N>
using Ic = std::integral_constant<int, N>;

template <typename T>
__device__ __forceinline__ T ldl(const unsigned char* base, int off)
{
    return *reinterpret_cast<const T*>(base + off);
}

// MINB: workgroups per CU the register budget is sized for (2 in FP32, 1 in FP64): at least
// MINB * W / 4 waves per SIMD.  P: trajectories per workgroup (P L <= 64).
template <typename Real, int W, int S, int Z, int MINB, int P>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(MINB * W / 4)))
    herm_seq_kernel(DevSystem<Real> s, RunShape r, int t0, int t1, Real* __restrict__ state,
                    Real* __restrict__ rec)
{
    using C = V2<Real>;
    using Ly = Lay<Real>;
    constexpr int kRowB = Ly::kRowB, kXo = Ly::kXo, kHB = Ly::kHB, kHSpin = Ly::kHSpin, kHPar = Ly::kHPar,
                  kDg = Ly::kDg, kEW = Ly::kEW, kSite = Ly::kSite, kNbrT = Ly::kNbrT, kNbrK = Ly::kNbrK,
                  kRows = Ly::kRows, kCB = Ly::kC, kRB = static_cast<int>(sizeof(Real));
    const Real half = Real(0.5);
    constexpr int kNT = 64 * W;
    constexpr int NCG = 64 / W;  // columns per lane in phase A
#define SEQ_LDS (sizeof(Real) == 4 ? seq_static_lds : seq_static_lds64)  // this kernel's LDS (a global array)
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int PL = P * L;                       // rows held: site i of trajectory t is row t L + i
    const int n_off = dev::class_rows(L);       // element rows of one spin's strict upper triangle (class order)
    const int n_el = P * n_off;                 // the same for all held trajectories, trajectory-major
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * P;  // first trajectory of the workgroup
    const int spare = kRows + PL * kRowB;
    auto live = [&](int t) { return P == 1 || b0 + t < r.B; };  // the last workgroup may hold fewer
    Real* ew = reinterpret_cast<Real*>(SEQ_LDS + kEW);

    auto yat = [&](int row, int col) -> C& {
        return *reinterpret_cast<C*>(SEQ_LDS + kRows + row * kRowB + col * kCB);
    };

    // ---- clear rows and tables (columns past L stay zero) ----
    {
        const int n16 = (kRows + (PL + 1) * kRowB) / 16;
        for (int e = tid; e < n16; e += kNT) reinterpret_cast<float4*>(SEQ_LDS)[e] = float4{0, 0, 0, 0};
    }
    __syncthreads();
    for (int rw = tid; rw < PL; rw += kNT) {  // site tables per row
        const int i = P == 1 ? rw : rw % L;
        *reinterpret_cast<Real*>(SEQ_LDS + kSite + rw * kRB) = s.tdiag[i];
        *reinterpret_cast<Real*>(SEQ_LDS + kSite + (64 + rw) * kRB) = s.epot[i];
        *reinterpret_cast<Real*>(SEQ_LDS + kSite + (128 + rw) * kRB) = s.coulomb_type == 1 ? s.u_diag[i] : Real(0);
        *reinterpret_cast<Real*>(SEQ_LDS + kSite + (192 + rw) * kRB) = s.pdiag[i];
    }
    // neighbour tables for lane = row t L + i: the neighbour row t L + k and the start of the row's column
    // block t L (rows past P L: row 0 with t = 0)
    for (int e = tid; e < Z * 64; e += kNT) {
        const int n = e >> 6, rw = e & 63;
        const bool on = rw < PL;
        const int tb = P == 1 || !on ? 0 : rw / L, i = rw - tb * L;
        *reinterpret_cast<Real*>(SEQ_LDS + kNbrT + e * kRB) = on ? s.nbr_t[n * L + i] : Real(0);
        *reinterpret_cast<int*>(SEQ_LDS + kNbrK + e * 4) =
            kRows + (on ? tb * L + s.nbr_k[n * L + i] : 0) * kRowB + tb * L * kCB;
    }

    // ---- first launch of a run: the initial state is sampled here (src/FHfTWASolver.cpp:342-362), the
    // same FP64 values as launch_sample's, rounded to Real: spin up into the Y half of row i, spin down
    // into its X half (free until the first phase A); the slots pick them up below ----
    const bool sampled = t0 == 0 && r.sample_occ != nullptr;
    if (sampled) {
        const double* occ = r.sample_occ;
        // the strict upper triangle folded onto rows i0 < ceil(L/2): (i0, c0) with c0 > i0 is itself, with
        // c0 < i0 it stands for (L-1-i0, L-1-c0) (the middle row of an odd L only once), so nearly every lane
        // draws a pair
        const int half = (L + 1) / 2;
        for (int e = tid; e < P * half * L; e += kNT) {
            const int tb = P == 1 ? 0 : e / (half * L);
            const int e2 = e - tb * half * L;
            const int i0 = e2 / L, c0 = e2 - i0 * L;
            int i = i0, c = c0;
            if (c0 < i0 && 2 * i0 != L - 1) {
                i = L - 1 - i0;
                c = L - 1 - c0;
            } else if (c0 <= i0) {
                continue;
            }
            const uint32_t p = static_cast<uint32_t>(dev::pair_index(L, i, c));
            for (int sg = 0; sg < 2; ++sg) {
                double wr, wi, re, im;
                dev::initial_noise(r.seed, r.b0 + b0 + tb, sg, p, wr, wi);
                dev::sampled_offdiag(occ[sg * L + i], occ[sg * L + c], wr, wi, re, im);
                *reinterpret_cast<C*>(SEQ_LDS + kRows + (tb * L + i) * kRowB + (tb * L + c) * kCB + sg * kXo) =
                    C{static_cast<Real>(re), static_cast<Real>(im)};
            }
        }
        __syncthreads();
    }

    // ---- element slots of the strict upper triangles (class order, class_slot; element row er holds
    // trajectory er / n_off): LDS addresses of Y_ij (aij) and Y_ji (aji) in the shared rows, h offsets
    // ah = (t L + i)*8 | ((t L + j)*8) << 16; padding lanes point at the spare row and at h entry 64.  The
    // diagonal is held apart, row t L + i by thread t L + i (< P L): real RK registers only.
    int aij[S], aji[S], ah[S];
    C y0[S], A0[S], p0[S], y1[S], A1[S], p1[S];
    const C* gs = reinterpret_cast<const C*>(state) + b0 * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int er = q * W + w;
        aij[q] = spare;
        aji[q] = spare;
        ah[q] = 64 * kRB | (64 * kRB) << 16;
        y0[q] = A0[q] = p0[q] = y1[q] = A1[q] = p1[q] = C{0, 0};
        const int tb = P == 1 ? 0 : er / n_off, e2 = er - tb * n_off;
        int i = 0, jj = 0;
        if (er < n_el && dev::class_slot(L, 2 * e2 + (j >> 5), j & 31, i, jj)) {
            const int ri = tb * L + i, rj = tb * L + jj;
            aij[q] = kRows + ri * kRowB + rj * kCB;
            aji[q] = kRows + rj * kRowB + ri * kCB;
            ah[q] = ri * kRB | (rj * kRB) << 16;
            if (sampled) {
                y0[q] = *reinterpret_cast<const C*>(SEQ_LDS + aij[q]);
                y1[q] = *reinterpret_cast<const C*>(SEQ_LDS + aij[q] + kXo);
            } else if (live(tb)) {
                y0[q] = gs[tb * 2 * LL + i * L + jj];
                y1[q] = gs[tb * 2 * LL + LL + i * L + jj];
            }
        }
    }
    const bool dslot = tid < PL;
    const int da = dslot ? kRows + tid * kRowB + tid * kCB : spare;
    const int dt = P == 1 || !dslot ? 0 : tid / L, di = tid - dt * L;  // the diagonal's trajectory and site
    Real dy0 = 0, dy1 = 0;
    if (dslot) {  // diagonal: n - 1/2
        if (sampled) {
            dy0 = static_cast<Real>(r.sample_occ[di] - 0.5);
            dy1 = static_cast<Real>(r.sample_occ[L + di] - 0.5);
        } else if (live(dt)) {
            dy0 = gs[dt * 2 * LL + di * L + di].x;
            dy1 = gs[dt * 2 * LL + LL + di * L + di].x;
        }
    }
    Real dA0 = 0, dp0 = 0, dA1 = 0, dp1 = 0;
    // ---- phase A: lane j is row j (site j - t L of trajectory t); wave w covers columns t L + w + W*m of
    // the lane's block.  The neighbour rows and hoppings come from the LDS tables at each use (not held
    // in registers across the time loop); lanes past P L read row 0 with t = 0 and write the spare row.
    const int xb = (j < PL ? kRows + j * kRowB + (P == 1 ? 0 : j / L * L * kCB) : spare) + kXo + w * kCB;
    // neighbour n of row rw: LDS offset of Y(k_n, t L) and the hopping
    auto nbr_row = [&](int n, int rw, int& kb, Real& tv) {
        int o = n * 64 + rw;
        asm volatile("" : "+v"(o));  // opaque: loaded here, not hoisted out of the time loop
        kb = *reinterpret_cast<const int*>(SEQ_LDS + kNbrK + o * 4);
        tv = *reinterpret_cast<const Real*>(SEQ_LDS + kNbrT + o * kRB);
    };
    auto nbr = [&](int n, int& kb, Real& tv) {
        nbr_row(n, j, kb, tv);
        kb += w * kCB;
    };

    const Real h1 = static_cast<Real>(r.h), h3 = static_cast<Real>(r.h / 3.0), h6 = static_cast<Real>(r.h / 6.0),
               h8 = static_cast<Real>(r.h / 8.0);

    // ---- M: the spin's stage input into the rows; in the spin-up pass, h of both spins ----
    auto materialize = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HP = kHB + (ST & 1) * kHPar;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * W + w >= n_el) continue;
            const C& yq = SG ? y1[q] : y0[q];
            const C& Aq = SG ? A1[q] : A0[q];
            const C& pq = SG ? p1[q] : p0[q];
            const Real yx = dev::stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
            const Real yy = dev::stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
            *reinterpret_cast<C*>(SEQ_LDS + aij[q]) = C{yx, yy};
            *reinterpret_cast<C*>(SEQ_LDS + aji[q]) = C{yx, -yy};
        }
        if (dslot) {
            const Real yx = SG ? dev::stage_input<ST>(dy1, dA1, dp1, h1, h3, h6)
                               : dev::stage_input<ST>(dy0, dA0, dp0, h1, h3, h6);
            *reinterpret_cast<C*>(SEQ_LDS + da) = C{yx, Real(0)};
            if constexpr (SG == 0) {
                // h_up,i = pdiag_i + U_ii n_dn,i and h_dn,i = pdiag_i + U_ii n_up,i, n = Re y + 1/2
                const Real y1x = dev::stage_input<ST>(dy1, dA1, dp1, h1, h3, h6);
                int o = tid * kRB;
                asm volatile("" : "+v"(o));
                const Real u = ldl<Real>(SEQ_LDS, kSite + 128 * kRB + o), pd = ldl<Real>(SEQ_LDS, kSite + 192 * kRB + o);
                *reinterpret_cast<Real*>(SEQ_LDS + HP + o) = u * (y1x + half) + pd;
                *reinterpret_cast<Real*>(SEQ_LDS + HP + kHSpin + o) = u * (yx + half) + pd;
                *reinterpret_cast<Real*>(SEQ_LDS + kDg + o) = yx;
                *reinterpret_cast<Real*>(SEQ_LDS + kDg + kHSpin + o) = y1x;
            }
        }
    };

    // ---- A: X = T Y for the spin in the rows ----
    auto phase_a = [&]() {
        int kb[Z];
        Real tv[Z];
#pragma unroll
        for (int n = 0; n < Z; ++n) nbr(n, kb[n], tv[n]);
        // software pipelined where the registers allow (z = 4, or at most 2 slots per thread): the next
        // column's Z reads are issued before this column's X store (the compiler keeps reads behind an
        // earlier store to the same array), so their latency overlaps the arithmetic; the first column
        // lies inside row k for every wave.  8x8 lattice, FP32: 2.93e7 -> 3.12e7 trajectory-steps/s.
        constexpr bool kPipe = Z == 4 || S <= 2;
        C v[Z];
        if constexpr (kPipe) {
#pragma unroll
            for (int n = 0; n < Z; ++n) v[n] = ldl<C>(SEQ_LDS, kb[n]);
        }
#pragma unroll
        for (int m = 0; m < NCG; ++m) {
            if (m * W + w >= L) break;  // columns past L are never read (wave-uniform)
            const int co = m * W * kCB;  // immediate column offset
            Real ax = 0, ay = 0;
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C vn = kPipe ? v[n] : ldl<C>(SEQ_LDS, kb[n] + co);
                ax += tv[n] * vn.x;
                ay += tv[n] * vn.y;
            }
            if (kPipe && m + 1 < NCG && (m + 1) * W + w < L) {
#pragma unroll
                for (int n = 0; n < Z; ++n) v[n] = ldl<C>(SEQ_LDS, kb[n] + co + W * kCB);
            }
            *reinterpret_cast<C*>(SEQ_LDS + xb + co) = C{ax, ay};
        }
    };

    // ---- B: derivative of the spin's slots, RK registers updated ----
    auto phase_b = [&](auto stc, auto sgc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int SG = decltype(sgc)::value;
        constexpr int HS = kHB + (ST & 1) * kHPar + SG * kHSpin;
        // FP64: one slot ahead -- the next slot's reads go out before this slot's arithmetic, and the
        // memory clobber keeps the slot after that behind it (every slot in flight at once would not fit
        // the registers; 8x8 lattice: 1.38e7 -> 1.40e7 trajectory-steps/s).  FP32 (two workgroups per
        // CU, 128 registers): one slot at a time.
        if constexpr (sizeof(Real) == 8) {
            C nij{0, 0}, nji{0, 0};
            Real nhj = 0, nhi = 0;
            auto fetch = [&](int q) {
                nij = ldl<C>(SEQ_LDS, aij[q] + kXo);
                nji = ldl<C>(SEQ_LDS, aji[q] + kXo);
                int ahq = ah[q];  // opaque: decoded here, not kept decoded across the time loop
                asm volatile("" : "+v"(ahq));
                nhj = ldl<Real>(SEQ_LDS, HS + (ahq >> 16));
                nhi = ldl<Real>(SEQ_LDS, HS + (ahq & 0xffff));
            };
            if (w < n_el) fetch(0);
#pragma unroll
            for (int q = 0; q < S; ++q) {
                if (q * W + w >= n_el) break;  // slot rows grow with q (wave-uniform)
                const C xij = nij, xji = nji;
                const Real dh = nhj - nhi;
                if (q + 1 < S && (q + 1) * W + w < n_el) fetch(q + 1);
                asm volatile("" ::: "memory");
                C& yq = SG ? y1[q] : y0[q];
                C& Aq = SG ? A1[q] : A0[q];
                C& pq = SG ? p1[q] : p0[q];
                const Real ycx = dev::stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
                const Real ycy = dev::stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
                const Real cx = xji.x - xij.x + dh * ycx;
                const Real cy = -xji.y - xij.y + dh * ycy;
                dev::rk_update<ST>(yq.x, Aq.x, pq.x, cy, h8);
                dev::rk_update<ST>(yq.y, Aq.y, pq.y, -cx, h8);
            }
        } else {
#pragma unroll
            for (int q = 0; q < S; ++q) {
                if (q * W + w >= n_el) continue;
                asm volatile("" ::: "memory");
                const C xij = ldl<C>(SEQ_LDS, aij[q] + kXo);
                const C xji = ldl<C>(SEQ_LDS, aji[q] + kXo);
                int ahq = ah[q];  // opaque: decoded here, not kept decoded across the time loop
                asm volatile("" : "+v"(ahq));
                const Real dh = ldl<Real>(SEQ_LDS, HS + (ahq >> 16)) - ldl<Real>(SEQ_LDS, HS + (ahq & 0xffff));
                C& yq = SG ? y1[q] : y0[q];
                C& Aq = SG ? A1[q] : A0[q];
                C& pq = SG ? p1[q] : p0[q];
                const Real ycx = dev::stage_input<ST>(yq.x, Aq.x, pq.x, h1, h3, h6);
                const Real ycy = dev::stage_input<ST>(yq.y, Aq.y, pq.y, h1, h3, h6);
                // c = conj(X_ji) - X_ij + (h_j - h_i) Y_ij;  k = -i c
                const Real cx = xji.x - xij.x + dh * ycx;
                const Real cy = -xji.y - xij.y + dh * ycy;
                dev::rk_update<ST>(yq.x, Aq.x, pq.x, cy, h8);
                dev::rk_update<ST>(yq.y, Aq.y, pq.y, -cx, h8);
            }
        }
        if (dslot) {  // k_ii = -2 Im X_ii
            const Real k = Real(-2) * ldl<C>(SEQ_LDS, da + kXo).y;
            if constexpr (SG == 0) dev::rk_update<ST>(dy0, dA0, dp0, k, h8);
            else dev::rk_update<ST>(dy1, dA1, dp1, k, h8);
        }
    };

    // ---- records of the time point, one spin per pass (rows hold that spin's stage-0 input); the
    // per-wave energy partials go through ew, so nothing stays live between the passes ----
    // (trajectory t: hopping terms on wave 2t, on-site terms on wave 2t + 1, lane = site)
    Real* rr = nullptr;  // records of the workgroup's first trajectory; trajectory t's follow at t R
    auto record_part = [&](auto sgc) {
        constexpr int SG = decltype(sgc)::value;
        int jt = j, tt = tid;
        asm volatile("" : "+v"(jt), "+v"(tt));
        for (int e = tt; e < PL; e += kNT) {
            const int te = P == 1 ? 0 : e / L;
            if (live(te)) rr[te * r.R + SG * L + e - te * L] = yat(e, e).x;
        }
        if (tid < P && live(tid)) {
            const C c = yat(tid * L + r.ci, tid * L + r.cj);
            rr[tid * r.R + R2 + 4 + 2 * SG] = c.x;
            rr[tid * r.R + R2 + 5 + 2 * SG] = c.y;
        }
        const int tw = w >> 1;  // the trajectory whose energy terms this wave holds (if tw < P)
        const int rw = tw * L + jt;
        Real eth = 0, eu = 0, epe = 0;
        if (tw < P && (w & 1) == 0 && jt < L) {  // hopping terms: Y(k_n(i), i) = conj(Y(i, k_n(i)))
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                int kb;
                Real tv;
                nbr_row(n, rw, kb, tv);
                eth += tv * ldl<C>(SEQ_LDS, kb + jt * kCB).x;
            }
        }
        if (SG == 0 && tw < P && (w & 1) == 1 && jt < L) {  // on-site terms from the stage-input diagonals
            const Real nu = ldl<Real>(SEQ_LDS, kDg + rw * kRB) + half;
            const Real nd = ldl<Real>(SEQ_LDS, kDg + kHSpin + rw * kRB) + half;
            eth += ldl<Real>(SEQ_LDS, kSite + rw * kRB) * (nu + nd);
            epe += ldl<Real>(SEQ_LDS, kSite + (64 + rw) * kRB) * (nu + nd);
            eu += ldl<Real>(SEQ_LDS, kSite + (128 + rw) * kRB) * nu * nd;
        }
        if (w < 2 * P) {  // the other waves hold no energy terms
            eth = wave_sum(eth);
            eu = wave_sum(eu);
            epe = wave_sum(epe);
        }
        if (j == 0) {  // the same wave wrote this entry in the spin-up pass, two barriers earlier
            const Real e0 = SG ? ew[w * 4 + 0] : Real(0), e1 = SG ? ew[w * 4 + 1] : Real(0),
                       e3 = SG ? ew[w * 4 + 3] : Real(0);
            ew[w * 4 + 0] = e0 + eth;
            ew[w * 4 + 1] = e1 + eu;
            ew[w * 4 + 2] = 0;
            ew[w * 4 + 3] = e3 + epe;
        }
    };
    auto flush_energies = [&]() {
        if (tid < P && live(tid)) {  // one trajectory per thread: its two waves (P = 1: every wave, the rest 0)
            Real e4[4] = {0, 0, 0, 0};
            const int w0 = P == 1 ? 0 : 2 * tid, w1 = P == 1 ? W : 2 * tid + 2;
            for (int ww = w0; ww < w1; ++ww)
                for (int qq = 0; qq < 4; ++qq) e4[qq] += ew[ww * 4 + qq];
            for (int qq = 0; qq < 4; ++qq) rr[tid * r.R + R2 + qq] = e4[qq];
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
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b0) * R;
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
    C* go = reinterpret_cast<C*>(state) + b0 * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        if (aij[q] == spare) continue;
        const int ri = (ah[q] & 0xffff) / kRB, rj = (ah[q] >> 16) / kRB;
        const int tb = P == 1 ? 0 : ri / L, i = ri - tb * L, col = rj - tb * L;
        if (!live(tb)) continue;
        C* g = go + tb * 2 * LL;
        g[i * L + col] = y0[q];
        g[col * L + i] = C{y0[q].x, -y0[q].y};
        g[LL + i * L + col] = y1[q];
        g[LL + col * L + i] = C{y1[q].x, -y1[q].y};
    }
    if (dslot && live(dt)) {
        C* g = go + dt * 2 * LL;
        g[di * L + di] = C{dy0, Real(0)};
        g[LL + di * L + di] = C{dy1, Real(0)};
    }
}
#undef SEQ_LDS

template <typename Real>
size_t seq_lds_bytes(int rows)
{
    return Lay<Real>::kRows + static_cast<size_t>(rows + 1) * Lay<Real>::kRowB;
}

template <typename Real, int W, int S, int Z, int P>
hipError_t run_seq(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                   hipStream_t st)
{
    constexpr int MINB = sizeof(Real) == 4 ? 2 : 1;
    auto kern = herm_seq_kernel<Real, W, S, Z, MINB, P>;
    constexpr int kStatic = sizeof(Real) == 4 ? kSeqStaticLds : kSeqStaticLds64;  // the kernel's static array
    if (seq_lds_bytes<Real>(P * s.L) > static_cast<size_t>(kStatic)) return hipErrorInvalidValue;
    const int64_t grid = (r.B + P - 1) / P;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(64 * W), 0, st, s, r,
                       t0, t1, state, rec);
    return hipGetLastError();
}

template <typename Real, int W, int Z, int P>
hipError_t run_seq_s(int S, const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                     hipStream_t st)
{
    // P L <= 64: at most 32 element rows, so S <= 2 with 16 waves and S <= 4 with 8
    if constexpr (W == 16) {
        switch (S) {
            case 1: return run_seq<Real, 16, 1, Z, P>(s, r, t0, t1, state, rec, st);
            case 2: return run_seq<Real, 16, 2, Z, P>(s, r, t0, t1, state, rec, st);
            default: return hipErrorInvalidValue;
        }
    } else if constexpr (P > 1) {  // L <= 32: S <= 2
        switch (S) {
            case 1: return run_seq<Real, 8, 1, Z, P>(s, r, t0, t1, state, rec, st);
            case 2: return run_seq<Real, 8, 2, Z, P>(s, r, t0, t1, state, rec, st);
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (S) {
            case 1: return run_seq<Real, 8, 1, Z, 1>(s, r, t0, t1, state, rec, st);
            case 2: return run_seq<Real, 8, 2, Z, 1>(s, r, t0, t1, state, rec, st);
            case 3: return run_seq<Real, 8, 3, Z, 1>(s, r, t0, t1, state, rec, st);
            case 4: return run_seq<Real, 8, 4, Z, 1>(s, r, t0, t1, state, rec, st);
            default: return hipErrorInvalidValue;
        }
    }
}

template <typename Real, int W, int Z>
hipError_t run_seq_p(int P, int S, const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                     hipStream_t st)
{
    switch (P) {
        case 1: return run_seq_s<Real, W, Z, 1>(S, s, r, t0, t1, state, rec, st);
        case 2: return run_seq_s<Real, W, Z, 2>(S, s, r, t0, t1, state, rec, st);
        case 4: return run_seq_s<Real, W, Z, 4>(S, s, r, t0, t1, state, rec, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename Real>
bool herm_seq_fits(const DevSystem<Real>& s)
{
    return s.L >= 2 && s.L <= 64 && (s.z == 4 || s.z == 8) && !s.amp && s.coulomb_type != 2 &&
           seq_lds_bytes<Real>(s.L) <= kLdsLimit;
}

template <typename Real>
hipError_t launch_herm_seq(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                           int nt_hint, hipStream_t st)
{
    if (!herm_seq_fits(s) || r.dens || r.B <= 0) return hipErrorInvalidValue;
    // 8 waves by default.  FP64: 4 slots per thread at L = 64 fit the 256 registers of a one-workgroup-
    // per-CU shape without spilling (1.34e7 trajectory-steps/s on the 8x8 workload, against 1.12e7 with
    // 16 waves, whose 2 slots exceed 128 registers).  FP32: two 8-wave workgroups per CU.  16 waves on
    // request (nt_hint 1024); one trajectory per workgroup for every L on request (nt_hint 512).
    const int W = nt_hint == 1024 ? 16 : 8;
    // trajectories per workgroup: as many as fit 64 rows, up to 4 (one on request, nt_hint 512)
    const int P = nt_hint == 512 || s.L > 32 ? 1 : (s.L <= 16 ? 4 : 2);
    const int rows = P * dev::class_rows(s.L);  // element rows of the strict upper triangles (at most 32)
    const int S = std::max(1, (rows + W - 1) / W);
    if (W == 16) return s.z == 4 ? run_seq_p<Real, 16, 4>(P, S, s, r, t0, t1, state, rec, st)
                                 : run_seq_p<Real, 16, 8>(P, S, s, r, t0, t1, state, rec, st);
    return s.z == 4 ? run_seq_p<Real, 8, 4>(P, S, s, r, t0, t1, state, rec, st)
                    : run_seq_p<Real, 8, 8>(P, S, s, r, t0, t1, state, rec, st);
}

template bool herm_seq_fits<float>(const DevSystem<float>&);
template bool herm_seq_fits<double>(const DevSystem<double>&);
template hipError_t launch_herm_seq<float>(const DevSystem<float>&, const RunShape&, int, int, float*, float*, int,
                                           hipStream_t);
template hipError_t launch_herm_seq<double>(const DevSystem<double>&, const RunShape&, int, int, double*, double*,
                                            int, hipStream_t);

}  // namespace fpsamd
