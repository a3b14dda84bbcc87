// ftwa_herm.hip -- FP32 batched fTWA integrator exploiting the Hermiticity of rho.
//
// Same physics and record format as ftwa_stencil.hip (src/FHfTWASolver.cpp:366-426, :568-600, :629-661),
// organised around X = T rho, the hopping applied from the left:
//
//   rho T - T rho = X^dagger - X      (rho and T Hermitian: (rho T)_ij = conj(X_ji))
//   k_ij = -i [ conj(X_ji) - X_ij + (h_j - h_i) rho_ij ]          k_ii = -2 Im X_ii
//
// One workgroup integrates one trajectory.  Per RK stage:
//   phase A  X_ij = sum_{k in nb(i)} t_ik Y_kj with lane = row i and a wave-uniform column: every
//            lane keeps the LDS addresses of its neighbour rows and t_ik in registers, the column
//            enters as an immediate offset, so the stage costs no address arithmetic.  X goes to
//            LDS.                                                                          -> barrier
//   phase B  the strict upper triangle of both spins is packed into 64-lane "element rows", followed
//            by the diagonals; element row r belongs to wave r % W, slot r / W.  Each thread keeps, per
//            slot, the RK4 "3/8" registers (y, A, p) and the LDS byte addresses of its element and of
//            the mirror, so a stage is: read X_ij, X_ji, h_i, h_j, form k, update, write Y_ij and
//            Y_ji = conj(Y_ij).  Slot kinds are wave-uniform; diagonal slots also write the other
//            spin's next Hartree potential.                                               -> barrier
// Only the upper triangle carries RK state (half the registers and update arithmetic of the
// full-matrix kernel) and the state stays exactly Hermitian (real diagonal).
//
// LDS (bytes):
//   [0, 1280)     h double buffer [2][160] floats: entry sg*64 + i, entries 128.. absorb padding lanes
//   [1280, 1536)  energy partials [W][4]
//   [1536, 2304)  per-site constants of the energies: tM_ii, potential, U_ii ([64] each)
//   [2304, ...)   rows r = sg*L + i (2L of them, plus a spare row 2L), 129 complex each:
//                 Y_r,0..63 | X_r,0..63 | pad.  X_ij sits 512 bytes after Y_ij; the odd stride keeps
//                 the transposed (column) accesses of phase B on distinct banks.
//   then          pulse hoppings t (1 + e^{i phi}) [Z][L] complex
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::wave_sum;
using C = float2;

constexpr int kRowBytes = 129 * 8;  // LDS row: Y (64 complex), X (64 complex), pad
constexpr int kXOff = 64 * 8;       // X_ij - Y_ij, bytes
constexpr int kHB = 0;              // h buffers
constexpr int kHBStride = 160 * 4;  // bytes per h buffer
constexpr int kEW = 1280;           // energy partials
constexpr int kSite = 1536;         // tM_ii [64], potential [64], U_ii [64]
constexpr int kRows = 2304;         // first row
constexpr size_t kLdsLimit = 160 * 1024;
// LDS of both kernels, a static array of the largest size (2L + 1 <= 129 rows and a pulse table of 8 x 64
// complex) whose offsets the compiler folds into the LDS instructions; one workgroup per CU at any
// L <= 64 as with a dynamic array.  The kernels name the array directly: a pointer variable captured by
// their lambdas costs VALU instructions.
constexpr int kHermStaticLds = kRows + 129 * kRowBytes + 8 * 64 * 8;
__shared__ __align__(16) unsigned char herm_static_lds[kHermStaticLds];

template <int N>
using Ic = std::integral_constant<int, N>;

__device__ __forceinline__ C lds_c(const unsigned char* base, int off)
{
    return *reinterpret_cast<const C*>(base + off);
}
__device__ __forceinline__ float lds_f(const unsigned char* base, int off)
{
    return *reinterpret_cast<const float*>(base + off);
}

// EPTA: column groups per wavefront in phase A (half per spin); S: element slots per thread in phase B;
// DENS: records carry the packed real density matrices (a compile-time switch, so the variant without
// them keeps its register allocation).
template <int W, int EPTA, int S, int Z, bool PULSE, bool DENS>
__global__ void __launch_bounds__(64 * W)
    herm_kernel(DevSystem<float> s, RunShape r, int t0, int t1, float* __restrict__ state, float* __restrict__ rec)
{
    constexpr int kNT = 64 * W;
#define smem herm_static_lds
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int NS = L * (L - 1) / 2;             // strict upper elements per spin
    const int n_off = (2 * NS + 63) / 64;       // element rows of off-diagonal elements
    const int n_dia = (R2 + 63) / 64;           // element rows of diagonal elements
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t b = blockIdx.x;
    const bool general = s.coulomb_type == 2;
    const int spare = kRows + R2 * kRowBytes;  // spare row: target of padding lanes and rows
    C* Tc = reinterpret_cast<C*>(smem + kRows + (R2 + 1) * kRowBytes);  // [Z][L]
    float* ew = reinterpret_cast<float*>(smem + kEW);

    auto yat = [&](int row, int col) -> C& {
        return *reinterpret_cast<C*>(smem + kRows + row * kRowBytes + col * 8);
    };

    // ---- clear the rows (columns past L stay finite), h buffers and tables ----
    {
        const int n16 = (kRows + (R2 + 1) * kRowBytes) / 16;
        for (int e = tid; e < n16; e += kNT) reinterpret_cast<float4*>(smem)[e] = float4{0, 0, 0, 0};
    }
    __syncthreads();
    for (int i = tid; i < L; i += kNT) {
        *reinterpret_cast<float*>(smem + kSite + i * 4) = s.tdiag[i];
        *reinterpret_cast<float*>(smem + kSite + 256 + i * 4) = s.epot[i];
        *reinterpret_cast<float*>(smem + kSite + 512 + i * 4) = s.u_diag[i];
    }

    // ---- element slots: LDS byte addresses of Y_ij (aij) and of the mirror Y_ji (aji), h offsets ----
    // Off-diagonal: ah = (sg*64 + i)*4 | (sg*64 + j)*4 << 16.  Diagonal: aji = aij,
    // ah = ((1 - sg)*64 + i)*4 (the other spin's next h), and the unused imaginary RK registers hold
    // the constants of that h: A.y = U_ii (on-site Coulomb, else 0), p.y = pdiag_i.  Padding lanes
    // point at the spare row and at h entry 128.
    int aij[S], aji[S], ah[S];
    C y[S], A[S], p[S];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int er = q * W + w;
        aij[q] = spare;
        aji[q] = spare;
        ah[q] = 128 * 4;
        y[q] = C{0, 0};
        A[q] = C{0, 0};
        p[q] = C{0, 0};
        if (er < n_off) {
            const int e = er * 64 + j;
            if (e < 2 * NS) {
                const int sg = e >= NS ? 1 : 0;
                const int u = e - sg * NS;
                // row i of the strict upper triangle: the largest i with i*(2L - i - 1)/2 <= u
                const float c = static_cast<float>(2 * L - 1);
                int i = static_cast<int>(0.5f * (c - sqrtf(fmaxf(c * c - 8.0f * static_cast<float>(u), 0.0f))));
                if (i < 0) i = 0;
                while (i > 0 && i * (2 * L - i - 1) / 2 > u) --i;
                while ((i + 1) * (2 * L - i - 2) / 2 <= u) ++i;
                const int jj = i + 1 + (u - i * (2 * L - i - 1) / 2);
                aij[q] = kRows + (sg * L + i) * kRowBytes + jj * 8;
                aji[q] = kRows + (sg * L + jj) * kRowBytes + i * 8;
                ah[q] = (sg * 64 + i) * 4 | ((sg * 64 + jj) * 4) << 16;
                y[q] = gs[sg * LL + i * L + jj];
                *reinterpret_cast<C*>(smem + aij[q]) = y[q];
                *reinterpret_cast<C*>(smem + aji[q]) = C{y[q].x, -y[q].y};
            }
        } else if (er < n_off + n_dia) {
            const int d = (er - n_off) * 64 + j;
            if (d < R2) {
                const int sg = d >= L ? 1 : 0;
                const int i = d - sg * L;
                aij[q] = kRows + (sg * L + i) * kRowBytes + i * 8;
                aji[q] = aij[q];
                ah[q] = ((1 - sg) * 64 + i) * 4;
                y[q] = C{gs[sg * LL + i * L + i].x, 0.0f};
                A[q].y = s.coulomb_type == 1 ? s.u_diag[i] : 0.0f;
                p[q].y = s.pdiag[i];
                *reinterpret_cast<C*>(smem + aij[q]) = y[q];
            }
        }
    }
    // ---- phase-A addresses: lane j is row i = j of each spin; wave w covers columns w + W*mm ----
    // kb{sg}[n]: byte address of Y(sg*L + k_n(i), w); xb{sg}: of X(sg*L + i, w).  Lanes past L read
    // row 0 with t = 0 and write the spare row.
    const int il = min(j, L - 1);
    int kb0[Z], kb1[Z];
    float tv[Z];
#pragma unroll
    for (int n = 0; n < Z; ++n) {
        const int k = j < L ? s.nbr_k[n * L + il] : 0;
        kb0[n] = kRows + k * kRowBytes + w * 8;
        kb1[n] = kb0[n] + L * kRowBytes;
        tv[n] = j < L ? s.nbr_t[n * L + il] : 0.0f;
    }
    const int xb0 = (j < L ? kRows + il * kRowBytes : spare) + kXOff + w * 8;
    const int xb1 = (j < L ? kRows + (L + il) * kRowBytes : spare) + kXOff + w * 8;

    const float h1 = static_cast<float>(r.h);
    const float h3 = static_cast<float>(r.h / 3.0), h6 = static_cast<float>(r.h / 6.0),
                h8 = static_cast<float>(r.h / 8.0);
    int64_t gst = static_cast<int64_t>(t0) * r.nsub * 4;

    auto pulse_table = [&](int64_t stage_index) {
        if constexpr (PULSE) {
            const float a = static_cast<float>(s.amp[stage_index]);
            for (int e = tid; e < Z * L; e += kNT) {
                float sn, cs;
                __sincosf(a * s.nbr_pd[e], &sn, &cs);
                const float t = s.nbr_t[e];
                Tc[e] = C{t * (1.0f + cs), t * sn};
            }
        }
    };
    // h of both spins from the diagonal in Y, into buffer buf
    auto h_phase = [&](int buf) {
        float* h = reinterpret_cast<float*>(smem + kHB + buf * kHBStride);
        for (int e = tid; e < R2; e += kNT) {
            const int sg = e >= L ? 1 : 0, i = e - sg * L;
            float v = s.pdiag[i];
            if (s.coulomb_type == 1) {
                v += s.u_diag[i] * (yat((1 - sg) * L + i, i).x + 0.5f);
            } else if (general) {
                float a = 0.0f;
                for (int jj = 0; jj < L; ++jj)
                    a += s.U[i * L + jj] * (yat((1 - sg) * L + jj, jj).x + 0.5f) +
                         s.V[i * L + jj] * (yat(sg * L + jj, jj).x + 0.5f);
                v += a;
            }
            h[sg * 64 + i] = v;
        }
    };
    __syncthreads();
    pulse_table(gst);
    h_phase(0);
    __syncthreads();

    float* rr = nullptr;
    auto flush_energies = [&]() {
        if (tid == 0) {
            float e4[4] = {0, 0, 0, 0};
            for (int ww = 0; ww < W; ++ww)
                for (int qq = 0; qq < 4; ++qq) e4[qq] += ew[ww * 4 + qq];
            for (int qq = 0; qq < 4; ++qq) rr[R2 + qq] = e4[qq];
        }
    };

    // Stage ST reads h buffer ST & 1 and writes buffer (ST + 1) & 1: four stages per step, so every
    // step starts on buffer 0.
    auto stage = [&](auto stc, int64_t gidx, bool flush) {
        constexpr int ST = decltype(stc)::value;
        constexpr int HC = kHB + (ST & 1) * kHBStride;
        constexpr int HN = kHB + ((ST + 1) & 1) * kHBStride;
        // ---- phase A: X = T Y; lane = row i, column group (spin sg, columns w + W*mm) per step ----
        {
            C T[Z];
            if constexpr (PULSE) {
#pragma unroll
                for (int n = 0; n < Z; ++n) T[n] = Tc[n * L + il];
            }
            // software pipelined where the registers allow: the next column group's reads go out before
            // this group's X store (the compiler keeps reads behind an earlier store to the same array)
            constexpr int NCG = EPTA / 2;
            constexpr bool kPipe = Z == 4 ? (S <= 4 || !PULSE) : (!PULSE && S <= 3);
            C v[Z];
            if constexpr (kPipe) {
#pragma unroll
                for (int n = 0; n < Z; ++n) v[n] = lds_c(smem, kb0[n]);
            }
#pragma unroll
            for (int m = 0; m < EPTA; ++m) {
                const int sg = m / NCG;
                const int co = (m % NCG) * W * 8;  // immediate column offset
                float ax = 0.0f, ay = 0.0f;
#pragma unroll
                for (int n = 0; n < Z; ++n) {
                    const C vn = kPipe ? v[n] : lds_c(smem, (sg ? kb1[n] : kb0[n]) + co);
                    if constexpr (PULSE) {
                        ax += T[n].x * vn.x - T[n].y * vn.y;
                        ay += T[n].x * vn.y + T[n].y * vn.x;
                    } else {
                        ax += tv[n] * vn.x;
                        ay += tv[n] * vn.y;
                    }
                }
                if (kPipe && m + 1 < EPTA) {
                    const int sn = (m + 1) / NCG, cn = ((m + 1) % NCG) * W * 8;
#pragma unroll
                    for (int n = 0; n < Z; ++n) v[n] = lds_c(smem, (sn ? kb1[n] : kb0[n]) + cn);
                }
                *reinterpret_cast<C*>(smem + (sg ? xb1 : xb0) + co) = C{ax, ay};
            }
        }
        __syncthreads();
        if (ST == 0 && flush) flush_energies();
        // ---- phase B: every slot's reads first, so they are in flight together ----
        C xij[S], xji[S];
        float hi[S], hj[S];
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * W + w < n_off + n_dia) {
                xij[q] = lds_c(smem, aij[q] + kXOff);
                xji[q] = lds_c(smem, aji[q] + kXOff);
                int ahq = ah[q];  // opaque: decoded here, not kept decoded across the time loop
                asm volatile("" : "+v"(ahq));
                hi[q] = lds_f(smem, HC + (ahq & 0xffff));
                hj[q] = lds_f(smem, HC + (ahq >> 16));
            }
        }
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int er = q * W + w;
            if (er < n_off) {
                const float dh = hj[q] - hi[q];
                C yc;  // current stage input
                if constexpr (ST == 0) {
                    yc = y[q];
                } else if constexpr (ST == 1) {
                    yc = C{y[q].x + h3 * p[q].x, y[q].y + h3 * p[q].y};
                } else if constexpr (ST == 2) {
                    yc = C{y[q].x + h6 * (A[q].x - 3.0f * p[q].x), y[q].y + h6 * (A[q].y - 3.0f * p[q].y)};
                } else {
                    yc = C{y[q].x + h1 * p[q].x, y[q].y + h1 * p[q].y};
                }
                // c = conj(X_ji) - X_ij + dh * Y_ij;  k = -i c
                const float cx = xji[q].x - xij[q].x + dh * yc.x;
                const float cy = -xji[q].y - xij[q].y + dh * yc.y;
                const float kx = cy, ky = -cx;
                C yn;
                if constexpr (ST == 0) {
                    A[q] = C{kx, ky};
                    p[q] = C{kx, ky};
                    yn = C{y[q].x + h3 * kx, y[q].y + h3 * ky};
                } else if constexpr (ST == 1) {
                    A[q] = C{A[q].x + 3.0f * kx, A[q].y + 3.0f * ky};
                    p[q] = C{p[q].x - kx, p[q].y - ky};
                    yn = C{y[q].x + h6 * (A[q].x - 3.0f * p[q].x), y[q].y + h6 * (A[q].y - 3.0f * p[q].y)};
                } else if constexpr (ST == 2) {
                    A[q] = C{A[q].x + 3.0f * kx, A[q].y + 3.0f * ky};
                    p[q] = C{p[q].x + kx, p[q].y + ky};
                    yn = C{y[q].x + h1 * p[q].x, y[q].y + h1 * p[q].y};
                } else {
                    A[q] = C{A[q].x + kx, A[q].y + ky};
                    y[q] = C{y[q].x + h8 * A[q].x, y[q].y + h8 * A[q].y};
                    yn = y[q];
                }
                *reinterpret_cast<C*>(smem + aij[q]) = yn;
                *reinterpret_cast<C*>(smem + aji[q]) = C{yn.x, -yn.y};
            } else if (er < n_off + n_dia) {
                // diagonal: k = -2 Im X_ii, real RK registers in .x; A.y = u_ii, p.y = pdiag_i
                const float k = -2.0f * xij[q].y;
                float yn;
                if constexpr (ST == 0) {
                    A[q].x = k;
                    p[q].x = k;
                    yn = y[q].x + h3 * k;
                } else if constexpr (ST == 1) {
                    A[q].x += 3.0f * k;
                    p[q].x -= k;
                    yn = y[q].x + h6 * (A[q].x - 3.0f * p[q].x);
                } else if constexpr (ST == 2) {
                    A[q].x += 3.0f * k;
                    p[q].x += k;
                    yn = y[q].x + h1 * p[q].x;
                } else {
                    A[q].x += k;
                    y[q].x += h8 * A[q].x;
                    yn = y[q].x;
                }
                *reinterpret_cast<float*>(smem + aij[q]) = yn;
                *reinterpret_cast<float*>(smem + HN + ah[q]) = A[q].y * (yn + 0.5f) + p[q].y;
            }
        }
        pulse_table(gidx + 1);
        __syncthreads();
        if (general) {
            h_phase((ST + 1) & 1);
            __syncthreads();
        }
    };

    const int64_t Bn = r.B;
    const int R = r.R;
    for (int t = t0; t < t1; ++t) {
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b) * R;
        {
            // lane indices made opaque here, so the addresses of this section are formed here and not
            // kept in registers across the whole time loop
            int jt = j, tt = tid;
            asm volatile("" : "+v"(jt), "+v"(tt));
            for (int e = tt; e < R2; e += kNT) {
                const int sg = e >= L ? 1 : 0, i = e - sg * L;
                rr[e] = yat(sg * L + i, i).x;
            }
            if (tid == 0) {
                const C cu = yat(r.ci, r.cj), cd = yat(L + r.ci, r.cj);
                rr[R2 + 4] = cu.x;
                rr[R2 + 5] = cu.y;
                rr[R2 + 6] = cd.x;
                rr[R2 + 7] = cd.y;
            }
            if constexpr (DENS)
                dev::record_density(rr + R2 + 8, L, jt, w, kNT / 64, [&](int row, int col) { return yat(row, col).x; });
            // energies of this time point, lane = site i: wave 0 the hopping terms (neighbour addresses
            // from the phase-A tables: Y(k, i) = conj(Y(i, k))), wave 1 the on-site terms
            float eth = 0, eu = 0, ev = 0, epe = 0;
            if (w == 0) {
#pragma unroll
                for (int n = 0; n < Z; ++n) {
                    const C a = lds_c(smem, kb0[n] + jt * 8), c = lds_c(smem, kb1[n] + jt * 8);
                    eth += (a.x + c.x) * tv[n];
                    if constexpr (PULSE) {
                        const C P = Tc[n * L + min(jt, L - 1)];
                        if (jt < L) epe += (a.x + c.x) * (P.x - tv[n]) + (a.y + c.y) * P.y;
                    }
                }
            } else if (w == 1 && jt < L) {
                const int i = jt;
                const float nu = yat(i, i).x + 0.5f;
                const float nd = yat(L + i, i).x + 0.5f;
                eth = lds_f(smem, kSite + i * 4) * (nu + nd);
                epe = lds_f(smem, kSite + 256 + i * 4) * (nu + nd);
                if (s.coulomb_type == 1) {
                    eu = lds_f(smem, kSite + 512 + i * 4) * nu * nd;
                } else if (general) {
                    float su = 0, sd = 0, vu = 0, vd = 0;
                    for (int jj = 0; jj < L; ++jj) {
                        const float nuj = yat(jj, jj).x + 0.5f;
                        const float ndj = yat(L + jj, jj).x + 0.5f;
                        su += s.U[i * L + jj] * ndj;
                        sd += s.U[i * L + jj] * nuj;
                        vu += s.V[i * L + jj] * nuj;
                        vd += s.V[i * L + jj] * ndj;
                    }
                    eu = (nu * su + nd * sd) / 2;
                    ev = (nu * vu + nd * vd) / 2;
                }
            }
            eth = wave_sum(eth);
            eu = wave_sum(eu);
            ev = wave_sum(ev);
            epe = wave_sum(epe);
            if (j == 0) {
                ew[w * 4 + 0] = eth;
                ew[w * 4 + 1] = eu;
                ew[w * 4 + 2] = ev;
                ew[w * 4 + 3] = epe;
            }
        }
        if (t == r.T - 1) {
            __syncthreads();
            flush_energies();
            break;
        }
        for (int sub = 0; sub < r.nsub; ++sub) {
            stage(Ic<0>{}, gst, sub == 0);
            stage(Ic<1>{}, gst + 1, false);
            stage(Ic<2>{}, gst + 2, false);
            stage(Ic<3>{}, gst + 3, false);
            gst += 4;
        }
    }
    // ---- state back to HBM, from the RK registers ----
    C* go = reinterpret_cast<C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        if (aij[q] == spare) continue;
        const int er = q * W + w;
        const int row = (aij[q] - kRows) / kRowBytes;
        const int col = (aij[q] - kRows - row * kRowBytes) / 8;
        const int sg = row >= L ? 1 : 0, i = row - sg * L;
        if (er < n_off) {
            go[sg * LL + i * L + col] = y[q];
            go[sg * LL + col * L + i] = C{y[q].x, -y[q].y};
        } else {
            go[sg * LL + i * L + i] = C{y[q].x, 0.0f};
        }
    }
}
#undef smem

// Spin-pipelined variant (no pulse; null or on-site Coulomb).  Waves 0-7 carry spin up, waves 8-15
// spin down, and the down group runs half a stage behind, so every interval between two barriers pairs
// one spin's LDS-heavy phase A with the other spin's VALU-heavy phase B:
//   even interval of stage s:  up: A(s)      down: B(s - 1)
//   odd interval of stage s:   up: B(s)      down: A(s)
// B(s) of one spin reads that spin's h(s) and writes the other spin's h(s + 1) (buffers by stage
// parity), so each h is produced one interval before it is read and overwritten two intervals after.
// Records are taken after the even interval of a step's first stage, when both spins sit at the time
// point.  Per spin: phase A covers 64 columns with 8 waves (8 column groups each), phase B packs the
// spin's strict upper triangle and its diagonal into element rows owned by wave r % 8, slot r / 8.
template <int S, int Z>
__global__ void __launch_bounds__(1024)
    herm_pipe_kernel(DevSystem<float> s, RunShape r, int t0, int t1, float* __restrict__ state,
                     float* __restrict__ rec)
{
    constexpr int kNT = 1024;
    constexpr int kNCG = 8;  // column groups per wave in phase A
#define smem herm_static_lds
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int NS = L * (L - 1) / 2;
    const int n_off = (NS + 63) / 64;  // element rows per spin
    const int n_dia = (L + 63) / 64;
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int sg = w >> 3;  // this wave's spin
    const int gw = w & 7;   // wave within the spin group
    const int64_t b = blockIdx.x;
    const int spare = kRows + R2 * kRowBytes;
    float* ew = reinterpret_cast<float*>(smem + kEW);
    auto yat = [&](int row, int col) -> C& {
        return *reinterpret_cast<C*>(smem + kRows + row * kRowBytes + col * 8);
    };
    {
        const int n16 = (kRows + (R2 + 1) * kRowBytes) / 16;
        for (int e = tid; e < n16; e += kNT) reinterpret_cast<float4*>(smem)[e] = float4{0, 0, 0, 0};
    }
    __syncthreads();
    for (int i = tid; i < L; i += kNT) {
        *reinterpret_cast<float*>(smem + kSite + i * 4) = s.tdiag[i];
        *reinterpret_cast<float*>(smem + kSite + 256 + i * 4) = s.epot[i];
        *reinterpret_cast<float*>(smem + kSite + 512 + i * 4) = s.u_diag[i];
    }
    // ---- element slots of this wave's spin (layout as in herm_kernel) ----
    int aij[S], aji[S], ah[S];
    C y[S], A[S], p[S];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int er = q * 8 + gw;
        aij[q] = spare;
        aji[q] = spare;
        ah[q] = 128 * 4;
        y[q] = C{0, 0};
        A[q] = C{0, 0};
        p[q] = C{0, 0};
        if (er < n_off) {
            const int u = er * 64 + j;
            if (u < NS) {
                const float c = static_cast<float>(2 * L - 1);
                int i = static_cast<int>(0.5f * (c - sqrtf(fmaxf(c * c - 8.0f * static_cast<float>(u), 0.0f))));
                if (i < 0) i = 0;
                while (i > 0 && i * (2 * L - i - 1) / 2 > u) --i;
                while ((i + 1) * (2 * L - i - 2) / 2 <= u) ++i;
                const int jj = i + 1 + (u - i * (2 * L - i - 1) / 2);
                aij[q] = kRows + (sg * L + i) * kRowBytes + jj * 8;
                aji[q] = kRows + (sg * L + jj) * kRowBytes + i * 8;
                ah[q] = (sg * 64 + i) * 4 | ((sg * 64 + jj) * 4) << 16;
                y[q] = gs[sg * LL + i * L + jj];
                *reinterpret_cast<C*>(smem + aij[q]) = y[q];
                *reinterpret_cast<C*>(smem + aji[q]) = C{y[q].x, -y[q].y};
            }
        } else if (er < n_off + n_dia) {
            const int i = (er - n_off) * 64 + j;
            if (i < L) {
                aij[q] = kRows + (sg * L + i) * kRowBytes + i * 8;
                aji[q] = aij[q];
                ah[q] = ((1 - sg) * 64 + i) * 4;
                y[q] = C{gs[sg * LL + i * L + i].x, 0.0f};
                A[q].y = s.coulomb_type == 1 ? s.u_diag[i] : 0.0f;
                p[q].y = s.pdiag[i];
                *reinterpret_cast<C*>(smem + aij[q]) = y[q];
            }
        }
    }
    // ---- phase-A addresses for this wave's spin: lane j = row i, columns gw + 8 mm ----
    const int il = min(j, L - 1);
    int kb[Z];
    float tv[Z];
#pragma unroll
    for (int n = 0; n < Z; ++n) {
        const int k = j < L ? s.nbr_k[n * L + il] : 0;
        kb[n] = kRows + (sg * L + k) * kRowBytes + gw * 8;
        tv[n] = j < L ? s.nbr_t[n * L + il] : 0.0f;
    }
    const int xb = (j < L ? kRows + (sg * L + il) * kRowBytes : spare) + kXOff + gw * 8;
    const float h1 = static_cast<float>(r.h);
    const float h3 = static_cast<float>(r.h / 3.0), h6 = static_cast<float>(r.h / 6.0),
                h8 = static_cast<float>(r.h / 8.0);
    // h of both spins at the first stage, buffer 0, from the diagonal the slots above wrote
    __syncthreads();
    for (int e = tid; e < R2; e += kNT) {
        const int sgh = e >= L ? 1 : 0, i = e - sgh * L;
        float v = s.pdiag[i];
        if (s.coulomb_type == 1) v += s.u_diag[i] * (yat((1 - sgh) * L + i, i).x + 0.5f);
        *reinterpret_cast<float*>(smem + kHB + (sgh * 64 + i) * 4) = v;
    }
    __syncthreads();

    auto phase_a = [&]() {  // software pipelined as in herm_kernel where the registers allow
        constexpr bool kPipe = Z == 4 || S <= 2;
        C v[Z];
        if constexpr (kPipe) {
#pragma unroll
            for (int n = 0; n < Z; ++n) v[n] = lds_c(smem, kb[n]);
        }
#pragma unroll
        for (int mm = 0; mm < kNCG; ++mm) {
            const int co = mm * 64;  // 8 columns of 8 bytes
            float ax = 0.0f, ay = 0.0f;
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C vn = kPipe ? v[n] : lds_c(smem, kb[n] + co);
                ax += tv[n] * vn.x;
                ay += tv[n] * vn.y;
            }
            if (kPipe && mm + 1 < kNCG) {
#pragma unroll
                for (int n = 0; n < Z; ++n) v[n] = lds_c(smem, kb[n] + co + 64);
            }
            *reinterpret_cast<C*>(smem + xb + co) = C{ax, ay};
        }
    };
    auto phase_b = [&](auto stc) {
        constexpr int ST = decltype(stc)::value;
        constexpr int HC = kHB + (ST & 1) * kHBStride;
        constexpr int HN = kHB + ((ST + 1) & 1) * kHBStride;
        C xij[S], xji[S];
        float hi[S], hj[S];
#pragma unroll
        for (int q = 0; q < S; ++q) {
            if (q * 8 + gw < n_off + n_dia) {
                xij[q] = lds_c(smem, aij[q] + kXOff);
                xji[q] = lds_c(smem, aji[q] + kXOff);
                int ahq = ah[q];
                asm volatile("" : "+v"(ahq));
                hi[q] = lds_f(smem, HC + (ahq & 0xffff));
                hj[q] = lds_f(smem, HC + (ahq >> 16));
            }
        }
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int er = q * 8 + gw;
            if (er < n_off) {
                const float dh = hj[q] - hi[q];
                C yc;
                if constexpr (ST == 0) {
                    yc = y[q];
                } else if constexpr (ST == 1) {
                    yc = C{y[q].x + h3 * p[q].x, y[q].y + h3 * p[q].y};
                } else if constexpr (ST == 2) {
                    yc = C{y[q].x + h6 * (A[q].x - 3.0f * p[q].x), y[q].y + h6 * (A[q].y - 3.0f * p[q].y)};
                } else {
                    yc = C{y[q].x + h1 * p[q].x, y[q].y + h1 * p[q].y};
                }
                const float cx = xji[q].x - xij[q].x + dh * yc.x;
                const float cy = -xji[q].y - xij[q].y + dh * yc.y;
                const float kx = cy, ky = -cx;
                C yn;
                if constexpr (ST == 0) {
                    A[q] = C{kx, ky};
                    p[q] = C{kx, ky};
                    yn = C{y[q].x + h3 * kx, y[q].y + h3 * ky};
                } else if constexpr (ST == 1) {
                    A[q] = C{A[q].x + 3.0f * kx, A[q].y + 3.0f * ky};
                    p[q] = C{p[q].x - kx, p[q].y - ky};
                    yn = C{y[q].x + h6 * (A[q].x - 3.0f * p[q].x), y[q].y + h6 * (A[q].y - 3.0f * p[q].y)};
                } else if constexpr (ST == 2) {
                    A[q] = C{A[q].x + 3.0f * kx, A[q].y + 3.0f * ky};
                    p[q] = C{p[q].x + kx, p[q].y + ky};
                    yn = C{y[q].x + h1 * p[q].x, y[q].y + h1 * p[q].y};
                } else {
                    A[q] = C{A[q].x + kx, A[q].y + ky};
                    y[q] = C{y[q].x + h8 * A[q].x, y[q].y + h8 * A[q].y};
                    yn = y[q];
                }
                *reinterpret_cast<C*>(smem + aij[q]) = yn;
                *reinterpret_cast<C*>(smem + aji[q]) = C{yn.x, -yn.y};
            } else if (er < n_off + n_dia) {
                const float k = -2.0f * xij[q].y;
                float yn;
                if constexpr (ST == 0) {
                    A[q].x = k;
                    p[q].x = k;
                    yn = y[q].x + h3 * k;
                } else if constexpr (ST == 1) {
                    A[q].x += 3.0f * k;
                    p[q].x -= k;
                    yn = y[q].x + h6 * (A[q].x - 3.0f * p[q].x);
                } else if constexpr (ST == 2) {
                    A[q].x += 3.0f * k;
                    p[q].x += k;
                    yn = y[q].x + h1 * p[q].x;
                } else {
                    A[q].x += k;
                    y[q].x += h8 * A[q].x;
                    yn = y[q].x;
                }
                *reinterpret_cast<float*>(smem + aij[q]) = yn;
                *reinterpret_cast<float*>(smem + HN + ah[q]) = A[q].y * (yn + 0.5f) + p[q].y;
            }
        }
    };

    float* rr = nullptr;
    auto record = [&]() {
        int jt = j, tt = tid;
        asm volatile("" : "+v"(jt), "+v"(tt));
        for (int e = tt; e < R2; e += kNT) {
            const int sgr = e >= L ? 1 : 0, i = e - sgr * L;
            rr[e] = yat(sgr * L + i, i).x;
        }
        if (tid == 0) {
            const C cu = yat(r.ci, r.cj), cd = yat(L + r.ci, r.cj);
            rr[R2 + 4] = cu.x;
            rr[R2 + 5] = cu.y;
            rr[R2 + 6] = cd.x;
            rr[R2 + 7] = cd.y;
        }
        if (r.dens)
            dev::record_density(rr + R2 + 8, L, jt, w, kNT / 64, [&](int row, int col) { return yat(row, col).x; });
        float eth = 0, eu = 0, epe = 0;
        if (w == 0) {  // spin-up tables of wave 0 (column 0): Y(k, i) = conj(Y(i, k)); spin down L rows on
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C a = lds_c(smem, kb[n] + jt * 8), c = lds_c(smem, kb[n] + L * kRowBytes + jt * 8);
                eth += (a.x + c.x) * tv[n];
            }
        } else if (w == 1 && jt < L) {
            const int i = jt;
            const float nu = yat(i, i).x + 0.5f;
            const float nd = yat(L + i, i).x + 0.5f;
            eth = lds_f(smem, kSite + i * 4) * (nu + nd);
            epe = lds_f(smem, kSite + 256 + i * 4) * (nu + nd);
            if (s.coulomb_type == 1) eu = lds_f(smem, kSite + 512 + i * 4) * nu * nd;
        }
        eth = wave_sum(eth);
        eu = wave_sum(eu);
        epe = wave_sum(epe);
        if (j == 0) {
            ew[w * 4 + 0] = eth;
            ew[w * 4 + 1] = eu;
            ew[w * 4 + 2] = 0.0f;
            ew[w * 4 + 3] = epe;
        }
    };
    auto flush_energies = [&]() {
        if (tid == 0) {
            float e4[4] = {0, 0, 0, 0};
            for (int ww = 0; ww < 16; ++ww)
                for (int qq = 0; qq < 4; ++qq) e4[qq] += ew[ww * 4 + qq];
            for (int qq = 0; qq < 4; ++qq) rr[R2 + qq] = e4[qq];
        }
    };

    const int64_t Bn = r.B;
    const int R = r.R;
    bool pending = false;  // the down spin still owes phase B of the previous step's last stage
    for (int t = t0; t < t1; ++t) {
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b) * R;
        if (t == r.T - 1) {
            if (pending) {
                if (sg) phase_b(Ic<3>{});
                __syncthreads();
                pending = false;
            }
            record();
            __syncthreads();
            flush_energies();
            break;
        }
        for (int sub = 0; sub < r.nsub; ++sub) {
            // stage 0
            if (sg == 0) phase_a();
            else if (pending) phase_b(Ic<3>{});
            __syncthreads();
            if (sub == 0) {
                record();
                __syncthreads();
            }
            if (sg == 0) phase_b(Ic<0>{});
            else phase_a();
            __syncthreads();
            if (sub == 0) flush_energies();
            // stage 1
            if (sg == 0) phase_a();
            else phase_b(Ic<0>{});
            __syncthreads();
            if (sg == 0) phase_b(Ic<1>{});
            else phase_a();
            __syncthreads();
            // stage 2
            if (sg == 0) phase_a();
            else phase_b(Ic<1>{});
            __syncthreads();
            if (sg == 0) phase_b(Ic<2>{});
            else phase_a();
            __syncthreads();
            // stage 3
            if (sg == 0) phase_a();
            else phase_b(Ic<2>{});
            __syncthreads();
            if (sg == 0) phase_b(Ic<3>{});
            else phase_a();
            __syncthreads();
            pending = true;
        }
    }
    if (pending) {
        if (sg) phase_b(Ic<3>{});
        __syncthreads();
    }
    C* go = reinterpret_cast<C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        if (aij[q] == spare) continue;
        const int er = q * 8 + gw;
        const int row = (aij[q] - kRows) / kRowBytes;
        const int col = (aij[q] - kRows - row * kRowBytes) / 8;
        const int i = row - sg * L;
        if (er < n_off) {
            go[sg * LL + i * L + col] = y[q];
            go[sg * LL + col * L + i] = C{y[q].x, -y[q].y};
        } else {
            go[sg * LL + i * L + i] = C{y[q].x, 0.0f};
        }
    }
}
#undef smem

size_t herm_lds_bytes(int L, int z, bool pulse)
{
    return kRows + static_cast<size_t>(2 * L + 1) * kRowBytes + (pulse ? static_cast<size_t>(z) * L * sizeof(C) : 0);
}

// Shape for L sites: 16 waves; column groups per wave in phase A (EA) and element slots per thread (S).
struct HShape {
    int W, ea, s;
};
HShape herm_shape(int L)
{
    if (L > 64 || L < 2) return {0, 0, 0};
    const int W = 16;
    const int need_a = (2 * L + W - 1) / W;
    const int rows_b = (L * (L - 1) + 63) / 64 + (2 * L + 63) / 64;
    const int need_s = (rows_b + W - 1) / W;
    const int ea = need_a <= 2 ? 2 : (need_a <= 4 ? 4 : (need_a <= 8 ? 8 : 0));
    if (!ea || need_s > 5) return {0, 0, 0};
    // compiled (EA, S) pairs: (2, 1), (4, 2), (8, 3..5)
    int s = need_s;
    if (ea == 2) s = need_s <= 1 ? 1 : 0;
    else if (ea == 4) s = need_s <= 2 ? 2 : 0;
    else s = need_s <= 3 ? 3 : need_s;
    if (!s) return {0, 0, 0};
    return {W, ea, s};
}

template <int W, int EA, int S, int Z, bool PULSE, bool DENS>
hipError_t run(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec, hipStream_t st)
{
    auto kern = herm_kernel<W, EA, S, Z, PULSE, DENS>;
    if (herm_lds_bytes(s.L, s.z, PULSE) > static_cast<size_t>(kHermStaticLds)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(64 * W), 0, st, s, r, t0, t1, state, rec);
    return hipGetLastError();
}

template <int EA, int S, bool PULSE>
hipError_t run_z(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                 hipStream_t st)
{
    if (r.dens) {
        if (s.z == 4) return run<16, EA, S, 4, PULSE, true>(s, r, t0, t1, state, rec, st);
        if (s.z == 8) return run<16, EA, S, 8, PULSE, true>(s, r, t0, t1, state, rec, st);
    } else {
        if (s.z == 4) return run<16, EA, S, 4, PULSE, false>(s, r, t0, t1, state, rec, st);
        if (s.z == 8) return run<16, EA, S, 8, PULSE, false>(s, r, t0, t1, state, rec, st);
    }
    return hipErrorInvalidValue;
}

template <bool PULSE>
hipError_t run_shape(const HShape& h, const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state,
                     float* rec, hipStream_t st)
{
    if (h.ea == 2 && h.s == 1) return run_z<2, 1, PULSE>(s, r, t0, t1, state, rec, st);
    if (h.ea == 4 && h.s == 2) return run_z<4, 2, PULSE>(s, r, t0, t1, state, rec, st);
    if (h.ea == 8 && h.s == 3) return run_z<8, 3, PULSE>(s, r, t0, t1, state, rec, st);
    if (h.ea == 8 && h.s == 4) return run_z<8, 4, PULSE>(s, r, t0, t1, state, rec, st);
    if (h.ea == 8 && h.s == 5) return run_z<8, 5, PULSE>(s, r, t0, t1, state, rec, st);
    return hipErrorInvalidValue;
}

int pipe_slots(int L)
{
    if (L < 2 || L > 64) return 0;
    const int rows = (L * (L - 1) / 2 + 63) / 64 + (L + 63) / 64;
    return (rows + 7) / 8;
}

template <int S, int Z>
hipError_t run_pipe(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                    hipStream_t st)
{
    auto kern = herm_pipe_kernel<S, Z>;
    if (herm_lds_bytes(s.L, s.z, false) > static_cast<size_t>(kHermStaticLds)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(1024), 0, st, s, r, t0, t1, state, rec);
    return hipGetLastError();
}

template <int Z>
hipError_t run_pipe_s(int S, const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                      hipStream_t st)
{
    switch (S) {
        case 1: return run_pipe<1, Z>(s, r, t0, t1, state, rec, st);
        case 2: return run_pipe<2, Z>(s, r, t0, t1, state, rec, st);
        case 3: return run_pipe<3, Z>(s, r, t0, t1, state, rec, st);
        case 4: return run_pipe<4, Z>(s, r, t0, t1, state, rec, st);
        case 5: return run_pipe<5, Z>(s, r, t0, t1, state, rec, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool herm_pipe_fits(const DevSystem<float>& s)
{
    return !s.amp && s.coulomb_type != 2 && (s.z == 4 || s.z == 8) && pipe_slots(s.L) >= 1 && pipe_slots(s.L) <= 5;
}

hipError_t launch_herm_pipe(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                            hipStream_t st)
{
    if (!herm_pipe_fits(s) || r.B <= 0) return hipErrorInvalidValue;
    const int S = pipe_slots(s.L);
    if (s.z == 4) return run_pipe_s<4>(S, s, r, t0, t1, state, rec, st);
    return run_pipe_s<8>(S, s, r, t0, t1, state, rec, st);
}

bool herm_fits(int L, int z)
{
    const HShape h = herm_shape(L);
    const int zp = stencil_padded_z(z);
    return h.W && zp && herm_lds_bytes(L, zp, true) <= kLdsLimit;
}

hipError_t launch_herm(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                       int nt_hint, hipStream_t st)
{
    (void)nt_hint;  // one workgroup shape (16 waves) for every L
    const HShape h = herm_shape(s.L);
    if (!h.W || (s.z != 4 && s.z != 8) || r.B <= 0) return hipErrorInvalidValue;
    if (s.amp) return run_shape<true>(h, s, r, t0, t1, state, rec, st);
    return run_shape<false>(h, s, r, t0, t1, state, rec, st);
}

}  // namespace fpsamd
