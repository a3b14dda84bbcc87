// ftwa_herm.hip -- FP32 batched fTWA integrator exploiting the Hermiticity of rho.
//
// Same physics and record format as ftwa_stencil.hip (src/FHfTWASolver.cpp:366-426, :568-600, :629-661),
// organised around X = T rho, the hopping applied from the left:
//
//   rho T - T rho = X^dagger - X      (rho and T Hermitian: (rho T)_ij = conj(X_ji))
//   k_ij = -i [ conj(X_ji) - X_ij + (h_j - h_i) rho_ij ]
//
// Per RK stage and trajectory (one workgroup):
//   phase A  wavefront w forms full rows of X, lane j = column j:  X_ij = sum_{k in nb(i)} t_ik Y_kj.
//            The row index is wave-uniform, so nb(i) and t_ik are one scalar load; every read of Y is a
//            contiguous row segment.  X goes to LDS.                                      -> barrier
//   phase B  each thread owns upper-triangle elements (i <= j) of both spins, balanced over the
//            workgroup, with their RK4 "3/8" registers (y, A, p).  It reads X_ij and X_ji, forms k,
//            updates (y, A, p), and writes the next stage input Y_ij and its mirror Y_ji = conj(Y_ij).
//            The owners of diagonal elements write the other spin's next Hartree potential into the
//            other half of a double-buffered h.                                           -> barrier
// Only the upper triangle carries RK state: half the registers and half the update arithmetic of the
// full-matrix kernel, and the state stays exactly Hermitian (real diagonal).
// LDS rows are padded to 65 complex so the transposed accesses of phase B hit distinct banks.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::wave_sum;
using C = float2;

constexpr int kLP = 65;  // LDS row stride in complex elements
constexpr size_t kLdsLimit = 160 * 1024;

template <int N>
using Ic = std::integral_constant<int, N>;

__device__ __forceinline__ C cmul(C a, C b) { return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// EPTA: rows per wavefront in phase A; EPTB: upper-triangle elements per thread in phase B.
template <int W, int EPTA, int EPTB, int Z, bool PULSE>
__global__ void __launch_bounds__(64 * W)
    herm_kernel(DevSystem<float> s, RunShape r, int t0, int t1, float* __restrict__ state, float* __restrict__ rec,
                const RowNbrs<float, Z>* __restrict__ rows)
{
    constexpr int kNT = 64 * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int NU = L * (L + 1) / 2;  // upper-triangle elements per spin
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t b = blockIdx.x;
    const bool general = s.coulomb_type == 2;

    C* Y = reinterpret_cast<C*>(smem);                // [2L][kLP] stage input, both triangles
    C* X = Y + R2 * kLP;                              // [2L][kLP] X = T Y
    C* Tc = X + R2 * kLP;                             // [Z][L] pulse hopping t (1 + e^{i phi})
    float* hb = reinterpret_cast<float*>(PULSE ? (Tc + Z * L) : Tc);  // [2 buffers][2 spins][64]
    float* ew = hb + 4 * 64;                          // [W][4]

    // ---- phase-B ownership: element e = tid + kNT*q, spin-major, row-major upper triangle ----
    int own[EPTB];  // packed (spin << 16) | (i << 8) | j, or -1
#pragma unroll
    for (int q = 0; q < EPTB; ++q) {
        const int e = tid + kNT * q;
        if (e < 2 * NU) {
            const int sg = e >= NU ? 1 : 0;
            int u = e - sg * NU;
            // row i: the largest i with i*L - i*(i-1)/2 <= u
            const float Lf = static_cast<float>(L) + 0.5f;
            int i = static_cast<int>(Lf - sqrtf(Lf * Lf - 2.0f * static_cast<float>(u)));
            if (i < 0) i = 0;
            while (i > 0 && i * L - i * (i - 1) / 2 > u) --i;
            while ((i + 1) * L - (i + 1) * i / 2 <= u) ++i;
            const int jj = i + (u - (i * L - i * (i - 1) / 2));
            own[q] = (sg << 16) | (i << 8) | jj;
        } else {
            own[q] = -1;
        }
    }

    for (int e = tid; e < 4 * 64; e += kNT) hb[e] = 0.0f;
    C y[EPTB], A[EPTB], p[EPTB];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < EPTB; ++q) {
        y[q] = C{0, 0};
        A[q] = C{0, 0};
        p[q] = C{0, 0};
        if (own[q] >= 0) {
            const int sg = own[q] >> 16, i = (own[q] >> 8) & 255, jj = own[q] & 255;
            y[q] = gs[sg * LL + i * L + jj];
            if (i == jj) y[q].y = 0.0f;
            Y[(sg * L + i) * kLP + jj] = y[q];
            Y[(sg * L + jj) * kLP + i] = C{y[q].x, -y[q].y};
        }
    }
    const float h1 = static_cast<float>(r.h);
    const float h3 = static_cast<float>(r.h / 3.0), h6 = static_cast<float>(r.h / 6.0),
                h8 = static_cast<float>(r.h / 8.0);
    int64_t gst = static_cast<int64_t>(t0) * r.nsub * 4;
    int hbuf = 0;  // which half of hb holds the current stage's Hartree potential

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
    auto h_phase = [&](int buf) {
        float* h = hb + buf * 128;
        for (int e = tid; e < R2; e += kNT) {
            const int sg = e >= L ? 1 : 0, i = e - sg * L;
            float v = s.pdiag[i];
            if (s.coulomb_type == 1) {
                v += s.u_diag[i] * (Y[((1 - sg) * L + i) * kLP + i].x + 0.5f);
            } else if (general) {
                float a = 0.0f;
                for (int jj = 0; jj < L; ++jj)
                    a += s.U[i * L + jj] * (Y[((1 - sg) * L + jj) * kLP + jj].x + 0.5f) +
                         s.V[i * L + jj] * (Y[(sg * L + jj) * kLP + jj].x + 0.5f);
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

    auto stage = [&](auto stc, int64_t gidx, bool flush) {
        constexpr int ST = decltype(stc)::value;
        // ---- phase A: X = T Y, full rows ----
        {
            int wo = w;
            int jo = j;
            asm volatile("" : "+s"(wo), "+v"(jo));
#pragma unroll
            for (int m = 0; m < EPTA; ++m) {
                const int row = wo + W * m;
                if (row >= R2) break;
                const int sg = row >= L ? 1 : 0;
                const int i = row - sg * L;
                const RowNbrs<float, Z> nb = rows[i];
                const C* Ys = Y + sg * L * kLP;
                float ax = 0.0f, ay = 0.0f;
#pragma unroll
                for (int n = 0; n < Z; ++n) {
                    const C v = Ys[nb.k[n] * kLP + jo];
                    if constexpr (PULSE) {
                        const C T = Tc[n * L + i];
                        ax += T.x * v.x - T.y * v.y;
                        ay += T.x * v.y + T.y * v.x;
                    } else {
                        ax += nb.t[n] * v.x;
                        ay += nb.t[n] * v.y;
                    }
                }
                X[row * kLP + jo] = C{ax, ay};
            }
        }
        __syncthreads();
        if (ST == 0 && flush) flush_energies();
        // ---- phase B: upper-triangle elements ----
        const float* hc = hb + hbuf * 128;
        float* hn = hb + (hbuf ^ 1) * 128;
        // element indices are made opaque per stage: their LDS addresses are formed here, not hoisted
        // out of the time loop into registers for every element
        int ow[EPTB];
#pragma unroll
        for (int q = 0; q < EPTB; ++q) {
            ow[q] = own[q];
            asm volatile("" : "+v"(ow[q]));
        }
#pragma unroll
        for (int q = 0; q < EPTB; ++q) {
            if (ow[q] < 0) continue;
            const int sg = ow[q] >> 16, i = (ow[q] >> 8) & 255, jj = ow[q] & 255;
            const C xij = X[(sg * L + i) * kLP + jj];
            const C xji = X[(sg * L + jj) * kLP + i];
            const float dh = hc[sg * 64 + jj] - hc[sg * 64 + i];
            // current stage input from the registers
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
            // c = conj(X_ji) - X_ij + dh * Y_ij;  k = -i c
            const float cx = xji.x - xij.x + dh * yc.x;
            const float cy = -xji.y - xij.y + dh * yc.y;
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
            if (i == jj) {
                yn.y = 0.0f;
                Y[(sg * L + i) * kLP + i] = yn;
                const float uc = (s.coulomb_type == 1) ? s.u_diag[i] : 0.0f;
                hn[(1 - sg) * 64 + i] = uc * (yn.x + 0.5f) + s.pdiag[i];
            } else {
                Y[(sg * L + i) * kLP + jj] = yn;
                Y[(sg * L + jj) * kLP + i] = C{yn.x, -yn.y};
            }
        }
        pulse_table(gidx + 1);
        hbuf ^= 1;
        __syncthreads();
        if (general) {
            h_phase(hbuf);
            __syncthreads();
        }
    };

    const int64_t Bn = r.B;
    const int R = r.R;
    for (int t = t0; t < t1; ++t) {
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b) * R;
        {
            for (int e = tid; e < R2; e += kNT) {
                const int sg = e >= L ? 1 : 0, i = e - sg * L;
                rr[e] = Y[(sg * L + i) * kLP + i].x;
            }
            if (tid == 0) {
                const C cu = Y[L - 1], cd = Y[L * kLP + L - 1];
                rr[R2 + 4] = cu.x;
                rr[R2 + 5] = cu.y;
                rr[R2 + 6] = cd.x;
                rr[R2 + 7] = cd.y;
            }
            float eth = 0, eu = 0, ev = 0, epe = 0;
            if (tid < L) {
                const int i = tid;
                const float nu = Y[i * kLP + i].x + 0.5f;
                const float nd = Y[(L + i) * kLP + i].x + 0.5f;
                for (int n = 0; n < Z; ++n) {
                    const int k = s.nbr_k[n * L + i];
                    const float tt = s.nbr_t[n * L + i];
                    const C a = Y[i * kLP + k], c = Y[(L + i) * kLP + k];
                    eth += (a.x + c.x) * tt;
                    if constexpr (PULSE) {
                        const C P = Tc[n * L + i];
                        epe += (a.x + c.x) * (P.x - tt) - (a.y + c.y) * P.y;
                    }
                }
                eth += s.tdiag[i] * (nu + nd);
                epe += s.pot[i] * (nu + nd);
                if (s.coulomb_type == 1) {
                    eu = s.u_diag[i] * nu * nd;
                } else if (general) {
                    float su = 0, sd = 0, vu = 0, vd = 0;
                    for (int jj = 0; jj < L; ++jj) {
                        const float nuj = Y[jj * kLP + jj].x + 0.5f;
                        const float ndj = Y[(L + jj) * kLP + jj].x + 0.5f;
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
    C* go = reinterpret_cast<C*>(state) + b * 2 * LL;
#pragma unroll
    for (int q = 0; q < EPTB; ++q) {
        if (own[q] < 0) continue;
        const int sg = own[q] >> 16, i = (own[q] >> 8) & 255, jj = own[q] & 255;
        go[sg * LL + i * L + jj] = y[q];
        if (i != jj) go[sg * LL + jj * L + i] = C{y[q].x, -y[q].y};
    }
}

size_t herm_lds_bytes(int L, int z, bool pulse, int W)
{
    return 2 * static_cast<size_t>(2 * L) * kLP * sizeof(C) + (pulse ? static_cast<size_t>(z) * L * sizeof(C) : 0) +
           4 * 64 * sizeof(float) + static_cast<size_t>(W) * 4 * sizeof(float);
}

// Shape (waves W, rows per wave in phase A, upper elements per thread in phase B) for L sites.
// 16 waves for any L <= 64; 13 waves (exactly 5 upper elements per thread at L = 64) and 8 waves are
// alternatives for the autotuner at L = 64 (block_threads 832 or 512).
struct HShape {
    int W, ea, eb;
};
HShape herm_shape(int L, int nt_hint)
{
    if (L > 64 || L < 1) return {0, 0, 0};
    if (nt_hint == 832 && L * (L + 1) <= 832 * 5 && 2 * L <= 13 * 10) return {13, 10, 5};
    if (nt_hint == 512 && L * (L + 1) <= 512 * 9 && 2 * L <= 8 * 16) return {8, 16, 9};
    const int need_a = (2 * L + 15) / 16;
    const int need_b = (L * (L + 1) + 1023) / 1024;
    const int ea = need_a <= 2 ? 2 : (need_a <= 4 ? 4 : (need_a <= 8 ? 8 : 0));
    const int eb = need_b <= 1 ? 1 : (need_b <= 2 ? 2 : (need_b <= 3 ? 3 : (need_b <= 5 ? 5 : 0)));
    if (!ea || !eb) return {0, 0, 0};
    return {16, ea, eb};
}

template <int W, int EA, int EB, int Z, bool PULSE>
hipError_t run(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec, hipStream_t st)
{
    auto kern = herm_kernel<W, EA, EB, Z, PULSE>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(64 * W), herm_lds_bytes(s.L, s.z, PULSE, W), st,
                       s, r, t0, t1, state, rec, static_cast<const RowNbrs<float, Z>*>(s.rows));
    return hipGetLastError();
}

template <int W, int EA, int EB, bool PULSE>
hipError_t run_z(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                 hipStream_t st)
{
    if (s.z == 4) return run<W, EA, EB, 4, PULSE>(s, r, t0, t1, state, rec, st);
    if (s.z == 8) return run<W, EA, EB, 8, PULSE>(s, r, t0, t1, state, rec, st);
    return hipErrorInvalidValue;
}

template <bool PULSE>
hipError_t run_shape(const HShape& h, const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state,
                     float* rec, hipStream_t st)
{
    if (h.W == 16) {
        if (h.ea == 2 && h.eb == 1) return run_z<16, 2, 1, PULSE>(s, r, t0, t1, state, rec, st);
        if (h.ea == 4 && h.eb == 1) return run_z<16, 4, 1, PULSE>(s, r, t0, t1, state, rec, st);
        if (h.ea == 4 && h.eb == 2) return run_z<16, 4, 2, PULSE>(s, r, t0, t1, state, rec, st);
        if (h.ea == 8 && h.eb == 2) return run_z<16, 8, 2, PULSE>(s, r, t0, t1, state, rec, st);
        if (h.ea == 8 && h.eb == 3) return run_z<16, 8, 3, PULSE>(s, r, t0, t1, state, rec, st);
        if (h.ea == 8 && h.eb == 5) return run_z<16, 8, 5, PULSE>(s, r, t0, t1, state, rec, st);
    }
    if (!PULSE && s.z == 4) {
        if (h.W == 13) return run<13, 10, 5, 4, false>(s, r, t0, t1, state, rec, st);
        if (h.W == 8) return run<8, 16, 9, 4, false>(s, r, t0, t1, state, rec, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

bool herm_fits(int L, int z)
{
    const HShape h = herm_shape(L, 0);
    const int zp = stencil_padded_z(z);
    return h.W && zp && herm_lds_bytes(L, zp, true, h.W) <= kLdsLimit;
}

hipError_t launch_herm(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                       int nt_hint, hipStream_t st)
{
    HShape h = herm_shape(s.L, nt_hint);
    if ((h.W == 13 || h.W == 8) && (s.amp || s.z != 4)) h = herm_shape(s.L, 0);
    if (!h.W || (s.z != 4 && s.z != 8) || r.B <= 0) return hipErrorInvalidValue;
    if (s.amp) return run_shape<true>(h, s, r, t0, t1, state, rec, st);
    return run_shape<false>(h, s, r, t0, t1, state, rec, st);
}

}  // namespace fpsamd
