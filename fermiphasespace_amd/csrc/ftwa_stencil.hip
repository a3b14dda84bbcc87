// ftwa_stencil.hip -- the batched fTWA integrator (the product kernel).
//
// Restates FHfTWASolver's trajectory loop body (src/FHfTWASolver.cpp:366-426) with derivative
// (:568-600) and step_RK4_38 (:629-661) for many trajectories at once:
//
//   * one workgroup integrates one trajectory over a chunk of time points, both spins' RK stage input
//     resident in LDS as a [2L][64] complex array (rows stacked spin-up then spin-down, each row padded
//     to one 64-lane wavefront; padding stays exactly zero);
//   * wavefront w owns the rows w, w+W, w+2W, ... and lane j owns column j, so the row index is
//     wave-uniform: neighbour lists of the row come through scalar loads, neighbour lists of the column
//     sit in registers, and the addresses of the row's own elements are immediate offsets;
//   * the commutator -i[rho, Delta] with Delta = tM + diag(h) is a sparse stencil over the hopping's
//     neighbour lists:  k_ij = -i [ sum_{k in nb(j)} Y_ik t_kj - sum_{k in nb(i)} t_ik Y_kj + (h_j - h_i) Y_ij ];
//   * the RK4 "3/8" state per element is (y, A, p) in registers: A accumulates b-weighted stages and
//     p = k1, k1 - k2, k1 - k2 + k3, so every stage input follows from (y, A, p) alone:
//        Y2 = y + h p/3,   Y3 = y + h (A - 3p)/6,   Y4 = y + h p,   y' = y + h A/8;
//   * the four stages are unrolled at compile time; each is compute (read LDS) -> barrier -> write the
//     next stage input (and, for an on-site U, the owners of diagonal elements refresh the other spin's
//     Hartree potential) -> barrier;
//   * every time point writes one record per trajectory (diagonals, energies, corner element) that the
//     reduction kernels sum in fixed order.
// Neighbour lists are padded on the host to exactly Z entries (padding: the site itself, zero hopping).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

#ifndef FTWA_REAL
#define FTWA_REAL 2
#endif

namespace fpsamd {

namespace {

using dev::V2;
using dev::wave_sum;

constexpr int kRow = 64;  // LDS row length: sites padded to one wavefront
constexpr size_t kLdsLimit = 160 * 1024;

template <int N>
using Ic = std::integral_constant<int, N>;

// Value of v held by lane `lane` (wave-uniform index), broadcast to the whole wavefront.
[[maybe_unused]] __device__ __forceinline__ float lane_value(float v, int lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
[[maybe_unused]] __device__ __forceinline__ double lane_value(double v, int lane)
{
    const long long bits = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// LT > 0 fixes the site count at compile time (all index arithmetic folds); LT = 0 reads it from s.L.
template <typename Real, int W, int EPT, int Z, bool PULSE, int LT>
__global__ void __launch_bounds__(64 * W)
    stencil_kernel(DevSystem<Real> s, RunShape r, int t0, int t1, Real* __restrict__ state, Real* __restrict__ rec,
                   const RowNbrs<Real, Z>* __restrict__ rows)
{
    using C = V2<Real>;
    constexpr int NT = 64 * W;
    // Folding the spin of each element into the code (when rows exactly tile the wavefronts) looks
    // attractive but makes the register allocator spill heavily; the spin stays a uniform runtime value.
    constexpr bool kExact = false;
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = LT > 0 ? LT : s.L;
    const int LL = L * L;
    const int R2 = 2 * L;
    const int tid = threadIdx.x;
    const int j = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    __builtin_assume(w >= 0 && w < W);
    const bool jon = (LT == kRow) ? true : (j < L);
    const int64_t b = blockIdx.x;
    const bool general = s.coulomb_type == 2;
    const int* __restrict__ nbk = s.nbr_k;
    const Real* __restrict__ nbt = s.nbr_t;
    const Real* __restrict__ udiag = s.u_diag;
    const Real* __restrict__ pdiag = s.pdiag;

    C* Y = reinterpret_cast<C*>(smem);                                  // [2L][64]
    C* Tc = Y + R2 * kRow;                                              // [Z][L] (pulse)
    Real* hb = reinterpret_cast<Real*>(PULSE ? (Tc + Z * L) : Tc);      // [2][64]
    Real* ew = hb + 2 * kRow;                                           // [W][4]

    for (int e = tid; e < R2 * kRow; e += NT) Y[e] = C{0, 0};
    for (int e = tid; e < 2 * kRow; e += NT) hb[e] = Real(0);

    int ck[Z];
    Real ct[Z];
#pragma unroll
    for (int n = 0; n < Z; ++n) {
        ck[n] = jon ? nbk[n * L + j] : j;
        ct[n] = jon ? nbt[n * L + j] : Real(0);
    }
    // lane j keeps U_jj and the diagonal potential of site j: the Hartree update of row i reads lane i
    const Real uj = (jon && s.coulomb_type == 1) ? udiag[j] : Real(0);
    const Real pj = jon ? pdiag[j] : Real(0);
    __syncthreads();

    C y[EPT], A[EPT], p[EPT];
    const C* gs = reinterpret_cast<const C*>(state) + b * 2 * LL;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int row = w + W * m;
        y[m] = C{0, 0};
        A[m] = C{0, 0};
        p[m] = C{0, 0};
        if ((kExact || row < R2) && jon) {
            const int sg = kExact ? (m >= EPT / 2 ? 1 : 0) : (row >= L ? 1 : 0);
            const int i = row - sg * L;
            y[m] = gs[sg * LL + i * L + j];
            Y[row * kRow + j] = y[m];
        }
    }
    const Real h1 = static_cast<Real>(r.h);
    const Real h3 = static_cast<Real>(r.h / 3.0), h6 = static_cast<Real>(r.h / 6.0), h8 = static_cast<Real>(r.h / 8.0);
    int64_t gst = static_cast<int64_t>(t0) * r.nsub * 4;  // global RK-stage index (pulse amplitude table)

    auto pulse_table = [&](int64_t stage_index) {
        if constexpr (PULSE) {
            const Real a = static_cast<Real>(s.amp[stage_index]);
            for (int e = tid; e < Z * L; e += NT) {
                Real sn, cs;
                if constexpr (std::is_same<Real, float>::value)
                    __sincosf(a * s.nbr_pd[e], &sn, &cs);
                else
                    sincos(a * s.nbr_pd[e], &sn, &cs);
                const Real t = nbt[e];
                Tc[e] = C{t * (Real(1) + cs), t * sn};
            }
        }
    };
    // Hartree potential h for both spins from the LDS diagonals (src/FHfTWASolver.cpp:574-583)
    auto h_phase = [&]() {
        for (int e = tid; e < R2; e += NT) {
            const int sg = e >= L ? 1 : 0, i = e - sg * L;
            Real v = pdiag[i];
            if (s.coulomb_type == 1) {
                v += udiag[i] * (Y[((1 - sg) * L + i) * kRow + i].x + Real(0.5));
            } else if (general) {
                Real a = 0;
                for (int jj = 0; jj < L; ++jj)
                    a += s.U[i * L + jj] * (Y[((1 - sg) * L + jj) * kRow + jj].x + Real(0.5)) +
                         s.V[i * L + jj] * (Y[(sg * L + jj) * kRow + jj].x + Real(0.5));
                v += a;
            }
            hb[sg * kRow + i] = v;
        }
    };

    __syncthreads();
    pulse_table(gst);
    h_phase();
    __syncthreads();

    Real* rr = nullptr;
    auto flush_energies = [&]() {
        if (tid == 0) {
            Real e4[4] = {0, 0, 0, 0};
            for (int ww = 0; ww < W; ++ww)
                for (int q = 0; q < 4; ++q) e4[q] += ew[ww * 4 + q];
            for (int q = 0; q < 4; ++q) rr[R2 + q] = e4[q];
        }
    };

    // One RK stage: compute k for every owned element, fold it into (A, p, y), then publish the next
    // stage input.  ST is the stage (0..3); flush writes the record's energies after the first barrier.
    auto stage = [&](auto stc, int64_t gidx, bool flush) {
        constexpr int ST = decltype(stc)::value;
        // the row records are read per element through the scalar cache; the row base is made opaque
        // per stage so the compiler does not pin every row's record in registers for the whole run.
        // Likewise the lane's column and its neighbour columns, so LDS addresses are formed per stage
        // instead of being hoisted, one register each per (element, neighbour), out of the time loop
        int jo = j;
        int cko[Z];
        int wo = w;
        asm volatile("" : "+v"(jo), "+s"(wo));
#pragma unroll
        for (int n = 0; n < Z; ++n) {
            cko[n] = ck[n];
            asm volatile("" : "+v"(cko[n]));
        }
        const Real hj0 = hb[j], hj1 = hb[kRow + j];
        C cT[Z];
        if constexpr (PULSE) {
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C v = jon ? Tc[n * L + j] : C{0, 0};
                cT[n] = C{v.x, -v.y};  // t_kj(t) = conj(t_jk(t))
            }
        }
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int row = wo + W * m;
            if (!kExact && row >= R2) continue;
            const int sg = kExact ? (m >= EPT / 2 ? 1 : 0) : (row >= L ? 1 : 0);
            const int i = row - sg * L;
            const RowNbrs<Real, Z> nb = rows[i];
            const C* Yr = Y + row * kRow;
            const C* Ys = Y + sg * L * kRow;
            const C yij = (ST == 0) ? y[m] : Yr[jo];
            const Real hjs = sg ? hj1 : hj0;
            const Real dh = hjs - lane_value(hjs, i);  // h_j - h_i, h_i from lane i
            Real ax = dh * yij.x, ay = dh * yij.y;
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C v = Yr[cko[n]];
                if constexpr (PULSE) {
                    ax += v.x * cT[n].x - v.y * cT[n].y;
                    ay += v.x * cT[n].y + v.y * cT[n].x;
                } else {
                    ax += v.x * ct[n];
                    ay += v.y * ct[n];
                }
            }
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const C v = Ys[nb.k[n] * kRow + jo];
                if constexpr (PULSE) {
                    const C T = Tc[n * L + i];
                    ax -= T.x * v.x - T.y * v.y;
                    ay -= T.x * v.y + T.y * v.x;
                } else {
                    const Real tt = nb.t[n];
                    ax -= tt * v.x;
                    ay -= tt * v.y;
                }
            }
            const Real kx = ay, ky = -ax;  // k = -i (ax + i ay)
            if constexpr (ST == 0) {
                A[m] = C{kx, ky};
                p[m] = C{kx, ky};
            } else if constexpr (ST == 1) {
                A[m] = C{A[m].x + Real(3) * kx, A[m].y + Real(3) * ky};
                p[m] = C{p[m].x - kx, p[m].y - ky};
            } else if constexpr (ST == 2) {
                A[m] = C{A[m].x + Real(3) * kx, A[m].y + Real(3) * ky};
                p[m] = C{p[m].x + kx, p[m].y + ky};
            } else {
                A[m] = C{A[m].x + kx, A[m].y + ky};
                y[m] = C{y[m].x + h8 * A[m].x, y[m].y + h8 * A[m].y};
            }
            // keep each element's loads and arithmetic together: unbounded interleaving of the
            // unrolled elements makes the scheduler hoist loads until registers spill
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        if (ST == 0 && flush) flush_energies();
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int row = w + W * m;
            if (!kExact && row >= R2) continue;
            C yn;
            if constexpr (ST == 0) {
                yn = C{y[m].x + h3 * p[m].x, y[m].y + h3 * p[m].y};
            } else if constexpr (ST == 1) {
                yn = C{y[m].x + h6 * (A[m].x - Real(3) * p[m].x), y[m].y + h6 * (A[m].y - Real(3) * p[m].y)};
            } else if constexpr (ST == 2) {
                yn = C{y[m].x + h1 * p[m].x, y[m].y + h1 * p[m].y};
            } else {
                yn = y[m];
            }
            Y[row * kRow + jo] = yn;
            {
                // on-site Hartree potential of the other spin from this row's diagonal element, which
                // lane i holds: fetched wave-wide and stored by every lane (same address, same value),
                // so there is no divergent branch.  With a general U/V, h_phase overwrites it below.
                const int sg = kExact ? (m >= EPT / 2 ? 1 : 0) : (row >= L ? 1 : 0);
                const int i = row - sg * L;
                hb[(1 - sg) * kRow + i] = lane_value(uj * (yn.x + Real(0.5)) + pj, i);
            }
        }
        pulse_table(gidx + 1);
        __syncthreads();
        if (general) {
            h_phase();
            __syncthreads();
        }
    };

    const int64_t Bn = r.B;
    const int R = r.R;
    for (int t = t0; t < t1; ++t) {
        rr = rec + (static_cast<int64_t>(t - t0) * Bn + b) * R;
        // ---- record time point t (src/FHfTWASolver.cpp:366-418) ----
        {
            for (int e = tid; e < R2; e += NT) {
                const int sg = e >= L ? 1 : 0, i = e - sg * L;
                rr[e] = Y[(sg * L + i) * kRow + i].x;
            }
            if (tid == 0) {
                const C cu = Y[r.ci * kRow + r.cj], cd = Y[(L + r.ci) * kRow + r.cj];
                rr[R2 + 4] = cu.x;
                rr[R2 + 5] = cu.y;
                rr[R2 + 6] = cd.x;
                rr[R2 + 7] = cd.y;
            }
            if (r.dens)
                dev::record_density(rr + R2 + 8, L, j, w, NT / 64, [&](int row, int col) { return Y[row * kRow + col].x; });
            Real eth = 0, eu = 0, ev = 0, epe = 0;
            if (tid < L) {
                const int i = tid;
                const Real nu = Y[i * kRow + i].x + Real(0.5);
                const Real nd = Y[(L + i) * kRow + i].x + Real(0.5);
                for (int n = 0; n < Z; ++n) {
                    const int k = nbk[n * L + i];
                    const Real tt = nbt[n * L + i];
                    const C a = Y[i * kRow + k], c = Y[(L + i) * kRow + k];
                    eth += (a.x + c.x) * tt;
                    if constexpr (PULSE) {
                        const C P = Tc[n * L + i];  // t (1 + e^{i phi}); the pulse part is P - t
                        epe += (a.x + c.x) * (P.x - tt) - (a.y + c.y) * P.y;
                    }
                }
                eth += s.tdiag[i] * (nu + nd);
                epe += s.epot[i] * (nu + nd);
                if (s.coulomb_type == 1) {
                    eu = udiag[i] * nu * nd;
                } else if (general) {
                    Real su = 0, sd = 0, vu = 0, vd = 0;
                    for (int jj = 0; jj < L; ++jj) {
                        const Real nuj = Y[jj * kRow + jj].x + Real(0.5);
                        const Real ndj = Y[(L + jj) * kRow + jj].x + Real(0.5);
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
        if (t == r.T - 1) {  // no step after the last record changes an output
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
    for (int m = 0; m < EPT; ++m) {
        const int row = w + W * m;
        if ((kExact || row < R2) && jon) {
            const int sg = kExact ? (m >= EPT / 2 ? 1 : 0) : (row >= L ? 1 : 0);
            const int i = row - sg * L;
            go[sg * LL + i * L + j] = y[m];
        }
    }
}

template <typename Real>
size_t lds_bytes(int L, int z, int W, bool pulse)
{
    const size_t C = 2 * sizeof(Real);
    return 2 * static_cast<size_t>(L) * kRow * C + (pulse ? static_cast<size_t>(z) * L * C : 0) +
           2 * kRow * sizeof(Real) + static_cast<size_t>(W) * 4 * sizeof(Real);
}

constexpr int kEpts[] = {2, 4, 8, 16};

// Waves per workgroup (8 or 16) and rows per wave for L sites; 0 when L does not fit.  Default:
// 16 waves in FP32 (8 rows each at L = 64 keeps y, A, p within 128 VGPRs), 8 waves in FP64.
template <typename Real>
int choose_shape(int L, int nt_hint, int* W)
{
    int w = (nt_hint == 1024) ? 16 : (nt_hint == 512 ? 8 : (sizeof(Real) == 4 ? 16 : 8));
    for (int e : kEpts)
        if (w * e >= 2 * L) {
            *W = w;
            return e;
        }
    return 0;
}

template <typename Real, int W, int EPT, int Z, bool PULSE, int LT>
hipError_t run(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec, hipStream_t st)
{
    auto kern = stencil_kernel<Real, W, EPT, Z, PULSE, LT>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const size_t lds = lds_bytes<Real>(s.L, s.z, W, PULSE);
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(64 * W), lds, st, s, r, t0, t1, state, rec,
                       static_cast<const RowNbrs<Real, Z>*>(s.rows));
    return hipGetLastError();
}

template <typename Real, int W, int EPT, bool PULSE>
hipError_t run_z(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec, hipStream_t st)
{
    if (s.z == 4) {
        // the 8x8-lattice shape: site count folded into the code
        if constexpr (!PULSE && ((W == 16 && EPT == 8) || (W == 8 && EPT == 16)))
            if (s.L == 64) return run<Real, W, EPT, 4, false, 64>(s, r, t0, t1, state, rec, st);
        return run<Real, W, EPT, 4, PULSE, 0>(s, r, t0, t1, state, rec, st);
    }
    if (s.z == 8) return run<Real, W, EPT, 8, PULSE, 0>(s, r, t0, t1, state, rec, st);
    return hipErrorInvalidValue;
}

template <typename Real, int W, bool PULSE>
hipError_t run_e(int ept, const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                 hipStream_t st)
{
    switch (ept) {
    case 2: return run_z<Real, W, 2, PULSE>(s, r, t0, t1, state, rec, st);
    case 4: return run_z<Real, W, 4, PULSE>(s, r, t0, t1, state, rec, st);
    case 8: return run_z<Real, W, 8, PULSE>(s, r, t0, t1, state, rec, st);
    case 16: return run_z<Real, W, 16, PULSE>(s, r, t0, t1, state, rec, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename Real>
bool stencil_fits(int L, int z, int* nt)
{
    if (L < 1 || L > kRow || stencil_padded_z(z) == 0) return false;
    int W = 0;
    if (!choose_shape<Real>(L, *nt, &W)) return false;
    if (lds_bytes<Real>(L, stencil_padded_z(z), W, true) > kLdsLimit) return false;
    *nt = 64 * W;
    return true;
}

template <typename Real>
hipError_t launch_stencil(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                          int nt_hint, hipStream_t st)
{
    int W = 0;
    const int ept = choose_shape<Real>(s.L, nt_hint, &W);
    if (!ept || (s.z != 4 && s.z != 8) || s.L > kRow || r.B <= 0) return hipErrorInvalidValue;
    const bool pulse = s.amp != nullptr;
    if (lds_bytes<Real>(s.L, s.z, W, pulse) > kLdsLimit) return hipErrorInvalidValue;
    if (W == 16) {
        return pulse ? run_e<Real, 16, true>(ept, s, r, t0, t1, state, rec, st)
                     : run_e<Real, 16, false>(ept, s, r, t0, t1, state, rec, st);
    }
    return pulse ? run_e<Real, 8, true>(ept, s, r, t0, t1, state, rec, st)
                 : run_e<Real, 8, false>(ept, s, r, t0, t1, state, rec, st);
}

#if FTWA_REAL == 1
template bool stencil_fits<float>(int, int, int*);
template hipError_t launch_stencil<float>(const DevSystem<float>&, const RunShape&, int, int, float*, float*, int,
                                          hipStream_t);
#else
template bool stencil_fits<double>(int, int, int*);
template hipError_t launch_stencil<double>(const DevSystem<double>&, const RunShape&, int, int, double*, double*, int,
                                           hipStream_t);
#endif

}  // namespace fpsamd
