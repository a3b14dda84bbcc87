// ftwa_gpsr.hip -- batched GPSR integrator (the reference's FHGPSRSolver trajectory loop) for small lattices.
//
// Restates FHGPSRSolver::runSimulation's trajectory body (src/FHGPSRSolver.cpp:770-826) with derivative
// (:110-133), step_Drift_RK4_38 (:168-201) and step (:608-637).  The diffusion is the analytic form of
// step_Diffusion (:211-232), whose second moments are the diffusion matrix the reference factorises
// (:241-287), rewritten with a single matrix product per spin (rhoH = I - rho, W_k = diag(xi_k)):
//     rhoH W1 rho + i rho W2 rhoH = W1 rho + i rho W2 - rho (W1 + i W2) rho, i.e. elementwise
//     d_ij = c [ (xi1_i + i xi2_j) rho_ij - sum_p rho_ip (xi1_p + i xi2_p) rho_pj ],   c = sqrt(i U(0,0)),
// spin down with conj(xi1), conj(xi2).
//
//   * one workgroup integrates one trajectory over a chunk of time points.  N = 8, 16 or 32 is the padded
//     lattice size (L <= N) and the workgroup has 2 RP N threads: thread t owns column j = t % N of spin
//     t / (RP N), rows r0, r0 + RP, ..., r0 = (t % RP N) / N -- N / RP elements of one spin (RP = 8 rows
//     per pass);
//   * GPSR's rho is not Hermitian, so every element of both spins carries RK registers (y, A, p)
//     (A the b-weighted sum, p = k1, k1 - k2, k1 - k2 + k3: stage inputs y, y + h p/3, y + h (A - 3p)/6,
//     y + h p; y' = y + h A/8);
//   * the stage input lives in LDS double buffered, [2][2 spins][N][N+1] complex: a stage reads one buffer,
//     writes the next stage input into the other, and ends with one barrier;
//   * the commutator -i[rho, Delta], Delta = tM + diag(h), runs over the hopping's neighbour lists (z padded
//     to 4 or 8): k_ij = -i [ sum_n Y_{i,k_n(j)} t_n(j) - sum_n t_n(i) Y_{k_n(i),j} + (h_j - h_i) Y_ij ],
//     h_i = pdiag_i + sum_l U_il Y^sbar_ll (complex, no 1/2);
//   * the diffusion increment d is formed in the first stage from the LDS copy of the step's initial state
//     and added in the last: y' = y + h A/8 + d.  The noise of a chunk's steps (2L complex values per step,
//     variance h) is drawn beforehand by gpsr_noise_kernel into HBM ([step][B][2][L] complex, 1 % of a
//     step's work in traffic) and each step's values are staged into LDS during the previous step's last
//     stage, which keeps the Philox / Box-Muller code out of the integrator's registers;
//   * a trajectory holding a NaN or Inf after a step is set to NaN at the start of the next (the flag is
//     the last stage's barrier, __syncthreads_or), as step() does at entry;
//   * every time point writes one record per trajectory: the record layout of the other integrators plus
//     Im rho_ii of both spins at the end.
// Padding rows and columns (i or j >= L) stay exactly zero: their hoppings and noise are zero.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::V2;
using dev::wave_sum;

constexpr int kLdsLimit = 160 * 1024;

template <typename C>
__device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Byte offsets of the dynamic LDS sections.
template <typename Real, int N, int Z>
struct GpsrLay {
    static constexpr int RS = N + 1;  // row stride (complex elements): odd, so columns spread over banks
    static constexpr int up16(int b) { return (b + 15) / 16 * 16; }
    static constexpr int oY = 0;                                                   // [2][2][N][RS] complex
    static constexpr int oK = up16(oY + 2 * 2 * N * RS * 2 * (int)sizeof(Real));   // [Z][N] int
    static constexpr int oT = up16(oK + Z * N * (int)sizeof(int));                 // [Z][N] hopping
    static constexpr int oS = up16(oT + Z * N * (int)sizeof(Real));                // [4][N] pdiag, U_ii, tM_ii, epot
    static constexpr int oU = up16(oS + 4 * N * (int)sizeof(Real));                // [N][N] U (general Coulomb)
    static constexpr int oX = up16(oU + N * N * (int)sizeof(Real));                // [2][N] complex step noise
    static constexpr int bytes = up16(oX + 2 * N * 2 * (int)sizeof(Real));
};

// RK registers after stage st with derivative k, and the next stage's input (the 3/8 rule in the
// (y, A, p) form): inputs y + h p/3, y + h (A - 3p)/6, y + h p after stages 0, 1, 2; y += h A/8 after 3.
template <typename Real>
__device__ __forceinline__ Real rk_stage(int st, Real& y, Real& A, Real& p, Real k, Real h1, Real h3, Real h6,
                                         Real h8)
{
    switch (st) {
    case 0:
        A = k;
        p = k;
        return y + h3 * p;
    case 1:
        A += Real(3) * k;
        p -= k;
        return y + h6 * (A - Real(3) * p);
    case 2:
        A += Real(3) * k;
        p += k;
        return y + h1 * p;
    default:
        A += k;
        y += h8 * A;
        return y;
    }
}

// Scaled diffusion noise of steps [g0, g0 + ns) for B trajectories: out[(gg * B + b) * 2L + k L + p] =
// sqrt(h) xi_k(p) of step g0 + gg (complex, interleaved).
template <typename Real>
__global__ void gpsr_noise_kernel(int L, int64_t B, int64_t b0, int64_t g0, int64_t ns, uint64_t seed, double h,
                                  Real* __restrict__ out)
{
    const int64_t total = ns * B * 2 * L;
    const double sh = sqrt(h);
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t gg = e / (B * 2 * L);
        const int64_t rem = e - gg * B * 2 * L;
        const int64_t b = rem / (2 * L);
        const int q = static_cast<int>(rem - b * 2 * L);
        double re, im;
        dev::gpsr_noise(seed, b0 + b, g0 + gg, q / L, q % L, re, im);
        out[2 * e] = static_cast<Real>(re * sh);
        out[2 * e + 1] = static_cast<Real>(im * sh);
    }
}

// RP rows per pass: 2 RP N threads, thread t owns column t % N of spin t / (RP N) and rows
// r0 + RP e (e < N / RP), r0 = (t % RP N) / N.
template <typename Real, int N, int Z, int RP>
__global__ void __launch_bounds__(2 * RP * N)
    gpsr_kernel(DevSystem<Real> s, RunShape r, int t0, int t1, Real* __restrict__ state, Real* __restrict__ rec,
                const Real* __restrict__ noise)
{
    using C = V2<Real>;
    using Lay = GpsrLay<Real, N, Z>;
    constexpr int E = N / RP, RS = Lay::RS, HALF = RP * N, NT = 2 * RP * N;
    static_assert(N % RP == 0 && 2 * N <= 64, "2N sites' energy terms are summed by wavefront 0");
    extern __shared__ __align__(16) unsigned char lds[];
    C* Ys = reinterpret_cast<C*>(lds + Lay::oY);
    int* nbk = reinterpret_cast<int*>(lds + Lay::oK);
    Real* nbt = reinterpret_cast<Real*>(lds + Lay::oT);
    Real* hd = reinterpret_cast<Real*>(lds + Lay::oS);
    Real* ud = hd + N;
    Real* tdg = ud + N;
    Real* epv = tdg + N;
    Real* Um = reinterpret_cast<Real*>(lds + Lay::oU);
    C* xis = reinterpret_cast<C*>(lds + Lay::oX);

    const int L = s.L;
    const int tid = threadIdx.x;
    const int sg = tid / HALF;
    const int lt = tid - sg * HALF;
    const int j = lt % N;
    const int r0 = lt / N;
    const int64_t b = blockIdx.x;
    const int64_t LL = static_cast<int64_t>(L) * L;
    const bool gen = s.coulomb_type == 2;
    const bool diffuse = s.gc_re != 0.0 || s.gc_im != 0.0;
    const Real gcr = static_cast<Real>(s.gc_re), gci = static_cast<Real>(s.gc_im);
    const Real h1 = static_cast<Real>(r.h), h3 = static_cast<Real>(r.h / 3.0), h6 = static_cast<Real>(r.h / 6.0),
               h8 = static_cast<Real>(r.h / 8.0);
    const Real cs = sg ? Real(-1) : Real(1);  // spin down takes conj(xi)

    // static tables
    for (int q = tid; q < Z * N; q += NT) {
        const int n = q / N, i = q - n * N;
        const bool on = i < L;
        nbk[q] = on ? s.nbr_k[n * L + i] : i;
        nbt[q] = on ? s.nbr_t[n * L + i] : Real(0);
    }
    for (int i = tid; i < N; i += NT) {
        const bool on = i < L;
        hd[i] = on ? s.pdiag[i] : Real(0);
        ud[i] = on && s.coulomb_type == 1 ? s.u_diag[i] : Real(0);
        tdg[i] = on ? s.tdiag[i] : Real(0);
        epv[i] = on ? s.epot[i] : Real(0);
    }
    for (int q = tid; q < N * N; q += NT) {
        const int i = q / N, l = q - i * N;
        Um[q] = gen && i < L && l < L ? s.U[i * L + l] : Real(0);
    }
    // noise slots of padding sites stay zero (padding elements multiply them by 0; garbage could be NaN)
    for (int q = tid; q < 2 * N; q += NT) xis[q] = C{Real(0), Real(0)};
    __syncthreads();
    // the thread's column neighbour list, in registers where the budget allows (it is the same for the whole
    // launch); FP64 at N = 32 reads it from the LDS tables at each use instead
    constexpr bool kColRegs = sizeof(Real) == 4 || N <= 16;
    int kc[Z];
    Real tc[Z];
#pragma unroll
    for (int n = 0; n < Z; ++n) {
        kc[n] = j < L ? s.nbr_k[n * L + j] : j;
        tc[n] = j < L ? s.nbr_t[n * L + j] : Real(0);
    }
    // state -> registers and the stage-input buffer 0
    C y[E], A[E], p[E], d[E];
    Real* gst = state + b * 2 * LL * 2 + static_cast<int64_t>(sg) * LL * 2;
    int bad = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = r0 + RP * e;
        C v{Real(0), Real(0)};
        if (i < L && j < L) {
            v.x = gst[2 * (i * L + j)];
            v.y = gst[2 * (i * L + j) + 1];
        }
        y[e] = v;
        A[e] = C{Real(0), Real(0)};
        p[e] = C{Real(0), Real(0)};
        d[e] = C{Real(0), Real(0)};
        bad |= !(isfinite(v.x) && isfinite(v.y));
        Ys[(sg * N + i) * RS + j] = v;
    }
    // stage the noise of step g (relative to the chunk's first step g0) into LDS
    const int64_t g0 = static_cast<int64_t>(t0) * r.nsub;
    const int64_t nsteps = static_cast<int64_t>(t1 < r.T ? t1 - t0 : r.T - 1 - t0) * r.nsub;
    auto draw_noise = [&](int64_t g) {
        if (tid < 2 * L && g - g0 < nsteps) {
            const Real* w = noise + ((g - g0) * r.B + b) * 2 * L * 2;
            xis[(tid / L) * N + tid % L] = C{w[2 * tid], w[2 * tid + 1]};
        }
    };
    if (diffuse) draw_noise(g0);
    bad = __syncthreads_or(bad);

    // Hartree field h_i of spin sp from buffer Y (both spins): pdiag + sum_l U_il Y^sbar_ll
    auto hval = [&](const C* Y, int sp, int i) -> C {
        const C* Yo = Y + (1 - sp) * N * RS;
        C v{hd[i], Real(0)};
        if (gen) {
#pragma unroll 1
            for (int l = 0; l < L; ++l) {
                const Real u = Um[i * N + l];
                const C o = Yo[l * RS + l];
                v.x += u * o.x;
                v.y += u * o.y;
            }
        } else {
            const C o = Yo[i * RS + i];
            v.x += ud[i] * o.x;
            v.y += ud[i] * o.y;
        }
        return v;
    };

    auto stage = [&](int st, int64_t g) {
        const C* Y = Ys + (st & 1) * 2 * N * RS;
        C* Yn = Ys + ((st + 1) & 1) * 2 * N * RS;
        const C* Ym = Y + sg * N * RS;
        const C hj = hval(Y, sg, j);
        C G[E];
        {
            if (st == 0 && diffuse) {
#pragma unroll
                for (int e = 0; e < E; ++e) G[e] = C{Real(0), Real(0)};
#pragma unroll 2
                for (int q = 0; q < L; ++q) {
                    const C a1 = xis[q], a2 = xis[N + q];
                    // eta = xi1 + i xi2 (conjugated noise for spin down)
                    const C eta{a1.x - cs * a2.y, cs * a1.y + a2.x};
                    const C z = cmul(eta, Ym[q * RS + j]);
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        const C w = Ym[(r0 + RP * e) * RS + q];
                        G[e].x += w.x * z.x - w.y * z.y;
                        G[e].y += w.x * z.y + w.y * z.x;
                    }
                }
            }
        }
        int flag = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            // the neighbour tables are read from LDS at each use through opaque offsets, so the compiler
            // does not hold E * Z decoded entries in registers across the time loop
            int i = r0 + RP * e, jc = j;
            asm volatile("" : "+v"(i), "+v"(jc));
            const C yij = Ym[i * RS + jc];
            C x{Real(0), Real(0)};
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const Real t = kColRegs ? tc[n] : nbt[n * N + jc];
                const C a = Ym[i * RS + (kColRegs ? kc[n] : nbk[n * N + jc])];
                x.x += a.x * t;
                x.y += a.y * t;
            }
#pragma unroll
            for (int n = 0; n < Z; ++n) {
                const Real t = nbt[n * N + i];
                const C a = Ym[nbk[n * N + i] * RS + jc];
                x.x -= t * a.x;
                x.y -= t * a.y;
            }
            const C hi = hval(Y, sg, i);
            const C dh{hj.x - hi.x, hj.y - hi.y};
            const C q = cmul(dh, yij);
            x.x += q.x;
            x.y += q.y;
            const C k{x.y, -x.x};  // -i x
            {
                if (st == 0 && diffuse) {
                    // c [ (xi1_i + i xi2_j) y_ij - G ]
                    const C a1 = xis[i], a2 = xis[N + j];
                    const C f{a1.x - cs * a2.y, cs * a1.y + a2.x};
                    const C m = cmul(f, yij);
                    const C sdf{m.x - G[e].x, m.y - G[e].y};
                    d[e] = C{gcr * sdf.x - gci * sdf.y, gcr * sdf.y + gci * sdf.x};
                }
            }
            C nxt{rk_stage(st, y[e].x, A[e].x, p[e].x, k.x, h1, h3, h6, h8),
                  rk_stage(st, y[e].y, A[e].y, p[e].y, k.y, h1, h3, h6, h8)};
            if (st == 3) {
                y[e].x += d[e].x;
                y[e].y += d[e].y;
                flag |= !(isfinite(y[e].x) && isfinite(y[e].y));
                nxt = y[e];
            }
            Yn[(sg * N + i) * RS + j] = nxt;
        }
        if (st == 3) {
            if (diffuse) draw_noise(g + 1);  // the next step's noise (step g's was read in its first stage)
            bad = __syncthreads_or(flag);
        } else {
            __syncthreads();
        }
    };

    const int P = L * (L + 1) / 2;
    for (int t = t0; t < t1; ++t) {
        // ---- record (src/FHGPSRSolver.cpp:785-817): buffer 0 holds y
        Real* rr = rec + (static_cast<int64_t>(t - t0) * r.B + b) * r.R;
        const C* Y0 = Ys + sg * N * RS;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = r0 + RP * e;
            if (i < L && j < L) {
                const C v = Y0[i * RS + j];
                if (i == j) {
                    rr[sg * L + i] = v.x;
                    rr[r.R - 2 * L + sg * L + i] = v.y;
                }
                if (i == r.ci && j == r.cj) {
                    rr[2 * L + 4 + 2 * sg] = v.x;
                    rr[2 * L + 5 + 2 * sg] = v.y;
                }
                if (r.dens && j >= i)
                    rr[2 * L + 8 + sg * P + dev::packed_row(L, i) + j] = (v.x + Y0[j * RS + i].x) / Real(2);
            }
        }
        if (tid < 64) {
            // E_tH = Re sum (rho_up + rho_dn) o tM, E_U = Re (rho_up . U rho_dn + rho_dn . U rho_up) / 2,
            // E_pe = Re sum (rho_up + rho_dn) o extPot (no E_V in GPSR)
            double eth = 0.0, eu = 0.0, epe = 0.0;
            if (tid < 2 * L) {
                const int sp = tid / L, i = tid - sp * L;
                const C* Yp = Ys + sp * N * RS;
#pragma unroll 1
                for (int n = 0; n < Z; ++n)
                    eth += static_cast<double>(Yp[i * RS + nbk[n * N + i]].x) * static_cast<double>(nbt[n * N + i]);
                const C vi = Yp[i * RS + i];
                eth += static_cast<double>(tdg[i]) * static_cast<double>(vi.x);
                epe += static_cast<double>(epv[i]) * static_cast<double>(vi.x);
                if (sp == 0) {
                    const C* Yd = Ys + N * RS;
                    const C di = Yd[i * RS + i];
                    if (gen) {
#pragma unroll 1
                        for (int l = 0; l < L; ++l) {
                            const C ul = Yp[l * RS + l], dl = Yd[l * RS + l];
                            const double u = static_cast<double>(Um[i * N + l]);
                            eu += u * (static_cast<double>(vi.x) * dl.x - static_cast<double>(vi.y) * dl.y +
                                       static_cast<double>(di.x) * ul.x - static_cast<double>(di.y) * ul.y) / 2.0;
                        }
                    } else {
                        eu += static_cast<double>(ud[i]) *
                              (static_cast<double>(vi.x) * di.x - static_cast<double>(vi.y) * di.y);
                    }
                }
            }
            eth = wave_sum(eth);
            eu = wave_sum(eu);
            epe = wave_sum(epe);
            if (tid == 0) {
                rr[2 * L + 0] = static_cast<Real>(eth);
                rr[2 * L + 1] = static_cast<Real>(eu);
                rr[2 * L + 2] = Real(0);
                rr[2 * L + 3] = static_cast<Real>(epe);
            }
        }
        if (t == r.T - 1) break;
        for (int sub = 0; sub < r.nsub; ++sub) {
            const int64_t g = static_cast<int64_t>(t) * r.nsub + sub;
            if (bad) {
#pragma unroll
                for (int e = 0; e < E; ++e) y[e] = C{__builtin_nan(""), __builtin_nan("")};
            }
#pragma unroll 1
            for (int st = 0; st < 4; ++st) stage(st, g);
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = r0 + RP * e;
        if (i < L && j < L) {
            gst[2 * (i * L + j)] = y[e].x;
            gst[2 * (i * L + j) + 1] = y[e].y;
        }
    }
}

template <typename Real, int N, int Z>
hipError_t run_gpsr(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                    const Real* noise, hipStream_t st)
{
    constexpr int bytes = GpsrLay<Real, N, Z>::bytes;
    static_assert(bytes <= kLdsLimit, "GPSR layout exceeds the LDS");
    // 8 rows per pass (16 rows, i.e. 1024 threads, cap the registers at 128 and spill more in FP64)
    constexpr int RP = 8;
    auto k = gpsr_kernel<Real, N, Z, RP>;
    if (bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(r.B)), dim3(2 * RP * N), bytes, st, s, r, t0, t1, state, rec,
                       noise);
    return hipGetLastError();
}

template <typename Real, int N>
hipError_t run_gpsr_z(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                      const Real* noise, hipStream_t st)
{
    return s.z == 4 ? run_gpsr<Real, N, 4>(s, r, t0, t1, state, rec, noise, st)
                    : run_gpsr<Real, N, 8>(s, r, t0, t1, state, rec, noise, st);
}

}  // namespace

template <typename Real>
bool gpsr_fits(const DevSystem<Real>& s)
{
    return s.L >= 1 && s.L <= 32 && (s.z == 4 || s.z == 8) && s.amp == nullptr;
}

template <typename Real>
size_t gpsr_noise_elems(int L, int64_t B, int nsub, int nt)
{
    return static_cast<size_t>(nt) * nsub * B * 2 * L * 2;
}

template <typename Real>
hipError_t launch_gpsr(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                       Real* noise, hipStream_t st)
{
    if (!gpsr_fits(s) || r.B < 1 || r.B > 0x7fffffff || t0 < 0 || t1 > r.T || t0 >= t1) return hipErrorInvalidValue;
    const int64_t nsteps = static_cast<int64_t>(t1 < r.T ? t1 - t0 : r.T - 1 - t0) * r.nsub;
    if ((s.gc_re != 0.0 || s.gc_im != 0.0) && nsteps > 0) {
        const int64_t total = nsteps * r.B * 2 * s.L;
        int64_t g = (total + 255) / 256;
        if (g > (1 << 20)) g = 1 << 20;
        hipLaunchKernelGGL(gpsr_noise_kernel<Real>, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, s.L, r.B, r.b0,
                           static_cast<int64_t>(t0) * r.nsub, nsteps, r.seed, r.h, noise);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (s.L <= 8) return run_gpsr_z<Real, 8>(s, r, t0, t1, state, rec, noise, st);
    if (s.L <= 16) return run_gpsr_z<Real, 16>(s, r, t0, t1, state, rec, noise, st);
    return run_gpsr_z<Real, 32>(s, r, t0, t1, state, rec, noise, st);
}

template bool gpsr_fits<float>(const DevSystem<float>&);
template bool gpsr_fits<double>(const DevSystem<double>&);
template size_t gpsr_noise_elems<float>(int, int64_t, int, int);
template size_t gpsr_noise_elems<double>(int, int64_t, int, int);
template hipError_t launch_gpsr<float>(const DevSystem<float>&, const RunShape&, int, int, float*, float*, float*,
                                       hipStream_t);
template hipError_t launch_gpsr<double>(const DevSystem<double>&, const RunShape&, int, int, double*, double*,
                                        double*, hipStream_t);

}  // namespace fpsamd
