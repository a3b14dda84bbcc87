// ftwa_kernels.hip -- gfx950 kernels of the fTWA trajectory engine.
//
// Reference path restated: FHfTWASolver::runSimulation (src/FHfTWASolver.cpp:222-556) with
// derivative (:568-600) and step_RK4_38 (:629-661).  Kernel families in this file:
//
//   sample   Gaussian initial states rho = diag(n - 1/2) + S.W (:342-362), noise from Philox4x32-10
//   (the batched stencil integrator lives in ftwa_stencil.hip)
//   plain    the FP64 reference-order baseline: dense Delta, rho*Delta - Delta*rho from global memory,
//            one launch per time point
//   reduce   fixed-order, compensated FP64 ensemble sums of per-trajectory records
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
using dev::Neumaier;
using dev::wave_sum;

using dev::pair_index;
__device__ __forceinline__ void noise_value(uint64_t seed, int64_t traj, int spin, uint32_t pair, double& re,
                                            double& im)
{
    dev::initial_noise(seed, traj, spin, pair, re, im);
}

// ---- sampling -------------------------------------------------------------------------
template <typename Real>
__global__ void sample_kernel(int L, const double* __restrict__ n_up, const double* __restrict__ n_dn,
                              const double* __restrict__ noise, uint64_t seed, int64_t b0, int64_t B, double shift,
                              Real* __restrict__ state)
{
    const int64_t LL = static_cast<int64_t>(L) * L;
    const int64_t total = B * 2 * LL;
    const int64_t npairs = static_cast<int64_t>(L) * (L - 1) / 2;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = e / (2 * LL);
        const int64_t rem = e - b * 2 * LL;
        const int sg = static_cast<int>(rem / LL);
        const int64_t ij = rem - sg * LL;
        const int i = static_cast<int>(ij / L);
        const int j = static_cast<int>(ij - static_cast<int64_t>(i) * L);
        const double* n = sg ? n_dn : n_up;
        double re = 0.0, im = 0.0;
        if (i == j) {
            re = n[i] - shift;
        } else {
            const int a = i < j ? i : j;
            const int c = i < j ? j : i;
            const int64_t p = pair_index(L, a, c);
            double wr, wi;
            if (noise) {
                const double* w = noise + ((b * 2 + sg) * npairs + p) * 2;
                wr = w[0];
                wi = w[1];
            } else {
                noise_value(seed, b0 + b, sg, static_cast<uint32_t>(p), wr, wi);
            }
            dev::sampled_offdiag(n[a], n[c], wr, wi, re, im);
            if (i > j) im = -im;
        }
        state[2 * e] = static_cast<Real>(re);
        state[2 * e + 1] = static_cast<Real>(im);
    }
}

#if FTWA_REAL != 1
__global__ void noise_kernel(int L, uint64_t seed, int64_t b0, int64_t B, double* __restrict__ out)
{
    const int64_t npairs = static_cast<int64_t>(L) * (L - 1) / 2;
    const int64_t total = B * 2 * npairs;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = e / (2 * npairs);
        const int64_t rem = e - b * 2 * npairs;
        const int sg = static_cast<int>(rem / npairs);
        const int64_t p = rem - sg * npairs;
        double re, im;
        noise_value(seed, b0 + b, sg, static_cast<uint32_t>(p), re, im);
        out[2 * e] = re;
        out[2 * e + 1] = im;
    }
}
#endif

// rho <- R rho R^dagger for one (trajectory, spin) per workgroup; tmp holds R rho.
template <typename Real>
__global__ void rotate_kernel(int L, const double* __restrict__ rot, Real* __restrict__ state, Real* __restrict__ tmp)
{
    const int64_t LL = static_cast<int64_t>(L) * L;
    const int sg = blockIdx.x & 1;
    Real* M = state + static_cast<int64_t>(blockIdx.x) * LL * 2;
    Real* W = tmp + static_cast<int64_t>(blockIdx.x) * LL * 2;
    const double* R = rot + sg * LL * 2;
    for (int64_t e = threadIdx.x; e < LL; e += blockDim.x) {
        const int i = static_cast<int>(e / L), j = static_cast<int>(e % L);
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < L; ++k) {
            const double rr = R[2 * (i * L + k)], ri = R[2 * (i * L + k) + 1];
            const double mr = M[2 * (k * L + j)], mi = M[2 * (k * L + j) + 1];
            ar += rr * mr - ri * mi;
            ai += rr * mi + ri * mr;
        }
        W[2 * e] = static_cast<Real>(ar);
        W[2 * e + 1] = static_cast<Real>(ai);
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < LL; e += blockDim.x) {
        const int i = static_cast<int>(e / L), j = static_cast<int>(e % L);
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < L; ++k) {
            const double wr = W[2 * (i * L + k)], wi = W[2 * (i * L + k) + 1];
            const double rr = R[2 * (j * L + k)], ri = -R[2 * (j * L + k) + 1];  // conj(R_jk)
            ar += wr * rr - wi * ri;
            ai += wr * ri + wi * rr;
        }
        M[2 * e] = static_cast<Real>(ar);
        M[2 * e + 1] = static_cast<Real>(ai);
    }
}

// ---- plain FP64 path ---------------------------------------------------------------------
//
// The reference's arithmetic as a straightforward GPU program: one workgroup per trajectory,
// the dense L x L Delta built per stage (src/FHfTWASolver.cpp:571-596), the two dense products
// rho*Delta - Delta*rho read from global memory (:598-599), the RK4 "3/8" combinations written back
// (:644-660), and one launch per time point.  It is the FP64 baseline the batched engine is timed
// against, and the GPU twin of the CPU oracle.
#if FTWA_REAL != 1
constexpr int kPlainThreads = 256;

__device__ __forceinline__ void record_energies_tail(double* rr, int L, double* ew, int nwave)
{
    if (threadIdx.x == 0) {
        double e4[4] = {0, 0, 0, 0};
        for (int w = 0; w < nwave; ++w)
            for (int q = 0; q < 4; ++q) e4[q] += ew[w * 4 + q];
        for (int q = 0; q < 4; ++q) rr[2 * L + q] = e4[q];
    }
}

__global__ void __launch_bounds__(kPlainThreads)
    plain_kernel(DevSystem<double> s, RunShape r, int t, double* __restrict__ state, double* __restrict__ work,
                 double* __restrict__ rec)
{
    __shared__ double ew[(kPlainThreads / 64) * 4];
    const int L = s.L;
    const int LL = L * L;
    const int NE = 2 * LL;
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    const int64_t b = blockIdx.x;
    double* y = state + b * NE * 2;
    double* wk = work + b * static_cast<int64_t>(plain_work_doubles(L));
    double* Ys = wk;
    double* K = Ys + NE * 2;       // [4][NE][2]
    double* D = K + 4 * NE * 2;    // [2][L][L][2]
    double* Dd = D + NE * 2;       // [2][L][L][2] GPSR diffusion increment
    double* hc = Dd + NE * 2;      // [2][L][2]
    double* xi = hc + 4 * L;       // [2][L][2] GPSR step noise xi_1, xi_2
    const bool gpsr = s.gpsr != 0;
    const bool general = s.coulomb_type == 2 && !gpsr;  // GPSR's Delta has no V (src/FHGPSRSolver.cpp:112-113)
    const double hf = gpsr ? 0.0 : 0.5;                 // N = rho + 1/2 (fTWA, HF) or rho (GPSR)
    const double h = r.h;

    // ---- record (src/FHfTWASolver.cpp:366-418) ----
    double* rr = rec + b * r.R;
    {
        for (int e = tid; e < 2 * L; e += NT) {
            const int sg = e / L, i = e - sg * L;
            rr[e] = y[2 * (sg * LL + i * L + i)];
            if (gpsr) rr[r.R - 2 * L + e] = y[2 * (sg * LL + i * L + i) + 1];
        }
        if (tid == 0) {
            const int c = r.ci * L + r.cj;
            rr[2 * L + 4] = y[2 * c];
            rr[2 * L + 5] = y[2 * c + 1];
            rr[2 * L + 6] = y[2 * (LL + c)];
            rr[2 * L + 7] = y[2 * (LL + c) + 1];
        }
        if (r.dens)  // GPSR's rho is not Hermitian: the symmetric part (Re rho_ij + Re rho_ji) / 2
            dev::record_density(rr + 2 * L + 8, L, tid & 63, tid >> 6, (NT + 63) / 64, [&](int row, int col) {
                const int sr = row / L, i = row - sr * L;
                return gpsr ? (y[2 * (row * L + col)] + y[2 * ((sr * L + col) * L + i)]) / 2.0 : y[2 * (row * L + col)];
            });
        double eth = 0, eu = 0, ev = 0, epe = 0;
        const double a = s.amp ? s.amp[static_cast<int64_t>(t) * r.nsub * 4] : 0.0;
        for (int i = tid; i < L; i += NT) {
            for (int jj = 0; jj < L; ++jj) {
                const double nr = y[2 * (i * L + jj)] + y[2 * (LL + i * L + jj)] + (i == jj ? 2.0 * hf : 0.0);
                const double ni = y[2 * (i * L + jj) + 1] + y[2 * (LL + i * L + jj) + 1];
                const double tt = s.tM[i * L + jj];
                eth += nr * tt;
                if (s.amp && tt != 0.0) {
                    double sn, cs;
                    sincos(a * s.pd[i * L + jj], &sn, &cs);
                    epe += nr * tt * cs - ni * tt * sn;
                }
            }
            const double nu = y[2 * (i * L + i)] + hf, nui = y[2 * (i * L + i) + 1];
            const double nd = y[2 * (LL + i * L + i)] + hf, ndi = y[2 * (LL + i * L + i) + 1];
            epe += s.pot[i] * (nu + nd);
            if (s.coulomb_type != 0) {
                for (int jj = 0; jj < L; ++jj) {
                    const double nuj = y[2 * (jj * L + jj)] + hf, nujj = y[2 * (jj * L + jj) + 1];
                    const double ndj = y[2 * (LL + jj * L + jj)] + hf, ndjj = y[2 * (LL + jj * L + jj) + 1];
                    const double u = s.U[i * L + jj];
                    eu += u * ((nu * ndj - nui * ndjj) + (nd * nuj - ndi * nujj)) / 2.0;
                    if (general) {
                        const double v = s.V[i * L + jj];
                        ev += v * ((nu * nuj - nui * nujj) + (nd * ndj - ndi * ndjj)) / 2.0;
                        if (s.fock) {  // - Re sum V o N o N^T / 2 per spin (src/FHHFSolver.cpp:356-357)
                            for (int sg = 0; sg < 2; ++sg) {
                                const double* Ym = y + 2 * sg * LL;
                                const double ar = Ym[2 * (i * L + jj)] + (i == jj ? 0.5 : 0.0);
                                const double ai = Ym[2 * (i * L + jj) + 1];
                                const double br = Ym[2 * (jj * L + i)] + (i == jj ? 0.5 : 0.0);
                                const double bi = Ym[2 * (jj * L + i) + 1];
                                ev -= v * (ar * br - ai * bi) / 2.0;
                            }
                        }
                    }
                }
            }
        }
        eth = wave_sum(eth);
        eu = wave_sum(eu);
        ev = wave_sum(ev);
        epe = wave_sum(epe);
        if ((tid & 63) == 0) {
            ew[(tid >> 6) * 4 + 0] = eth;
            ew[(tid >> 6) * 4 + 1] = eu;
            ew[(tid >> 6) * 4 + 2] = ev;
            ew[(tid >> 6) * 4 + 3] = epe;
        }
        __syncthreads();
        record_energies_tail(rr, L, ew, (NT + 63) / 64);
    }
    if (t == r.T - 1) return;

    for (int sub = 0; sub < r.nsub; ++sub) {
        const int64_t g = static_cast<int64_t>(t) * r.nsub + sub;
        if (gpsr) {
            // FHGPSRSolver::step (src/FHGPSRSolver.cpp:608-637): a trajectory holding a NaN or Inf becomes
            // all NaN and is not stepped; otherwise the diffusion increment from the step's initial state,
            // the analytic form of step_Diffusion (:211-232): c (rhoH W1 rho + i rho W2 rhoH), spin down
            // with conj(W); the noise has variance h per step
            int bad = 0;
            for (int e = tid; e < 2 * NE; e += NT) bad |= !isfinite(y[e]);
            if (__syncthreads_or(bad)) {
                for (int e = tid; e < 2 * NE; e += NT) y[e] = __builtin_nan("");
                continue;
            }
            const double sh = sqrt(h);
            for (int e = tid; e < 2 * L; e += NT) {
                double re, im;
                dev::gpsr_noise(r.seed, r.b0 + b, g, e / L, e % L, re, im);
                xi[2 * e] = re * sh;
                xi[2 * e + 1] = im * sh;
            }
            __syncthreads();
            for (int e = tid; e < NE; e += NT) {
                const int sg = e / LL, ij = e - sg * LL, i = ij / L, jj = ij - i * L;
                const double* Ym = y + 2 * sg * LL;
                const double cs = sg ? -1.0 : 1.0;
                double ar = 0, ai = 0, br = 0, bi = 0;
                for (int k = 0; k < L; ++k) {
                    const double hr = (i == k ? 1.0 : 0.0) - Ym[2 * (i * L + k)], hi = -Ym[2 * (i * L + k) + 1];
                    const double w1r = xi[2 * k], w1i = cs * xi[2 * k + 1];
                    const double m1r = hr * w1r - hi * w1i, m1i = hr * w1i + hi * w1r;
                    const double pr = Ym[2 * (k * L + jj)], pi = Ym[2 * (k * L + jj) + 1];
                    ar += m1r * pr - m1i * pi;
                    ai += m1r * pi + m1i * pr;
                    const double qr = Ym[2 * (i * L + k)], qi = Ym[2 * (i * L + k) + 1];
                    const double w2r = xi[2 * (L + k)], w2i = cs * xi[2 * (L + k) + 1];
                    const double m2r = qr * w2r - qi * w2i, m2i = qr * w2i + qi * w2r;
                    const double gr = (k == jj ? 1.0 : 0.0) - Ym[2 * (k * L + jj)], gi = -Ym[2 * (k * L + jj) + 1];
                    br += m2r * gr - m2i * gi;
                    bi += m2r * gi + m2i * gr;
                }
                const double sr = ar - bi, si = ai + br;  // (rhoH W1 rho) + i (rho W2 rhoH)
                Dd[2 * e] = s.gc_re * sr - s.gc_im * si;
                Dd[2 * e + 1] = s.gc_re * si + s.gc_im * sr;
            }
        }
        for (int st = 0; st < 4; ++st) {
            const double* Yin = (st == 0) ? y : Ys;
            // Hartree potential, complex as in the reference (U (rho_diag + 1/2))
            for (int e = tid; e < 2 * L; e += NT) {
                const int sg = e / L, i = e - sg * L;
                double hr = 0.0, hi = 0.0;
                if (s.coulomb_type != 0) {
                    for (int jj = 0; jj < L; ++jj) {
                        const double u = s.U[i * L + jj];
                        hr += u * (Yin[2 * ((1 - sg) * LL + jj * L + jj)] + hf);
                        hi += u * Yin[2 * ((1 - sg) * LL + jj * L + jj) + 1];
                    }
                    if (general)
                        for (int jj = 0; jj < L; ++jj) {
                            const double v = s.V[i * L + jj];
                            hr += v * (Yin[2 * (sg * LL + jj * L + jj)] + 0.5);
                            hi += v * Yin[2 * (sg * LL + jj * L + jj) + 1];
                        }
                }
                hc[2 * e] = hr + s.pot[i];
                hc[2 * e + 1] = hi;
            }
            __syncthreads();
            const double a = s.amp ? s.amp[g * 4 + st] : 0.0;
            for (int e = tid; e < NE; e += NT) {
                const int sg = e / LL, ij = e - sg * LL, i = ij / L, jj = ij - i * L;
                const double tt = s.tM[ij];
                double dr = tt, di = 0.0;
                if (i == jj) {
                    dr += hc[2 * (sg * L + i)];
                    di += hc[2 * (sg * L + i) + 1];
                }
                if (s.amp && tt != 0.0) {
                    double sn, cs;
                    sincos(a * s.pd[ij], &sn, &cs);
                    dr += tt * cs;
                    di += tt * sn;
                }
                if (s.fock) {  // Fock exchange -N_sigma o V (src/FHHFSolver.cpp:587-588), N = rho + I/2
                    const double v = s.V[ij];
                    dr -= v * (Yin[2 * e] + (i == jj ? 0.5 : 0.0));
                    di -= v * Yin[2 * e + 1];
                }
                D[2 * e] = dr;
                D[2 * e + 1] = di;
            }
            __syncthreads();
            double* Ko = K + st * NE * 2;
            for (int e = tid; e < NE; e += NT) {
                const int sg = e / LL, ij = e - sg * LL, i = ij / L, jj = ij - i * L;
                const double* Ym = Yin + 2 * sg * LL;
                const double* Dm = D + 2 * sg * LL;
                double ar = 0, ai = 0, br = 0, bi = 0;
                for (int k = 0; k < L; ++k) {
                    const double r1r = Ym[2 * (i * L + k)], r1i = Ym[2 * (i * L + k) + 1];
                    const double d1r = Dm[2 * (k * L + jj)], d1i = Dm[2 * (k * L + jj) + 1];
                    ar += r1r * d1r - r1i * d1i;
                    ai += r1r * d1i + r1i * d1r;
                    const double d2r = Dm[2 * (i * L + k)], d2i = Dm[2 * (i * L + k) + 1];
                    const double r2r = Ym[2 * (k * L + jj)], r2i = Ym[2 * (k * L + jj) + 1];
                    br += d2r * r2r - d2i * r2i;
                    bi += d2r * r2i + d2i * r2r;
                }
                Ko[2 * e] = ai - bi;       // -i (c_r + i c_i) = c_i - i c_r
                Ko[2 * e + 1] = -(ar - br);
            }
            __syncthreads();
            const double* K1 = K;
            const double* K2 = K + NE * 2;
            const double* K3 = K + 2 * NE * 2;
            const double* K4 = K + 3 * NE * 2;
            for (int e = tid; e < 2 * NE; e += NT) {
                if (st == 0)
                    Ys[e] = y[e] + K1[e] * h / 3.0;
                else if (st == 1)
                    Ys[e] = y[e] + (-K1[e] / 3.0 + K2[e]) * h;
                else if (st == 2)
                    Ys[e] = y[e] + (K1[e] - K2[e] + K3[e]) * h;
                else if (gpsr)
                    y[e] = y[e] + (K1[e] + K2[e] * 3.0 + K3[e] * 3.0 + K4[e]) * h / 8.0 + Dd[e];
                else
                    y[e] = y[e] + (K1[e] + K2[e] * 3.0 + K3[e] * 3.0 + K4[e]) * h / 8.0;
            }
            __syncthreads();
        }
    }
}

#endif

// ---- ensemble reduction ---------------------------------------------------------------------

// im_off >= 0 (GPSR records): Im rho_up_ii, Im rho_dn_ii sit at [im_off, im_off + 2L) and the sums are
// of the complex diagonals: S_uu = sum |rho_up_ii|^2, S_ud = sum Re(rho_up_ii rho_dn_ii), the last slot
// sum Re(rho_up_ii rho_dn_ii)^2.
template <typename Real>
__global__ void reduce_partial_kernel(int L, int64_t B, int R, int nchunk, int im_off, const Real* __restrict__ rec,
                                      double* __restrict__ part)
{
    const int t = blockIdx.x;
    const int c = blockIdx.y;
    const int64_t CB = (B + nchunk - 1) / nchunk;
    const int64_t bb0 = static_cast<int64_t>(c) * CB;
    const int64_t bb1 = (bb0 + CB < B) ? bb0 + CB : B;
    const int Q = 6 * L + 12;
    constexpr int kRowsB = 16;  // rows loaded per batch
    const Real* rt = rec + static_cast<int64_t>(t) * B * R;
    double* po = part + (static_cast<int64_t>(t) * nchunk + c) * Q * 2;
    for (int slot = threadIdx.x; slot < L + 8; slot += blockDim.x) {
        if (slot < L) {
            Neumaier su, sd, suu, sdd, sud, snn;
            if (im_off < 0) {
                auto add = [&](double u, double d) {
                    su.add(u);
                    sd.add(d);
                    suu.add(u * u);
                    sdd.add(d * d);
                    sud.add(u * d);
                    const double nn = (u + 0.5) * (d + 0.5);
                    snn.add(nn * nn);
                };
                // rows read 16 at a time before they are added (the loads overlap instead of each waiting
                // out the memory latency behind the previous row's sums); a batch's terms summed in plain
                // FP64 in row order, the batch sums compensated
                int64_t bi = bb0;
                for (; bi + kRowsB <= bb1; bi += kRowsB) {
                    Real ub[kRowsB], db[kRowsB];
#pragma unroll
                    for (int q = 0; q < kRowsB; ++q) {
                        ub[q] = rt[(bi + q) * R + slot];
                        db[q] = rt[(bi + q) * R + L + slot];
                    }
                    double bu = 0, bd = 0, buu = 0, bdd = 0, bud = 0, bnn = 0;
#pragma unroll
                    for (int q = 0; q < kRowsB; ++q) {
                        const double u = static_cast<double>(ub[q]), d = static_cast<double>(db[q]);
                        bu += u;
                        bd += d;
                        buu += u * u;
                        bdd += d * d;
                        bud += u * d;
                        const double nn = (u + 0.5) * (d + 0.5);
                        bnn += nn * nn;
                    }
                    su.add(bu);
                    sd.add(bd);
                    suu.add(buu);
                    sdd.add(bdd);
                    sud.add(bud);
                    snn.add(bnn);
                }
                for (; bi < bb1; ++bi) add(static_cast<double>(rt[bi * R + slot]), static_cast<double>(rt[bi * R + L + slot]));
            } else {
                for (int64_t bi = bb0; bi < bb1; ++bi) {
                    const double u = static_cast<double>(rt[bi * R + slot]);
                    const double d = static_cast<double>(rt[bi * R + L + slot]);
                    const double ui = static_cast<double>(rt[bi * R + im_off + slot]);
                    const double di = static_cast<double>(rt[bi * R + im_off + L + slot]);
                    su.add(u);
                    sd.add(d);
                    suu.add(u * u + ui * ui);
                    sdd.add(d * d + di * di);
                    const double ud = u * d - ui * di;
                    sud.add(ud);
                    snn.add(ud * ud);
                }
            }
            const Neumaier* all[6] = {&su, &sd, &suu, &sdd, &sud, &snn};
            for (int k = 0; k < 6; ++k) {
                po[(6 * slot + k) * 2] = all[k]->s;
                po[(6 * slot + k) * 2 + 1] = all[k]->c;
            }
        } else if (slot < L + 4) {
            const int k = slot - L;
            Neumaier se, se2;
            int64_t bi = bb0;
            for (; bi + kRowsB <= bb1; bi += kRowsB) {
                Real eb[kRowsB];
#pragma unroll
                for (int q = 0; q < kRowsB; ++q) eb[q] = rt[(bi + q) * R + 2 * L + k];
                double be = 0, be2 = 0;
#pragma unroll
                for (int q = 0; q < kRowsB; ++q) {
                    const double e = static_cast<double>(eb[q]);
                    be += e;
                    be2 += e * e;
                }
                se.add(be);
                se2.add(be2);
            }
            for (; bi < bb1; ++bi) {
                const double e = static_cast<double>(rt[bi * R + 2 * L + k]);
                se.add(e);
                se2.add(e * e);
            }
            po[(6 * L + k) * 2] = se.s;
            po[(6 * L + k) * 2 + 1] = se.c;
            po[(6 * L + 4 + k) * 2] = se2.s;
            po[(6 * L + 4 + k) * 2 + 1] = se2.c;
        } else {
            const int k = slot - L - 4;
            Neumaier sc;
            int64_t bi = bb0;
            for (; bi + kRowsB <= bb1; bi += kRowsB) {
                Real cb[kRowsB];
#pragma unroll
                for (int q = 0; q < kRowsB; ++q) cb[q] = rt[(bi + q) * R + 2 * L + 4 + k];
                double bc = 0;
#pragma unroll
                for (int q = 0; q < kRowsB; ++q) bc += static_cast<double>(cb[q]);
                sc.add(bc);
            }
            for (; bi < bb1; ++bi) sc.add(static_cast<double>(rt[bi * R + 2 * L + 4 + k]));
            po[(6 * L + 8 + k) * 2] = sc.s;
            po[(6 * L + 8 + k) * 2 + 1] = sc.c;
        }
    }
}

__global__ void reduce_final_kernel(int Q, int nchunk, int nt, int t0, const double* __restrict__ part,
                                    double* __restrict__ acc)
{
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= static_cast<int64_t>(nt) * Q) return;
    const int t = static_cast<int>(idx / Q);
    const int q = static_cast<int>(idx - static_cast<int64_t>(t) * Q);
    Neumaier a;
    // chunks loaded 16 at a time before they are added in order (the loads overlap; the sum is unchanged)
    constexpr int kC = 16;
    const double2* pp = reinterpret_cast<const double2*>(part) + static_cast<int64_t>(t) * nchunk * Q + q;
    int c = 0;
    for (; c + kC <= nchunk; c += kC) {
        double2 v[kC];
#pragma unroll
        for (int k = 0; k < kC; ++k) v[k] = pp[static_cast<int64_t>(c + k) * Q];
#pragma unroll
        for (int k = 0; k < kC; ++k) {
            a.add(v[k].x);
            a.add(v[k].y);
        }
    }
    for (; c < nchunk; ++c) {
        const double2 v = pp[static_cast<int64_t>(c) * Q];
        a.add(v.x);
        a.add(v.y);
    }
    double* o = acc + (static_cast<int64_t>(t0 + t) * Q + q) * 2;
    o[0] = a.s;
    o[1] = a.c;
}

// Ensemble sums of the density part of the records ([2L+8, R): packed Re rho, both spins), per time
// point: partial sums over fixed trajectory chunks (grid nt x nchunk x column blocks, one column per
// thread, reads coalesced along a record), then a fixed-order compensated sum over the chunks.
template <typename Real>
__global__ void density_partial_kernel(int L, int64_t B, int R, int nchunk, const Real* __restrict__ rec,
                                       double* __restrict__ part)
{
    const int t = blockIdx.x;
    const int c = blockIdx.y;
    const int P2 = L * (L + 1);
    const int col = blockIdx.z * blockDim.x + threadIdx.x;
    if (col >= P2) return;
    const int64_t CB = (B + nchunk - 1) / nchunk;
    const int64_t bb0 = static_cast<int64_t>(c) * CB;
    const int64_t bb1 = (bb0 + CB < B) ? bb0 + CB : B;
    const Real* rt = rec + static_cast<int64_t>(t) * B * R + 2 * L + 8 + col;
    Neumaier a;
    for (int64_t bi = bb0; bi < bb1; ++bi) a.add(static_cast<double>(rt[bi * R]));
    double* po = part + ((static_cast<int64_t>(t) * nchunk + c) * P2 + col) * 2;
    po[0] = a.s;
    po[1] = a.c;
}

__global__ void density_final_kernel(int P2, int nchunk, int nt, const double* __restrict__ part,
                                     double* __restrict__ S)
{
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= static_cast<int64_t>(nt) * P2) return;
    const int t = static_cast<int>(idx / P2);
    const int q = static_cast<int>(idx - static_cast<int64_t>(t) * P2);
    Neumaier a;
    for (int c = 0; c < nchunk; ++c) {
        const double* pp = part + ((static_cast<int64_t>(t) * nchunk + c) * P2 + q) * 2;
        a.add(pp[0]);
        a.add(pp[1]);
    }
    S[idx] = a.s + a.c;
}

inline unsigned grid_for(int64_t total, int block)
{
    int64_t g = (total + block - 1) / block;
    if (g > (1 << 20)) g = 1 << 20;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

}  // namespace

template <typename Real>
hipError_t launch_sample(const DevSystem<Real>& s, const double* n_up, const double* n_dn, const double* noise,
                         uint64_t seed, int64_t b0, int64_t B, Real* state, hipStream_t st)
{
    const int64_t total = B * 2 * static_cast<int64_t>(s.L) * s.L;
    const double shift = s.gpsr ? 0.0 : 0.5;  // GPSR: the normally ordered diag(n), no 1/2
    hipLaunchKernelGGL(sample_kernel<Real>, dim3(grid_for(total, 256)), dim3(256), 0, st, s.L, n_up, n_dn, noise, seed,
                       b0, B, shift, state);
    return hipGetLastError();
}

#if FTWA_REAL != 1
hipError_t launch_noise(int L, uint64_t seed, int64_t b0, int64_t B, double* out, hipStream_t st)
{
    const int64_t total = B * 2 * (static_cast<int64_t>(L) * (L - 1) / 2);
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(noise_kernel, dim3(grid_for(total, 256)), dim3(256), 0, st, L, seed, b0, B, out);
    return hipGetLastError();
}
#endif

template <typename Real>
hipError_t launch_rotate(int L, const double* rot, int64_t B, Real* state, Real* scratch, hipStream_t st)
{
    hipLaunchKernelGGL(rotate_kernel<Real>, dim3(static_cast<unsigned>(B * 2)), dim3(256), 0, st, L, rot, state,
                       scratch);
    return hipGetLastError();
}

#if FTWA_REAL != 1
hipError_t launch_plain(const DevSystem<double>& s, const RunShape& r, int t, double* state, double* work, double* rec,
                        hipStream_t st)
{
    hipLaunchKernelGGL(plain_kernel, dim3(static_cast<unsigned>(r.B)), dim3(kPlainThreads), 0, st, s, r, t, state,
                       work, rec);
    return hipGetLastError();
}
#endif

#if FTWA_REAL != 1
int reduce_chunks(int64_t B)
{
    // chunks of 256 trajectories (at most 1024 chunks): enough workgroups to cover the records' read
    int64_t n = (B + 255) / 256;
    if (n > 1024) n = 1024;
    if (n < 1) n = 1;
    return static_cast<int>(n);
}
#endif

template <typename Real>
hipError_t launch_density_reduce(int L, int64_t B, int R, int nt, const Real* rec, double* part, int nchunk,
                                 double* S, hipStream_t st)
{
    const int P2 = L * (L + 1);
    hipLaunchKernelGGL(density_partial_kernel<Real>, dim3(nt, nchunk, (P2 + 255) / 256), dim3(256), 0, st, L, B, R,
                       nchunk, rec, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(density_final_kernel, dim3(grid_for(static_cast<int64_t>(nt) * P2, 256)), dim3(256), 0, st, P2,
                       nchunk, nt, part, S);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_reduce(int L, int64_t B, int R, int nt, int t0, const Real* rec, double* part, int nchunk,
                         double* acc, hipStream_t st, int im_off)
{
    const int Q = 6 * L + 12;
    const int threads = ((L + 8 + 63) / 64) * 64 > 256 ? 256 : ((L + 8 + 63) / 64) * 64;
    hipLaunchKernelGGL(reduce_partial_kernel<Real>, dim3(nt, nchunk), dim3(threads), 0, st, L, B, R, nchunk, im_off,
                       rec, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t total = static_cast<int64_t>(nt) * Q;
    hipLaunchKernelGGL(reduce_final_kernel, dim3(grid_for(total, 256)), dim3(256), 0, st, Q, nchunk, nt, t0, part, acc);
    return hipGetLastError();
}

// explicit instantiations: FTWA_REAL selects the precision of this translation unit (1 = float,
// 2 = double); the FP64-only entry points live in the double unit.
#if FTWA_REAL == 1
template hipError_t launch_sample<float>(const DevSystem<float>&, const double*, const double*, const double*, uint64_t,
                                         int64_t, int64_t, float*, hipStream_t);
template hipError_t launch_rotate<float>(int, const double*, int64_t, float*, float*, hipStream_t);
template hipError_t launch_reduce<float>(int, int64_t, int, int, int, const float*, double*, int, double*, hipStream_t,
                                          int);
template hipError_t launch_density_reduce<float>(int, int64_t, int, int, const float*, double*, int, double*,
                                                 hipStream_t);
#else
template hipError_t launch_sample<double>(const DevSystem<double>&, const double*, const double*, const double*,
                                          uint64_t, int64_t, int64_t, double*, hipStream_t);
template hipError_t launch_rotate<double>(int, const double*, int64_t, double*, double*, hipStream_t);
template hipError_t launch_reduce<double>(int, int64_t, int, int, int, const double*, double*, int, double*,
                                          hipStream_t, int);
template hipError_t launch_density_reduce<double>(int, int64_t, int, int, const double*, double*, int, double*,
                                                  hipStream_t);
#endif

}  // namespace fpsamd
