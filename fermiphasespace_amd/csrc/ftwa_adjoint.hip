// ftwa_adjoint.hip -- discrete adjoint of the fTWA RK4 "3/8" integration, with checkpointing.
//
// Objective: the ensemble double occupancy at the last time point,
//   J = (1/B) sum_b J_b,   J_b = (1/L) sum_i (rho_up,ii + 1/2)(rho_dn,ii + 1/2)
// (the site mean of NuNd, src/FHfTWASolver.cpp:395-398 and :524-531).  Parameters: the on-site U_ii,
// the hopping entries tM_kl on the hopping pattern (and the diagonal tM_kk), and the pulse amplitude
// a(tau) at every RK stage time (the host maps it onto knots or the Gaussian amplitude).
//
// Forward model, per trajectory and spin (the reference's derivative, src/FHfTWASolver.cpp:568-600):
//   f_s(rho) = -i [rho_s, Delta_s],  Delta_s = T(a) + diag(h_s),  h_s,i = pdiag_i + U_ii (Re rho_s',ii + 1/2)
//   T(a)_kl = tM_kl (1 + e^{i a pd_kl}) with a pulse, tM_kl without (k != l);  s' the other spin
// integrated with the textbook 3/8 tableau (src/FHfTWASolver.cpp:629-661):
//   Y1 = y, Y2 = y + h k1/3, Y3 = y + h (k2 - k1/3), Y4 = y + h (k1 - k2 + k3),
//   y' = y + h (k1 + 3 k2 + 3 k3 + k4)/8.
//
// Adjoint (real inner product <a, b> = Re sum conj(a_ij) b_ij): for a cotangent kappa of f_s,
//   rho_bar_s   += -i [Delta_s, kappa_s]
//   rho_bar_s',kk (real part) += U_kk G_s,k,      G_s,k = Re Q_s,kk
//   dJ/dtheta   += Re tr(Q_s dDelta_s/dtheta),    Q_s = i (rho_s kappa_s^dagger - kappa_s^dagger rho_s)
// and the step is reversed stage by stage:
//   kbar_s = h b_s lambda;  stage 4..1: Ybar = VJP(Y_s, kbar_s); lambda += Ybar; kbar_r -= ... (tableau)
// Checkpoints: the state every K steps is kept in HBM; each segment is recomputed forward (storing its
// K states) and then reversed.  One workgroup per trajectory; per-trajectory gradients in FP64, summed
// over trajectories in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::V2;
using dev::wave_sum;

constexpr int kAdjThreads = 256;

template <typename Real>
__device__ __forceinline__ V2<Real> cmul(V2<Real> a, V2<Real> b)
{
    return V2<Real>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename Real>
__device__ __forceinline__ V2<Real> cmulc(V2<Real> a, V2<Real> b)  // a * conj(b)
{
    return V2<Real>{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename Real>
__device__ __forceinline__ V2<Real> ccmul(V2<Real> a, V2<Real> b)  // conj(a) * b
{
    return V2<Real>{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// Per-trajectory work area, in complex elements of NE = 2 L^2 (both spins, row-major).
template <typename Real>
struct Work {
    V2<Real>* Y[4];  // stage inputs
    V2<Real>* K[4];  // stage slopes (forward) / their cotangents (reverse)
    V2<Real>* lam;   // cotangent of the state
    V2<Real>* tmp;   // VJP output
};
template <typename Real>
__device__ Work<Real> work_of(V2<Real>* base, int64_t b, int NE)
{
    V2<Real>* p = base + b * 10 * static_cast<int64_t>(NE);
    Work<Real> w;
    for (int s = 0; s < 4; ++s) {
        w.Y[s] = p + s * NE;
        w.K[s] = p + (4 + s) * NE;
    }
    w.lam = p + 8 * NE;
    w.tmp = p + 9 * NE;
    return w;
}

// hopping T_ik of the n-th neighbour of site i: from the stage's table Tst [z][L] with a pulse,
// tM_ik without
template <typename Real>
__device__ __forceinline__ V2<Real> hop(const DevSystem<Real>& s, const V2<Real>* Tst, int n, int i)
{
    if (Tst) return Tst[n * s.L + i];
    return V2<Real>{s.nbr_t[n * s.L + i], Real(0)};
}

// h of both spins from the stage input Y, into sh[2L]
template <typename Real>
__device__ void hartree(const DevSystem<Real>& s, const V2<Real>* Y, Real* sh)
{
    const int L = s.L, LL = L * L;
    for (int e = threadIdx.x; e < 2 * L; e += blockDim.x) {
        const int sg = e >= L ? 1 : 0, i = e - sg * L;
        Real v = s.pdiag[i];
        if (s.coulomb_type == 1) v += s.u_diag[i] * (Y[(1 - sg) * LL + i * L + i].x + Real(0.5));
        sh[e] = v;
    }
    __syncthreads();
}

// out = alpha [Delta, X] (alpha = +i: the derivative -i [X, Delta]; alpha = -i: the adjoint of its
// linear part).
// ACC adds into out instead of overwriting.
template <typename Real, bool PLUS_I, bool ACC>
__device__ void commutator(const DevSystem<Real>& s, const V2<Real>* Tst, const V2<Real>* X, const Real* sh,
                           V2<Real>* out)
{
    const int L = s.L, LL = L * L, NE = 2 * LL, z = s.z;
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        const int sg = e >= LL ? 1 : 0, ij = e - sg * LL, i = ij / L, j = ij - i * L;
        const V2<Real>* Xs = X + sg * LL;
        V2<Real> acc{Real(0), Real(0)};
        for (int n = 0; n < z; ++n) {  // (Delta X)_ij
            const V2<Real> T = hop(s, Tst, n, i);
            const V2<Real> v = cmul<Real>(T, Xs[s.nbr_k[n * L + i] * L + j]);
            acc.x += v.x;
            acc.y += v.y;
        }
        for (int n = 0; n < z; ++n) {  // (X Delta)_ij, T_kj = conj(T_jk)
            const V2<Real> T = hop(s, Tst, n, j);
            const V2<Real> v = cmulc<Real>(Xs[i * L + s.nbr_k[n * L + j]], T);
            acc.x -= v.x;
            acc.y -= v.y;
        }
        const Real dh = sh[sg * L + i] - sh[sg * L + j];
        acc.x += dh * Xs[ij].x;
        acc.y += dh * Xs[ij].y;
        const V2<Real> r = PLUS_I ? V2<Real>{-acc.y, acc.x} : V2<Real>{acc.y, -acc.x};
        if (ACC) {
            out[e].x += r.x;
            out[e].y += r.y;
        } else {
            out[e] = r;
        }
    }
    __syncthreads();
}

// dst = y + c1 k1 + c2 k2 + c3 k3 (null slopes skipped)
template <typename Real>
__device__ void combine(int NE, V2<Real>* dst, const V2<Real>* y, Real c1, const V2<Real>* k1, Real c2,
                        const V2<Real>* k2, Real c3, const V2<Real>* k3)
{
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        V2<Real> v = y[e];
        v.x += c1 * k1[e].x;
        v.y += c1 * k1[e].y;
        if (k2) {
            v.x += c2 * k2[e].x;
            v.y += c2 * k2[e].y;
        }
        if (k3) {
            v.x += c3 * k3[e].x;
            v.y += c3 * k3[e].y;
        }
        dst[e] = v;
    }
    __syncthreads();
}

// Stage inputs Y1..Y4 and slopes k1..k3 of the step from y (k4 too when want_k4).
template <typename Real>
__device__ void stages(const DevSystem<Real>& s, const V2<Real>* ttab, const RunShape& r, int64_t g, const V2<Real>* y,
                       Work<Real>& w, Real* sh, bool want_k4)
{
    const int NE = 2 * s.L * s.L;
    const Real h = static_cast<Real>(r.h);
    auto tst = [&](int st) { return ttab ? ttab + (g * 4 + st) * s.z * s.L : nullptr; };
    for (int e = threadIdx.x; e < NE; e += blockDim.x) w.Y[0][e] = y[e];
    __syncthreads();
    hartree(s, w.Y[0], sh);
    commutator<Real, true, false>(s, tst(0), w.Y[0], sh, w.K[0]);
    combine<Real>(NE, w.Y[1], y, h / Real(3), w.K[0], Real(0), nullptr, Real(0), nullptr);
    hartree(s, w.Y[1], sh);
    commutator<Real, true, false>(s, tst(1), w.Y[1], sh, w.K[1]);
    combine<Real>(NE, w.Y[2], y, -h / Real(3), w.K[0], h, w.K[1], Real(0), nullptr);
    hartree(s, w.Y[2], sh);
    commutator<Real, true, false>(s, tst(2), w.Y[2], sh, w.K[2]);
    combine<Real>(NE, w.Y[3], y, h, w.K[0], -h, w.K[1], h, w.K[2]);
    if (want_k4) {
        hartree(s, w.Y[3], sh);
        commutator<Real, true, false>(s, tst(3), w.Y[3], sh, w.K[3]);
    }
}

// One forward step in place: y <- y + h (k1 + 3 k2 + 3 k3 + k4) / 8.
template <typename Real>
__device__ void step_forward(const DevSystem<Real>& s, const V2<Real>* ttab, const RunShape& r, int64_t g,
                             V2<Real>* y, Work<Real>& w, Real* sh)
{
    const int NE = 2 * s.L * s.L;
    stages(s, ttab, r, g, y, w, sh, true);
    const Real h8 = static_cast<Real>(r.h / 8.0), h38 = static_cast<Real>(3.0 * r.h / 8.0);
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        V2<Real> v = y[e];
        v.x += h8 * (w.K[0][e].x + w.K[3][e].x) + h38 * (w.K[1][e].x + w.K[2][e].x);
        v.y += h8 * (w.K[0][e].y + w.K[3][e].y) + h38 * (w.K[1][e].y + w.K[2][e].y);
        y[e] = v;
    }
    __syncthreads();
}

// block-wide sum of one double per thread (fixed order), returned on every thread
__device__ double block_sum(double v, double* red)
{
    v = wave_sum(v);
    const int nw = (blockDim.x + 63) / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];
    __syncthreads();
    return t;
}

// Gradient accumulator layout per trajectory (FP64): [U_ii (L)] [tM_kk (L)] [tM pattern (z L, entry-major)]
// [dJ/da per stage (n_stage), pulse only]
struct GradLayout {
    int L, z;
    int64_t n_stage;
    __host__ __device__ int64_t size() const { return 2 * L + static_cast<int64_t>(z) * L + n_stage; }
};

// VJP of f at stage input Y for cotangent kappa: w.tmp = rho_bar, gradient terms into gacc.
template <typename Real>
__device__ void vjp(const DevSystem<Real>& s, const V2<Real>* Tst, const V2<Real>* Y, const V2<Real>* kap,
                    V2<Real>* out, Real* sh, double* gsh, double* red, double* gacc, const GradLayout& gl,
                    int64_t stage_index)
{
    const int L = s.L, LL = L * L, z = s.z;
    hartree(s, Y, sh);
    const bool pulse = Tst != nullptr;
    commutator<Real, false, false>(s, Tst, kap, sh, out);
    // G_s,k = Re Q_s,kk = Im( sum_m conj(kap_mk) Y_mk - sum_m Y_km conj(kap_km) )
    for (int e = threadIdx.x; e < 2 * L; e += blockDim.x) {
        const int sg = e >= L ? 1 : 0, k = e - sg * L;
        const V2<Real>* Ys = Y + sg * LL;
        const V2<Real>* Ks = kap + sg * LL;
        double g = 0.0;
        for (int m = 0; m < L; ++m) {
            const V2<Real> p = cmulc<Real>(Ys[k * L + m], Ks[k * L + m]);
            const V2<Real> q = ccmul<Real>(Ks[m * L + k], Ys[m * L + k]);
            g += static_cast<double>(q.y) - static_cast<double>(p.y);
        }
        gsh[e] = g;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        const double gu = gsh[k], gd = gsh[L + k];
        // Hartree: h_up,k depends on Re rho_dn,kk and h_dn,k on Re rho_up,kk
        if (s.coulomb_type == 1) {
            const Real u = s.u_diag[k];
            out[LL + k * L + k].x += static_cast<Real>(u * gu);
            out[k * L + k].x += static_cast<Real>(u * gd);
            gacc[k] += (static_cast<double>(Y[LL + k * L + k].x) + 0.5) * gu +
                       (static_cast<double>(Y[k * L + k].x) + 0.5) * gd;
        }
        gacc[L + k] += (gu + gd) * (pulse ? 2.0 : 1.0);  // Delta_kk holds tM_kk twice with a pulse
    }
    // hopping pattern: Q_lk = i S,  S = sum_m (Y_lm conj(kap_km) - conj(kap_ml) Y_mk);  Re(Q F) = -Im(S F)
    double sa = 0.0;
    for (int e = threadIdx.x; e < z * L; e += blockDim.x) {
        const int n = e / L, k = e - n * L;
        const Real t = s.nbr_t[n * L + k];
        if (t == Real(0)) continue;
        const int l = s.nbr_k[n * L + k];
        double sn = 0.0, cs = 1.0;  // e^{i phi} = T / tM - 1
        if (pulse) {
            const V2<Real> T = Tst[n * L + k];
            cs = static_cast<double>(T.x) / static_cast<double>(t) - 1.0;
            sn = static_cast<double>(T.y) / static_cast<double>(t);
        }
        double gt = 0.0;
        for (int sg = 0; sg < 2; ++sg) {
            const V2<Real>* Ys = Y + sg * LL;
            const V2<Real>* Ks = kap + sg * LL;
            double sx = 0.0, sy = 0.0;
            for (int m = 0; m < L; ++m) {
                const V2<Real> p = cmulc<Real>(Ys[l * L + m], Ks[k * L + m]);
                const V2<Real> q = ccmul<Real>(Ks[m * L + l], Ys[m * L + k]);
                sx += static_cast<double>(p.x) - static_cast<double>(q.x);
                sy += static_cast<double>(p.y) - static_cast<double>(q.y);
            }
            // F = dDelta_kl/dtM_kl = 1 + e^{i phi} with a pulse, 1 without
            const double fx = pulse ? 1.0 + static_cast<double>(cs) : 1.0, fy = pulse ? static_cast<double>(sn) : 0.0;
            gt -= sx * fy + sy * fx;  // -Im(S F)
            if (pulse) {
                // dDelta_kl/da = tM_kl i pd_kl e^{i phi}
                const double pd = static_cast<double>(s.nbr_pd[n * L + k]);
                const double dx = -static_cast<double>(t) * pd * static_cast<double>(sn);
                const double dy = static_cast<double>(t) * pd * static_cast<double>(cs);
                sa -= sx * dy + sy * dx;
            }
        }
        gacc[2 * L + e] += gt;
    }
    if (pulse) {
        const double tot = block_sum(sa, red);
        if (threadIdx.x == 0) gacc[2 * L + static_cast<int64_t>(z) * L + stage_index] += tot;
    }
    __syncthreads();
}

template <typename Real>
__global__ void __launch_bounds__(kAdjThreads)
    adj_forward_kernel(DevSystem<Real> s, const V2<Real>* ttab, RunShape r, int64_t n0, int64_t n1, V2<Real>* state,
                       V2<Real>* work)
{
    __shared__ Real sh[2 * 256];
    const int NE = 2 * s.L * s.L;
    const int64_t b = blockIdx.x;
    V2<Real>* y = state + b * NE;
    Work<Real> w = work_of<Real>(work, b, NE);
    for (int64_t g = n0; g < n1; ++g) step_forward(s, ttab, r, g, y, w, sh);
}

// J_b and the terminal cotangent lambda = dJ_b/d rho at the last time point
template <typename Real>
__global__ void __launch_bounds__(kAdjThreads)
    adj_terminal_kernel(DevSystem<Real> s, const V2<Real>* state, V2<Real>* work, double* jb)
{
    __shared__ double red[kAdjThreads / 64];
    const int L = s.L, LL = L * L, NE = 2 * LL;
    const int64_t b = blockIdx.x;
    const V2<Real>* y = state + b * NE;
    Work<Real> w = work_of<Real>(work, b, NE);
    for (int e = threadIdx.x; e < NE; e += blockDim.x) w.lam[e] = V2<Real>{Real(0), Real(0)};
    __syncthreads();
    double part = 0.0;
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
        const double nu = static_cast<double>(y[k * L + k].x) + 0.5;
        const double nd = static_cast<double>(y[LL + k * L + k].x) + 0.5;
        part += nu * nd;
        w.lam[k * L + k].x = static_cast<Real>(nd / L);
        w.lam[LL + k * L + k].x = static_cast<Real>(nu / L);
    }
    const double tot = block_sum(part, red);
    if (threadIdx.x == 0) jb[b] = tot / L;
}

// One segment [n0, n1): recompute its states from the checkpoint, then reverse it.
template <typename Real>
__global__ void __launch_bounds__(kAdjThreads)
    adj_backward_kernel(DevSystem<Real> s, const V2<Real>* ttab, RunShape r, int64_t n0, int64_t n1, const V2<Real>* ckpt,
                        V2<Real>* segbuf, V2<Real>* work, double* gacc_all, GradLayout gl)
{
    __shared__ Real sh[2 * 256];
    __shared__ double gsh[2 * 256];
    __shared__ double red[kAdjThreads / 64];
    const int NE = 2 * s.L * s.L;
    const int64_t b = blockIdx.x;
    const int64_t K = n1 - n0;
    Work<Real> w = work_of<Real>(work, b, NE);
    double* gacc = gacc_all + b * gl.size();
    V2<Real>* seg = segbuf + b * (K + 1) * NE;
    // states y_n0 .. y_{n1-1} of the segment
    for (int e = threadIdx.x; e < NE; e += blockDim.x) seg[e] = ckpt[b * NE + e];
    __syncthreads();
    for (int64_t g = n0; g + 1 < n1; ++g) {
        V2<Real>* cur = seg + (g - n0) * NE;
        V2<Real>* nxt = cur + NE;
        for (int e = threadIdx.x; e < NE; e += blockDim.x) nxt[e] = cur[e];
        __syncthreads();
        step_forward(s, ttab, r, g, nxt, w, sh);
    }
    const Real h = static_cast<Real>(r.h);
    const Real h8 = static_cast<Real>(r.h / 8.0), h38 = static_cast<Real>(3.0 * r.h / 8.0);
    for (int64_t g = n1 - 1; g >= n0; --g) {
        const V2<Real>* yg = seg + (g - n0) * NE;
        stages(s, ttab, r, g, yg, w, sh, false);
        // kbar_s = h b_s lambda (the K buffers now hold cotangents)
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const V2<Real> l = w.lam[e];
            w.K[0][e] = V2<Real>{h8 * l.x, h8 * l.y};
            w.K[1][e] = V2<Real>{h38 * l.x, h38 * l.y};
            w.K[2][e] = V2<Real>{h38 * l.x, h38 * l.y};
            w.K[3][e] = V2<Real>{h8 * l.x, h8 * l.y};
        }
        __syncthreads();
        auto tst = [&](int st) { return ttab ? ttab + (g * 4 + st) * s.z * s.L : nullptr; };
        // stage 4: Y4 = y + h (k1 - k2 + k3)
        vjp(s, tst(3), w.Y[3], w.K[3], w.tmp, sh, gsh, red, gacc, gl, g * 4 + 3);
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const V2<Real> v = w.tmp[e];
            w.lam[e].x += v.x;
            w.lam[e].y += v.y;
            w.K[0][e].x += h * v.x;
            w.K[0][e].y += h * v.y;
            w.K[1][e].x -= h * v.x;
            w.K[1][e].y -= h * v.y;
            w.K[2][e].x += h * v.x;
            w.K[2][e].y += h * v.y;
        }
        __syncthreads();
        // stage 3: Y3 = y + h (k2 - k1/3)
        vjp(s, tst(2), w.Y[2], w.K[2], w.tmp, sh, gsh, red, gacc, gl, g * 4 + 2);
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const V2<Real> v = w.tmp[e];
            w.lam[e].x += v.x;
            w.lam[e].y += v.y;
            w.K[0][e].x -= h / Real(3) * v.x;
            w.K[0][e].y -= h / Real(3) * v.y;
            w.K[1][e].x += h * v.x;
            w.K[1][e].y += h * v.y;
        }
        __syncthreads();
        // stage 2: Y2 = y + h k1/3
        vjp(s, tst(1), w.Y[1], w.K[1], w.tmp, sh, gsh, red, gacc, gl, g * 4 + 1);
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const V2<Real> v = w.tmp[e];
            w.lam[e].x += v.x;
            w.lam[e].y += v.y;
            w.K[0][e].x += h / Real(3) * v.x;
            w.K[0][e].y += h / Real(3) * v.y;
        }
        __syncthreads();
        // stage 1: Y1 = y
        vjp(s, tst(0), w.Y[0], w.K[0], w.tmp, sh, gsh, red, gacc, gl, g * 4 + 0);
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            w.lam[e].x += w.tmp[e].x;
            w.lam[e].y += w.tmp[e].y;
        }
        __syncthreads();
    }
}

// Hopping with the pulse at every RK stage: ttab[q][n][i] = tM_ik (1 + e^{i a(tau_q) pd_ik}), k = nb_n(i).
template <typename Real>
__global__ void adj_table_kernel(DevSystem<Real> s, int64_t n_stage, V2<Real>* ttab)
{
    const int64_t zl = static_cast<int64_t>(s.z) * s.L;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n_stage * zl) return;
    const int64_t q = idx / zl, e = idx - q * zl;
    const double t = static_cast<double>(s.nbr_t[e]);
    double sn, cs;
    sincos(s.amp[q] * static_cast<double>(s.nbr_pd[e]), &sn, &cs);
    ttab[idx] = V2<Real>{static_cast<Real>(t * (1.0 + cs)), static_cast<Real>(t * sn)};
}

// Fixed-order sum over trajectories: out[q] = sum_b src[b][q] (Neumaier), for n quantities.
__global__ void adj_reduce_kernel(const double* src, int64_t B, int64_t n, double* out)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double sum = 0.0, c = 0.0;
    for (int64_t b = 0; b < B; ++b) {
        const double v = src[b * n + q];
        const double t = sum + v;
        c += (fabs(sum) >= fabs(v)) ? (sum - t) + v : (v - t) + sum;
        sum = t;
    }
    out[q] = sum + c;
}

}  // namespace

template <typename Real>
size_t adjoint_work_bytes(int L, int64_t B, int64_t seg_steps)
{
    const size_t NE = 2 * static_cast<size_t>(L) * L;
    return static_cast<size_t>(B) * NE * sizeof(V2<Real>) * (10 + static_cast<size_t>(seg_steps) + 1);
}

template <typename Real>
hipError_t launch_adjoint_table(const DevSystem<Real>& s, int64_t n_stage, Real* ttab, hipStream_t st)
{
    const int64_t n = n_stage * s.z * s.L;
    if (n <= 0 || !s.amp) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_table_kernel<Real>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, s,
                       n_stage, reinterpret_cast<V2<Real>*>(ttab));
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_adjoint_forward(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int64_t n0,
                                  int64_t n1, Real* state, Real* work, hipStream_t st)
{
    if (s.L > 256 || r.B <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_forward_kernel<Real>, dim3(static_cast<unsigned>(r.B)), dim3(kAdjThreads), 0, st, s,
                       reinterpret_cast<const V2<Real>*>(ttab), r, n0, n1, reinterpret_cast<V2<Real>*>(state),
                       reinterpret_cast<V2<Real>*>(work));
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_adjoint_terminal(const DevSystem<Real>& s, int64_t B, const Real* state, Real* work, double* jb,
                                   hipStream_t st)
{
    hipLaunchKernelGGL(adj_terminal_kernel<Real>, dim3(static_cast<unsigned>(B)), dim3(kAdjThreads), 0, st, s,
                       reinterpret_cast<const V2<Real>*>(state), reinterpret_cast<V2<Real>*>(work), jb);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_adjoint_backward(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int64_t n0,
                                   int64_t n1, const Real* ckpt, Real* segbuf, Real* work, double* gacc,
                                   int64_t n_stage, hipStream_t st)
{
    if (s.L > 256 || r.B <= 0 || n1 <= n0) return hipErrorInvalidValue;
    GradLayout gl{s.L, s.z, n_stage};
    hipLaunchKernelGGL(adj_backward_kernel<Real>, dim3(static_cast<unsigned>(r.B)), dim3(kAdjThreads), 0, st, s,
                       reinterpret_cast<const V2<Real>*>(ttab), r, n0, n1, reinterpret_cast<const V2<Real>*>(ckpt),
                       reinterpret_cast<V2<Real>*>(segbuf), reinterpret_cast<V2<Real>*>(work), gacc, gl);
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const double* src, int64_t B, int64_t n, double* out, hipStream_t st)
{
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(adj_reduce_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, st, src, B, n, out);
    return hipGetLastError();
}

template size_t adjoint_work_bytes<float>(int, int64_t, int64_t);
template size_t adjoint_work_bytes<double>(int, int64_t, int64_t);
template hipError_t launch_adjoint_table<float>(const DevSystem<float>&, int64_t, float*, hipStream_t);
template hipError_t launch_adjoint_table<double>(const DevSystem<double>&, int64_t, double*, hipStream_t);
template hipError_t launch_adjoint_forward<float>(const DevSystem<float>&, const float*, const RunShape&, int64_t,
                                                  int64_t, float*, float*, hipStream_t);
template hipError_t launch_adjoint_forward<double>(const DevSystem<double>&, const double*, const RunShape&, int64_t,
                                                   int64_t, double*, double*, hipStream_t);
template hipError_t launch_adjoint_terminal<float>(const DevSystem<float>&, int64_t, const float*, float*, double*,
                                                   hipStream_t);
template hipError_t launch_adjoint_terminal<double>(const DevSystem<double>&, int64_t, const double*, double*,
                                                    double*, hipStream_t);
template hipError_t launch_adjoint_backward<float>(const DevSystem<float>&, const float*, const RunShape&, int64_t,
                                                   int64_t, const float*, float*, float*, double*, int64_t,
                                                   hipStream_t);
template hipError_t launch_adjoint_backward<double>(const DevSystem<double>&, const double*, const RunShape&, int64_t,
                                                    int64_t, const double*, double*, double*, double*, int64_t,
                                                    hipStream_t);


// ==== LDS-resident adjoint ==============================================================================
//
// The same discrete adjoint as above, organised for the chip: one workgroup per trajectory, each thread
// owning E element slots of the upper triangle (i <= j) of both spins.  rho, the stage inputs and every
// cotangent stay Hermitian (f and its vector-Jacobian product map Hermitian matrices to Hermitian ones;
// the terminal cotangent is real diagonal), so the upper triangle carries them all.  The stage inputs,
// slopes and cotangents live in registers; only the operand of a commutator (and, in a vector-Jacobian
// product, the stage input beside the cotangent) is written to LDS, mirrored to a full [2][L][L] matrix.
// Two such LDS matrices double-buffer the forward stages (one barrier per stage).  Per-trajectory state
// between launches sits in HBM in slot order [E][NT] (coalesced per thread): the checkpoints, the states
// of the segment being reversed and the cotangent handed from one segment to the one before.
// The gradient accumulators are per item (site k: the U and diagonal-hopping terms; bond (n, k): the
// hopping entry) in FP64 registers, loaded from and stored to the per-trajectory gradient array around
// each segment, so the order of the additions does not depend on the checkpoint spacing.
namespace {

struct AdjLdsLayout {
    int buf[2];  // [2 spins][L][L] complex
    int h[2];    // [2 spins][L] Real, the Hartree fields of buf[k]
    int tab[2];  // [z][L] complex: hopping of the stage (pulse), or tM
    int nbk;     // [z][L] int: neighbour k_n(i)
    int gsh;     // [2][L] double: G_sigma,k of a vector-Jacobian product
    int red;     // [NT / 64] double: per-wave pulse partials
    int dg;      // [2][L] Real: diagonal exchange (terminal cotangent, objective)
    int total;
};

template <typename Real>
__host__ __device__ inline AdjLdsLayout adj_lds_layout(int L, int z, int nt)
{
    auto up = [](int b) { return (b + 15) / 16 * 16; };
    const int C = 2 * static_cast<int>(sizeof(Real));
    AdjLdsLayout a{};
    int o = 0;
    for (int k = 0; k < 2; ++k) {
        a.buf[k] = o;
        o += up(2 * L * L * C);
    }
    for (int k = 0; k < 2; ++k) {
        a.h[k] = o;
        o += up(2 * L * static_cast<int>(sizeof(Real)));
    }
    for (int k = 0; k < 2; ++k) {
        a.tab[k] = o;
        o += up(z * L * C);
    }
    a.nbk = o;
    o += up(z * L * 4);
    a.gsh = o;
    o += up(2 * L * 8);
    a.red = o;
    o += up(nt / 64 * 8);
    a.dg = o;
    o += up(2 * L * static_cast<int>(sizeof(Real)));
    a.total = o;
    return a;
}

struct AdjLdsArgs {
    int mode;          // 0: forward pass (checkpoints, objective); 1: one segment reversed
    int64_t n0, n1;    // steps of the segment (mode 1)
    int64_t nst;       // steps up to the last time point
    int64_t K;         // checkpoint spacing
    int64_t n_stage;   // pulse stages recorded in the gradient array (0 without a pulse)
    const void* state; // mode 0: initial states [B][2][L][L] complex
    void* ckpt;        // [nseg][B][E*NT] complex
    void* seg;         // [B][K][E*NT] complex
    void* lam;         // [B][E*NT] complex
    double* gacc;      // [B][2L + zL + n_stage]
    double* jb;        // [B]
};

template <typename Real, int NT, int E>
__global__ void __launch_bounds__(NT) adj_lds_kernel(DevSystem<Real> s, const V2<Real>* ttab, RunShape r, AdjLdsArgs a)
{
    using C = V2<Real>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = s.L, LL = L * L, z = s.z;
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    const AdjLdsLayout ly = adj_lds_layout<Real>(L, z, NT);
    const int P = L * (L + 1) / 2;
    const bool pulse = ttab != nullptr;
    const Real h = static_cast<Real>(r.h), h3 = static_cast<Real>(r.h / 3.0), h8 = static_cast<Real>(r.h / 8.0),
               h38 = static_cast<Real>(3.0 * r.h / 8.0);

    // ---- slots: packed upper-triangle index p of spin sg, slot q of this thread holds g = q*NT + tid ----
    int sl[E];  // (sg << 16) | (i << 8) | j, or -1
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int g = q * NT + tid;
        sl[q] = -1;
        if (g < 2 * P) {
            const int sg = g >= P ? 1 : 0;
            int p = g - sg * P, i = 0;
            while (p >= L - i) {  // row i holds L - i entries
                p -= L - i;
                ++i;
            }
            sl[q] = (sg << 16) | (i << 8) | (i + p);
        }
    }
    // the slot code is re-read opaquely at every use, so the addresses derived from it are formed there
    // and not kept live across the time loop (about 28 registers per slot otherwise)
    auto slot = [&](int q) {
        int v = sl[q];
        asm volatile("" : "+v"(v));
        return v;
    };
    auto sg_of = [](int v) { return v >> 16; };
    auto i_of = [](int v) { return (v >> 8) & 255; };
    auto j_of = [](int v) { return v & 255; };
    auto at = [&](int bf, int sg, int i, int j) -> C& {
        return *reinterpret_cast<C*>(smem + ly.buf[bf] + ((sg * L + i) * L + j) * static_cast<int>(sizeof(C)));
    };
    Real* hh[2] = {reinterpret_cast<Real*>(smem + ly.h[0]), reinterpret_cast<Real*>(smem + ly.h[1])};
    C* tb[2] = {reinterpret_cast<C*>(smem + ly.tab[0]), reinterpret_cast<C*>(smem + ly.tab[1])};
    int* nbk = reinterpret_cast<int*>(smem + ly.nbk);
    double* gsh = reinterpret_cast<double*>(smem + ly.gsh);
    double* red = reinterpret_cast<double*>(smem + ly.red);
    Real* dg = reinterpret_cast<Real*>(smem + ly.dg);

    for (int e = tid; e < z * L; e += NT) {
        nbk[e] = s.nbr_k[e];
        if (!pulse) tb[0][e] = tb[1][e] = C{s.nbr_t[e], Real(0)};
    }
    __syncthreads();

    // hopping table of RK stage index q into tab[k] (pulse only; without one the table is tM throughout)
    auto load_tab = [&](int k, int64_t q) {
        if (!pulse) return;
        const C* src = ttab + q * z * L;
        for (int e = tid; e < z * L; e += NT) tb[k][e] = src[e];
    };
    // v -> LDS matrix bf (mirrored) and, from the diagonal slots, the Hartree fields of bf
    auto put = [&](int bf, const C* v) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            if (sl[q] < 0) continue;
            const int sv = slot(q), sg = sg_of(sv), i = i_of(sv), j = j_of(sv);
            at(bf, sg, i, j) = v[q];
            if (i != j) {
                at(bf, sg, j, i) = C{v[q].x, -v[q].y};
            } else {  // h_{1-sg},i = pdiag_i + U_ii (Re rho_sg,ii + 1/2)
                Real hv = s.pdiag[i];
                if (s.coulomb_type == 1) hv += s.u_diag[i] * (v[q].x + Real(0.5));
                hh[bf][(1 - sg) * L + i] = hv;
            }
        }
    };
    // acc = (Delta X - X Delta)_ij for the slots, X = buf[bf] (own value from v), hopping tab[k]
    auto comm = [&](int bf, int k, const C* v, C* out, bool plus_i) {
        const C* T = tb[k];
        const Real* hs = hh[bf];
#pragma unroll
        for (int q = 0; q < E; ++q) {
            if (sl[q] < 0) continue;
            const int sv = slot(q), sg = sg_of(sv), i = i_of(sv), j = j_of(sv);
            Real ax = 0, ay = 0;
            for (int n = 0; n < z; ++n) {
                const C t = T[n * L + i];
                const C x = at(bf, sg, nbk[n * L + i], j);
                ax += t.x * x.x - t.y * x.y;
                ay += t.x * x.y + t.y * x.x;
            }
            for (int n = 0; n < z; ++n) {  // X_ik T_kj with T_kj = conj(T_jk)
                const C t = T[n * L + j];
                const C x = at(bf, sg, i, nbk[n * L + j]);
                ax -= x.x * t.x + x.y * t.y;
                ay -= x.y * t.x - x.x * t.y;
            }
            const Real dh = hs[sg * L + i] - hs[sg * L + j];
            ax += dh * v[q].x;
            ay += dh * v[q].y;
            out[q] = plus_i ? C{-ay, ax} : C{ay, -ax};
        }
    };

    // one forward step from y in place, stage inputs Y1..Y4 into Ys when given; q0 = 4 g
    C k1[E], k2[E], k3[E], tmp[E];
    auto stages = [&](const C* y, C (*Ys)[E], int64_t q0, bool want_y4, C* k4) {
        // Y1 = y
        load_tab(0, q0);
        put(0, y);
        __syncthreads();
        comm(0, 0, y, k1, true);
#pragma unroll
        for (int q = 0; q < E; ++q) tmp[q] = C{y[q].x + h3 * k1[q].x, y[q].y + h3 * k1[q].y};
        if (Ys) {
#pragma unroll
            for (int q = 0; q < E; ++q) Ys[1][q] = tmp[q];
        }
        load_tab(1, q0 + 1);
        put(1, tmp);
        __syncthreads();
        comm(1, 1, tmp, k2, true);
#pragma unroll
        for (int q = 0; q < E; ++q)
            tmp[q] = C{y[q].x - h3 * k1[q].x + h * k2[q].x, y[q].y - h3 * k1[q].y + h * k2[q].y};
        if (Ys) {
#pragma unroll
            for (int q = 0; q < E; ++q) Ys[2][q] = tmp[q];
        }
        load_tab(0, q0 + 2);
        put(0, tmp);
        __syncthreads();
        comm(0, 0, tmp, k3, true);
#pragma unroll
        for (int q = 0; q < E; ++q)
            tmp[q] = C{y[q].x + h * k1[q].x - h * k2[q].x + h * k3[q].x, y[q].y + h * k1[q].y - h * k2[q].y + h * k3[q].y};
        if (Ys) {
#pragma unroll
            for (int q = 0; q < E; ++q) Ys[3][q] = tmp[q];
        }
        if (want_y4) {
            load_tab(1, q0 + 3);
            put(1, tmp);
            __syncthreads();
            comm(1, 1, tmp, k4, true);
        }
    };
    auto step = [&](C* y, int64_t g) {
        C k4[E];
        stages(y, nullptr, 4 * g, true, k4);
#pragma unroll
        for (int q = 0; q < E; ++q) {
            y[q].x += h8 * (k1[q].x + k4[q].x) + h38 * (k2[q].x + k3[q].x);
            y[q].y += h8 * (k1[q].y + k4[q].y) + h38 * (k2[q].y + k3[q].y);
        }
    };

    // per-trajectory gradient items: site k (item < L) and bond (n, k) (item = L + n L + k), each summed by
    // a group of kIG lanes; item p * (NT / kIG) + tid / kIG is this thread's p-th, its partial sums combined
    // across the group in a fixed order and kept by the group's first lane
    constexpr int kIG = 4, kIP = 1024 / NT;
    const int n_item = L + z * L;
    const int grp = tid / kIG, part = tid % kIG;
    double acc_a[kIP], acc_b[kIP];  // site: U_kk and tM_kk terms; bond: the entry's term (group leaders)
    double* ga = a.gacc + b * (2 * static_cast<int64_t>(L) + static_cast<int64_t>(z) * L + a.n_stage);
    C y[E];
    const int SL = E * NT;
    auto slot_ptr = [&](void* base, int64_t idx) { return reinterpret_cast<C*>(base) + idx * SL; };

    if (a.mode == 0) {
        // ---- forward pass: states every K steps, the objective at the end ----
        const C* gs = reinterpret_cast<const C*>(a.state) + b * 2 * LL;
#pragma unroll
        for (int q = 0; q < E; ++q) {
            y[q] = C{0, 0};
            if (sl[q] >= 0) {
                const int sv = slot(q);
                y[q] = gs[sg_of(sv) * LL + i_of(sv) * L + j_of(sv)];
            }
        }
        for (int64_t g = 0; g < a.nst; ++g) {
            if (g % a.K == 0) {
                C* ck = slot_ptr(a.ckpt, (g / a.K) * r.B + b);
#pragma unroll
                for (int q = 0; q < E; ++q) ck[q * NT + tid] = y[q];
            }
            step(y, g);
        }
        // J_b = (1/L) sum_k (rho_up,kk + 1/2)(rho_dn,kk + 1/2)
#pragma unroll
        for (int q = 0; q < E; ++q)
            if (sl[q] >= 0 && i_of(slot(q)) == j_of(slot(q))) dg[sg_of(slot(q)) * L + i_of(slot(q))] = y[q].x;
        __syncthreads();
        double part = 0.0;
        for (int k = tid; k < L; k += NT)
            part += (static_cast<double>(dg[k]) + 0.5) * (static_cast<double>(dg[L + k]) + 0.5);
        const double tot = block_sum(part, red);
        if (tid == 0) a.jb[b] = tot / L;
        return;
    }

    // ---- one segment [n0, n1): recompute its states from the checkpoint, then reverse it ----
    {
        const C* ck = slot_ptr(a.ckpt, (a.n0 / a.K) * r.B + b);
#pragma unroll
        for (int q = 0; q < E; ++q) y[q] = ck[q * NT + tid];
    }
    for (int64_t g = a.n0; g < a.n1; ++g) {
        C* sp = slot_ptr(a.seg, b * a.K + (g - a.n0));
#pragma unroll
        for (int q = 0; q < E; ++q) sp[q * NT + tid] = y[q];
        step(y, g);
    }
    C lam[E];
    if (a.n1 == a.nst) {  // terminal cotangent dJ_b/d rho: lambda_up,kk = n_dn,k / L, lambda_dn,kk = n_up,k / L
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q)
            if (sl[q] >= 0 && i_of(slot(q)) == j_of(slot(q))) dg[sg_of(slot(q)) * L + i_of(slot(q))] = y[q].x;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            lam[q] = C{0, 0};
            if (sl[q] >= 0 && i_of(slot(q)) == j_of(slot(q))) {
                const int sv = slot(q), sg = sg_of(sv), k = i_of(sv);
                lam[q].x = static_cast<Real>((static_cast<double>(dg[(1 - sg) * L + k]) + 0.5) / L);
            }
        }
    } else {
        const C* lp = slot_ptr(a.lam, b);
#pragma unroll
        for (int q = 0; q < E; ++q) lam[q] = lp[q * NT + tid];
    }
#pragma unroll
    for (int p = 0; p < kIP; ++p) {
        const int it = p * (NT / kIG) + grp;
        acc_a[p] = acc_b[p] = 0.0;
        if (part == 0 && it < L) {
            acc_a[p] = ga[it];
            acc_b[p] = ga[L + it];
        } else if (part == 0 && it < n_item) {
            acc_a[p] = ga[2 * L + (it - L)];
        }
    }

    // vector-Jacobian product of f at Ys for the cotangent kap (stage index q): out = -i [Delta, kap] plus
    // the Hartree term; the gradient items gain their terms of Re tr(Q dDelta/dtheta)
    auto vjp = [&](const C* Ysv, const C* kap, C* out, int64_t qst) {
        load_tab(0, qst);
        put(0, Ysv);  // h of the stage input into h[0]
        put(1, kap);
        __syncthreads();
        {
            // commutator of kap (buffer 1) with Delta (tab[0], h of buffer 0)
            const C* T = tb[0];
            const Real* hs = hh[0];
#pragma unroll
            for (int q = 0; q < E; ++q) {
                if (sl[q] < 0) continue;
                const int sv = slot(q), sg = sg_of(sv), i = i_of(sv), j = j_of(sv);
                Real ax = 0, ay = 0;
                for (int n = 0; n < z; ++n) {
                    const C t = T[n * L + i];
                    const C x = at(1, sg, nbk[n * L + i], j);
                    ax += t.x * x.x - t.y * x.y;
                    ay += t.x * x.y + t.y * x.x;
                }
                for (int n = 0; n < z; ++n) {
                    const C t = T[n * L + j];
                    const C x = at(1, sg, i, nbk[n * L + j]);
                    ax -= x.x * t.x + x.y * t.y;
                    ay -= x.y * t.x - x.x * t.y;
                }
                const Real dh = hs[sg * L + i] - hs[sg * L + j];
                ax += dh * kap[q].x;
                ay += dh * kap[q].y;
                out[q] = C{ay, -ax};
            }
        }
        double sa = 0.0;
        // sum over the group's lanes, fixed order; the first lane's value is the item's
        auto gsum = [](double v) {
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            return v;
        };
#pragma unroll
        for (int p = 0; p < kIP; ++p) {
            int it = p * (NT / kIG) + grp;  // re-read opaquely: addresses formed here, not held across calls
            asm volatile("" : "+v"(it));
            if (it < L) {
                // G_sg,k = Re Q_sg,kk = Im(sum_m conj(kap_mk) Y_mk - Y_km conj(kap_km)) = -2 sum_{m != k}
                // Im(Y_km conj(kap_km)) for Hermitian Y and kap (the m = k terms cancel)
                const int k = it;
                double gg[2];
                for (int sg = 0; sg < 2; ++sg) {
                    double g = 0.0;
#pragma unroll 2
                    for (int m = part; m < L; m += kIG) {
                        if (m == k) continue;
                        const C yk = at(0, sg, k, m), kk = at(1, sg, k, m);
                        g += static_cast<double>(yk.y * kk.x - yk.x * kk.y);
                    }
                    gg[sg] = -2.0 * gsum(g);
                }
                if (part == 0) {
                    gsh[k] = gg[0];
                    gsh[L + k] = gg[1];
                    if (s.coulomb_type == 1)
                        acc_a[p] += (static_cast<double>(at(0, 1, k, k).x) + 0.5) * gg[0] +
                                    (static_cast<double>(at(0, 0, k, k).x) + 0.5) * gg[1];
                    acc_b[p] += (gg[0] + gg[1]) * (pulse ? 2.0 : 1.0);
                }
            } else if (it < n_item) {
                // bond (n, k), l = k_n(k): Q_lk = i S, S = sum_m (Y_lm conj(kap_km) - conj(kap_ml) Y_mk)
                //                                          = sum_m (Y_lm conj(kap_km) - kap_lm conj(Y_km))
                const int e = it - L, n = e / L, k = e - n * L;
                (void)n;
                const double t = static_cast<double>(s.nbr_t[e]);
                if (t != 0.0) {
                    const int l = nbk[e];
                    double sn = 0.0, cs = 1.0;  // e^{i phi} = T / tM - 1
                    if (pulse) {
                        const C T = tb[0][e];
                        cs = static_cast<double>(T.x) / t - 1.0;
                        sn = static_cast<double>(T.y) / t;
                    }
                    const double fx = pulse ? 1.0 + cs : 1.0, fy = pulse ? sn : 0.0;
                    double gt = 0.0;
                    for (int sg = 0; sg < 2; ++sg) {
                        double sx = 0.0, sy = 0.0;
#pragma unroll 2
                        for (int m = part; m < L; m += kIG) {
                            const C yl = at(0, sg, l, m), kk = at(1, sg, k, m);
                            const C kl = at(1, sg, l, m), yk = at(0, sg, k, m);
                            const Real px = yl.x * kk.x + yl.y * kk.y, py = yl.y * kk.x - yl.x * kk.y;  // Y_lm conj(kap_km)
                            const Real qx = kl.x * yk.x + kl.y * yk.y, qy = kl.y * yk.x - kl.x * yk.y;  // kap_lm conj(Y_km)
                            sx += static_cast<double>(px) - static_cast<double>(qx);
                            sy += static_cast<double>(py) - static_cast<double>(qy);
                        }
                        sx = gsum(sx);
                        sy = gsum(sy);
                        gt -= sx * fy + sy * fx;  // -Im(S F), F = dDelta_kl / dtM_kl
                        if (pulse) {  // dDelta_kl/da = tM_kl i pd_kl e^{i phi}
                            const double pd = static_cast<double>(s.nbr_pd[e]);
                            const double dx = -t * pd * sn, dy = t * pd * cs;
                            if (part == 0) sa -= sx * dy + sy * dx;
                        }
                    }
                    if (part == 0) acc_a[p] += gt;
                }
            }
        }
        if (pulse) {
            sa = wave_sum(sa);
            if ((tid & 63) == 0) red[tid >> 6] = sa;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) {  // Hartree: h_up,k depends on Re rho_dn,kk, h_dn,k on Re rho_up,kk
            if (sl[q] < 0 || s.coulomb_type != 1) continue;
            const int sv = slot(q), sg = sg_of(sv), k = i_of(sv);
            if (k == j_of(sv)) out[q].x += static_cast<Real>(static_cast<double>(s.u_diag[k]) * gsh[(1 - sg) * L + k]);
        }
        if (pulse && tid == 0) {
            double tot = 0.0;
            for (int w = 0; w < NT / 64; ++w) tot += red[w];
            ga[2 * L + static_cast<int64_t>(z) * L + qst] += tot;
        }
    };

    C Ys[4][E];
    C c1[E], c2[E], c3[E], yb[E];
    for (int64_t g = a.n1 - 1; g >= a.n0; --g) {
        const C* sp = slot_ptr(a.seg, b * a.K + (g - a.n0));
#pragma unroll
        for (int q = 0; q < E; ++q) Ys[0][q] = sp[q * NT + tid];
        stages(Ys[0], Ys, 4 * g, false, nullptr);
        __syncthreads();  // the last stage's readers are done with the buffers the products write
        // stage 4: kbar4 = h/8 lambda
#pragma unroll
        for (int q = 0; q < E; ++q) c3[q] = C{h8 * lam[q].x, h8 * lam[q].y};
        vjp(Ys[3], c3, yb, 4 * g + 3);
#pragma unroll
        for (int q = 0; q < E; ++q) {  // kbar3 = 3h/8 lambda + h Yb4, kbar2 = 3h/8 lambda - h Yb4, kbar1 = h/8 lambda + h Yb4
            c3[q] = C{h38 * lam[q].x + h * yb[q].x, h38 * lam[q].y + h * yb[q].y};
            c2[q] = C{h38 * lam[q].x - h * yb[q].x, h38 * lam[q].y - h * yb[q].y};
            c1[q] = C{h8 * lam[q].x + h * yb[q].x, h8 * lam[q].y + h * yb[q].y};
            lam[q].x += yb[q].x;
            lam[q].y += yb[q].y;
        }
        vjp(Ys[2], c3, yb, 4 * g + 2);
#pragma unroll
        for (int q = 0; q < E; ++q) {  // Y3 = y + h (k2 - k1/3)
            c2[q].x += h * yb[q].x;
            c2[q].y += h * yb[q].y;
            c1[q].x -= h3 * yb[q].x;
            c1[q].y -= h3 * yb[q].y;
            lam[q].x += yb[q].x;
            lam[q].y += yb[q].y;
        }
        vjp(Ys[1], c2, yb, 4 * g + 1);
#pragma unroll
        for (int q = 0; q < E; ++q) {  // Y2 = y + h k1/3
            c1[q].x += h3 * yb[q].x;
            c1[q].y += h3 * yb[q].y;
            lam[q].x += yb[q].x;
            lam[q].y += yb[q].y;
        }
        vjp(Ys[0], c1, yb, 4 * g);
#pragma unroll
        for (int q = 0; q < E; ++q) {
            lam[q].x += yb[q].x;
            lam[q].y += yb[q].y;
        }
    }
    {
        C* lp = slot_ptr(a.lam, b);
#pragma unroll
        for (int q = 0; q < E; ++q) lp[q * NT + tid] = lam[q];
    }
#pragma unroll
    for (int p = 0; p < kIP; ++p) {
        const int it = p * (NT / kIG) + grp;
        if (part == 0 && it < L) {
            ga[it] = acc_a[p];
            ga[L + it] = acc_b[p];
        } else if (part == 0 && it < n_item) {
            ga[2 * L + (it - L)] = acc_a[p];
        }
    }
}

template <typename Real, int NT, int E>
hipError_t launch_adj_lds_t(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, const AdjLdsArgs& a,
                            hipStream_t st)
{
    const AdjLdsLayout ly = adj_lds_layout<Real>(s.L, s.z, NT);
    auto kern = adj_lds_kernel<Real, NT, E>;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(r.B)), dim3(NT), ly.total, st, s,
                       reinterpret_cast<const V2<Real>*>(ttab), r, a);
    return hipGetLastError();
}

}  // namespace

// Threads per workgroup and slots per thread of the LDS-resident adjoint for L sites, or 0 when it does
// not apply (the HBM form above then runs): both spins' upper triangles in at most 8 slots per thread of
// 256 / 512 / 1024 threads, the site and bond items within the threads, two matrices in LDS.
template <typename Real>
int adjoint_lds_threads(int L, int z, int* slots, int want_nt)
{
    // The smallest workgroup that holds both spins' upper triangles in at most 2 slots per thread (about
    // 40 registers per FP32 slot in the reverse step; on the 6x6 workload 1,024 threads with 2 slots run
    // 5.0e6 trajectory-steps/s against 3.8e6 for 512 threads with 3); FP32 takes 3 slots when 2 do not
    // fit.  The gradient items (site and bond terms, 4 lanes each) must fit 256 per workgroup.
    const int P2 = L * (L + 1);
    if (L > 255 || L * (1 + z) > 256) return 0;
    const int emax = sizeof(Real) == 4 ? 3 : 2;
    auto fits = [&](int nt, int e) {
        return e <= emax && (P2 + nt - 1) / nt <= e && adj_lds_layout<Real>(L, z, nt).total <= 160 * 1024;
    };
    const int tries[4][2] = {{256, 2}, {512, 2}, {1024, 2}, {512, 3}};
    for (const auto& t : tries) {
        const int nt = t[0];
        if (want_nt && nt != want_nt) continue;
        const int e = std::max(1, (P2 + nt - 1) / nt);
        if (e > t[1] || !fits(nt, e)) continue;
        if (slots) *slots = e;
        return nt;
    }
    return 0;
}

template <typename Real>
hipError_t launch_adjoint_lds(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int want_nt, int mode,
                              int64_t n0, int64_t n1, int64_t nst, int64_t K, int64_t n_stage, const Real* state,
                              Real* ckpt, Real* seg, Real* lam, double* gacc, double* jb, hipStream_t st)
{
    int e = 0;
    const int nt = adjoint_lds_threads<Real>(s.L, s.z, &e, want_nt);
    if (!nt || r.B <= 0) return hipErrorInvalidValue;
    AdjLdsArgs a{mode, n0, n1, nst, K, n_stage, state, ckpt, seg, lam, gacc, jb};
#define FPS_ADJ_CASE(NT_, E_) \
    if (nt == NT_ && e == E_) return launch_adj_lds_t<Real, NT_, E_>(s, ttab, r, a, st);
    if constexpr (sizeof(Real) == 4) {
        FPS_ADJ_CASE(256, 1) FPS_ADJ_CASE(256, 2) FPS_ADJ_CASE(256, 3) FPS_ADJ_CASE(512, 2) FPS_ADJ_CASE(512, 3)
        FPS_ADJ_CASE(1024, 2) FPS_ADJ_CASE(1024, 3)
    } else {
        FPS_ADJ_CASE(256, 1) FPS_ADJ_CASE(256, 2) FPS_ADJ_CASE(512, 2) FPS_ADJ_CASE(1024, 2)
    }
#undef FPS_ADJ_CASE
    return hipErrorInvalidValue;
}

template int adjoint_lds_threads<float>(int, int, int*, int);
template int adjoint_lds_threads<double>(int, int, int*, int);
template hipError_t launch_adjoint_lds<float>(const DevSystem<float>&, const float*, const RunShape&, int, int, int64_t,
                                              int64_t, int64_t, int64_t, int64_t, const float*, float*, float*, float*,
                                              double*, double*, hipStream_t);
template hipError_t launch_adjoint_lds<double>(const DevSystem<double>&, const double*, const RunShape&, int, int,
                                               int64_t, int64_t, int64_t, int64_t, int64_t, const double*, double*,
                                               double*, double*, double*, double*, hipStream_t);

}  // namespace fpsamd
