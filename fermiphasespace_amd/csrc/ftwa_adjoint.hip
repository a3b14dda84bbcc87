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

}  // namespace fpsamd
