/*
 * ftwa_oracle.c -- CPU FP64 restatement of the reference fTWA path.
 *
 * TEST INFRASTRUCTURE ONLY (see ftwa_oracle.h).  The product never links this file.
 *
 * Citations are to the reference FermiPhaseSpace sources.  The reference itself cannot be
 * built in this container (it needs Armadillo/BLAS/LAPACK), so this is a restatement:
 * dense Delta, rho*Delta - Delta*rho as plain loops, the RK4 "3/8 rule" tableau, the
 * reference's observable formulas and normalisation.  Operation order is fixed and FP64
 * throughout; build with -ffp-contract=off so the compiler adds no FMAs.
 */
#include "ftwa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* nTimePoints = floor(tMax/dt + 1) in FP64 (src/FHSolver.cpp:100) */
int64_t oracle_n_time_points(double dt, double t_max)
{
    return (int64_t)floor(t_max / dt + 1.0);
}

/* One complex noise value w with E|w|^2 = 1, Re and Im independent N(0, 1/2):
 * Box-Muller (1958) on two 53-bit uniforms from one Philox block. */
static void box_muller(const uint32_t ctr[4], uint64_t seed, double* re, double* im)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    oracle_philox4x32_10(ctr, key, x);
    const double two_m53 = 1.1102230246251565404e-16; /* 2^-53 */
    double u1 = ((double)(x[0] >> 5) * 67108864.0 + (double)(x[1] >> 6) + 1.0) * two_m53; /* (0,1] */
    double u2 = ((double)(x[2] >> 5) * 67108864.0 + (double)(x[3] >> 6)) * two_m53;       /* [0,1) */
    double r = sqrt(-log(u1));
    double th = 6.283185307179586232 * u2;
    *re = r * cos(th);
    *im = r * sin(th);
}

/* initial-state noise of pair `pair` of spin `spin` (second counter word 0 or 1) */
static void noise_value(uint64_t seed, int64_t traj, int spin, uint32_t pair, double* re, double* im)
{
    uint32_t ctr[4] = {pair, (uint32_t)spin, (uint32_t)(uint64_t)traj, (uint32_t)((uint64_t)traj >> 32)};
    box_muller(ctr, seed, re, im);
}

/* GPSR diffusion noise xi_k(p) of step `step` (bit 31 of the second counter word set: a stream apart
 * from the initial-state noise) */
static void gpsr_noise_value(uint64_t seed, int64_t traj, int64_t step, int k, int p, double* re, double* im)
{
    uint32_t ctr[4] = {(uint32_t)(uint64_t)step, 0x80000000u | ((uint32_t)k << 16) | (uint32_t)p,
                       (uint32_t)(uint64_t)traj, (uint32_t)((uint64_t)traj >> 32)};
    box_muller(ctr, seed, re, im);
}

void oracle_gpsr_noise(int L, int64_t traj, int64_t step, uint64_t seed, double* out)
{
    for (int k = 0; k < 2; ++k)
        for (int p = 0; p < L; ++p) gpsr_noise_value(seed, traj, step, k, p, &out[2 * (k * L + p)], &out[2 * (k * L + p) + 1]);
}

void oracle_noise(int L, int64_t first, int64_t count, uint64_t seed, double* out)
{
    const int64_t npairs = (int64_t)L * (L - 1) / 2;
    for (int64_t b = 0; b < count; ++b)
        for (int s = 0; s < 2; ++s)
            for (int64_t p = 0; p < npairs; ++p) {
                double* w = out + (((b * 2 + s) * npairs) + p) * 2;
                noise_value(seed, first + b, s, (uint32_t)p, &w[0], &w[1]);
            }
}

/* A(t) scalar amplitude: src/FHExternalPotentialPulse.cpp:88 (Gaussian-enveloped cosine),
 * or piecewise-linear knots t_k = k*t_max/(n-1) (an extension; no reference counterpart). */
double oracle_pulse_amplitude(const or_system* s, double t, double t_max)
{
    if (s->pulse_type == 1) {
        double x = (t - s->t0) / s->td;
        return s->A0 * exp(-(x * x) / 2.0) * cos(s->w0 * (t - s->t0) + s->phi);
    }
    if (s->pulse_type == 2) {
        int n = s->n_knots;
        if (n <= 0) return 0.0;
        if (n == 1) return s->knots[0];
        double step = t_max / (double)(n - 1);
        double u = t / step;
        if (u <= 0.0) return s->knots[0];
        int k = (int)floor(u);
        if (k >= n - 1) return s->knots[n - 1];
        double f = u - (double)k;
        return s->knots[k] * (1.0 - f) + s->knots[k + 1] * f;
    }
    return 0.0;
}

/* extPot(t)_ij = tM_ij exp(i A(t).(r_i - r_j)) where tM_ij != 0 (src/FHExternalPotentialPulse.cpp:86-110) */
static void ext_pot(const or_system* s, double t, double t_max, double* P /* [L][L][2] */)
{
    const int L = s->L;
    double a = oracle_pulse_amplitude(s, t, t_max);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double tij = s->tM[i * L + j];
            double* pij = P + 2 * (i * L + j);
            if (tij != 0.0) {
                const double* d = s->disp + 3 * (i * L + j);
                double ph = a * (s->pol[0] * d[0] + s->pol[1] * d[1] + s->pol[2] * d[2]);
                pij[0] = tij * cos(ph);
                pij[1] = tij * sin(ph);
            } else {
                pij[0] = 0.0;
                pij[1] = 0.0;
            }
        }
}

/* Delta_sigma = tM + diag(U (rho_sbar_ii + 1/2)) [+ diag(V (rho_s_ii + 1/2))] + extPot
 * (src/FHfTWASolver.cpp:571-596).  Complex [L][L][2]. */
static void build_delta(const or_system* s, const double* rho_same, const double* rho_other,
                        const double* P /* pulse extPot or NULL */, double* D)
{
    const int L = s->L;
    for (int i = 0; i < L * L; ++i) {
        D[2 * i] = s->tM[i];
        D[2 * i + 1] = 0.0;
    }
    if (s->gpsr) {
        /* Delta_sigma = tM + diag(U rho_sbar.diag()) (+ extPot): no 1/2, no V (src/FHGPSRSolver.cpp:112-131) */
        if (s->coulomb_type != 0)
            for (int i = 0; i < L; ++i) {
                double hr = 0.0, hi = 0.0;
                for (int j = 0; j < L; ++j) {
                    double u = s->U[i * L + j];
                    hr += u * rho_other[2 * (j * L + j)];
                    hi += u * rho_other[2 * (j * L + j) + 1];
                }
                D[2 * (i * L + i)] += hr;
                D[2 * (i * L + i) + 1] += hi;
            }
    } else if (s->coulomb_type != 0) {
        for (int i = 0; i < L; ++i) {
            double hr = 0.0, hi = 0.0;
            for (int j = 0; j < L; ++j) {
                double u = s->U[i * L + j];
                hr += u * (rho_other[2 * (j * L + j)] + 0.5);
                hi += u * rho_other[2 * (j * L + j) + 1];
            }
            D[2 * (i * L + i)] += hr;
            D[2 * (i * L + i) + 1] += hi;
        }
        if (s->coulomb_type == 2) {
            for (int i = 0; i < L; ++i) {
                double hr = 0.0, hi = 0.0;
                for (int j = 0; j < L; ++j) {
                    double v = s->V[i * L + j];
                    hr += v * (rho_same[2 * (j * L + j)] + 0.5);
                    hi += v * rho_same[2 * (j * L + j) + 1];
                }
                D[2 * (i * L + i)] += hr;
                D[2 * (i * L + i) + 1] += hi;
            }
            if (s->fock) /* - N_same o V, N = rho + I/2 */
                for (int i = 0; i < L; ++i)
                    for (int j = 0; j < L; ++j) {
                        double v = s->V[i * L + j];
                        D[2 * (i * L + j)] -= v * (rho_same[2 * (i * L + j)] + (i == j ? 0.5 : 0.0));
                        D[2 * (i * L + j) + 1] -= v * rho_same[2 * (i * L + j) + 1];
                    }
        }
    }
    if (s->pot)
        for (int i = 0; i < L; ++i) D[2 * (i * L + i)] += s->pot[i];
    if (P)
        for (int i = 0; i < L * L; ++i) {
            D[2 * i] += P[2 * i];
            D[2 * i + 1] += P[2 * i + 1];
        }
}

/* f = -i (rho D - D rho)   (src/FHfTWASolver.cpp:598-599) */
static void commutator(int L, const double* rho, const double* D, double* f)
{
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double ar = 0.0, ai = 0.0; /* (rho D)_ij */
            double br = 0.0, bi = 0.0; /* (D rho)_ij */
            for (int k = 0; k < L; ++k) {
                const double* r1 = rho + 2 * (i * L + k);
                const double* d1 = D + 2 * (k * L + j);
                ar += r1[0] * d1[0] - r1[1] * d1[1];
                ai += r1[0] * d1[1] + r1[1] * d1[0];
                const double* d2 = D + 2 * (i * L + k);
                const double* r2 = rho + 2 * (k * L + j);
                br += d2[0] * r2[0] - d2[1] * r2[1];
                bi += d2[0] * r2[1] + d2[1] * r2[0];
            }
            double cr = ar - br, ci = ai - bi;
            f[2 * (i * L + j)] = ci;      /* -i (cr + i ci) = ci - i cr */
            f[2 * (i * L + j) + 1] = -cr;
        }
}

typedef struct work {
    double *D, *P, *k1u, *k1d, *k2u, *k2d, *k3u, *k3d, *k4u, *k4d, *yu, *yd;
} work;

static int work_alloc(work* w, int L)
{
    size_t n = (size_t)2 * L * L;
    double* base = (double*)calloc(n * 12, sizeof(double));
    if (!base) return -1;
    w->D = base; w->P = base + n;
    w->k1u = base + 2 * n; w->k1d = base + 3 * n;
    w->k2u = base + 4 * n; w->k2d = base + 5 * n;
    w->k3u = base + 6 * n; w->k3d = base + 7 * n;
    w->k4u = base + 8 * n; w->k4d = base + 9 * n;
    w->yu = base + 10 * n; w->yd = base + 11 * n;
    return 0;
}

static void derivative(const or_system* s, double t, double t_max, const double* ru, const double* rd,
                       double* fu, double* fd, work* w)
{
    const int L = s->L;
    const double* P = NULL;
    if (s->pulse_type != 0) {
        ext_pot(s, t, t_max, w->P);
        P = w->P;
    }
    build_delta(s, ru, rd, P, w->D);
    commutator(L, ru, w->D, fu);
    build_delta(s, rd, ru, P, w->D);
    commutator(L, rd, w->D, fd);
}

void oracle_derivative(const or_system* s, double t, double t_max, const double* rho_u, const double* rho_d,
                       double* f_u, double* f_d)
{
    work w;
    if (work_alloc(&w, s->L)) return;
    derivative(s, t, t_max, rho_u, rho_d, f_u, f_d, &w);
    free(w.D);
}

/* step_RK4_38 (src/FHfTWASolver.cpp:629-661).  Stage times t, t+h/3, t+2h/3, t+h
 * (the reference's third stage uses the member dt instead of h, :653; it only matters for a
 * time-dependent potential, which the reference never reaches -- this follows the tableau). */
static void step_rk4_38(const or_system* s, double t, double t_max, double h, double* ru, double* rd, work* w)
{
    const int n = 2 * s->L * s->L;
    derivative(s, t, t_max, ru, rd, w->k1u, w->k1d, w);
    for (int i = 0; i < n; ++i) {
        w->yu[i] = ru[i] + w->k1u[i] * h / 3.0;
        w->yd[i] = rd[i] + w->k1d[i] * h / 3.0;
    }
    derivative(s, t + h / 3.0, t_max, w->yu, w->yd, w->k2u, w->k2d, w);
    for (int i = 0; i < n; ++i) {
        w->yu[i] = ru[i] + (-w->k1u[i] / 3.0 + w->k2u[i]) * h;
        w->yd[i] = rd[i] + (-w->k1d[i] / 3.0 + w->k2d[i]) * h;
    }
    derivative(s, t + 2.0 * h / 3.0, t_max, w->yu, w->yd, w->k3u, w->k3d, w);
    for (int i = 0; i < n; ++i) {
        w->yu[i] = ru[i] + (w->k1u[i] - w->k2u[i] + w->k3u[i]) * h;
        w->yd[i] = rd[i] + (w->k1d[i] - w->k2d[i] + w->k3d[i]) * h;
    }
    derivative(s, t + h, t_max, w->yu, w->yd, w->k4u, w->k4d, w);
    for (int i = 0; i < n; ++i) {
        ru[i] = ru[i] + (w->k1u[i] + w->k2u[i] * 3.0 + w->k3u[i] * 3.0 + w->k4u[i]) * h / 8.0;
        rd[i] = rd[i] + (w->k1d[i] + w->k2d[i] * 3.0 + w->k3d[i] * 3.0 + w->k4d[i]) * h / 8.0;
    }
}

/* GPSR diffusion increment, the analytic form of FHGPSRSolver::step_Diffusion (src/FHGPSRSolver.cpp:211-232):
 *   d_up = c (rhoH W1 rho + i rho W2 rhoH),   d_dn = c (rhoH conj(W1) rho + i rho conj(W2) rhoH),
 * c = sqrt(i U0), U0 = Re U(0,0) (:630), rhoH = I - rho, W_k = diag(xi_k) with the step's scaled noise.
 * The reference runs the rsvd factorisation of the same diffusion matrix (:241-287; initiateSolver
 * :77 selects it), and adds diffusion only for an on-site Coulomb interaction (:248). */
static void gpsr_diffusion_one(int L, const double* rho, const double* x1, const double* x2, int conj_xi,
                               double cr, double ci, double* d, double* tmp)
{
    const double sg = conj_xi ? -1.0 : 1.0;
    /* tmp = rhoH W1 rho */
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double ar = 0.0, ai = 0.0;
            for (int k = 0; k < L; ++k) {
                double hr = (i == k ? 1.0 : 0.0) - rho[2 * (i * L + k)], hi = -rho[2 * (i * L + k) + 1];
                double wr = x1[2 * k], wi = sg * x1[2 * k + 1];
                double mr = hr * wr - hi * wi, mi = hr * wi + hi * wr; /* (rhoH W1)_ik */
                double pr = rho[2 * (k * L + j)], pi = rho[2 * (k * L + j) + 1];
                ar += mr * pr - mi * pi;
                ai += mr * pi + mi * pr;
            }
            tmp[2 * (i * L + j)] = ar;
            tmp[2 * (i * L + j) + 1] = ai;
        }
    /* d = c (tmp + i rho W2 rhoH) */
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double ar = 0.0, ai = 0.0;
            for (int k = 0; k < L; ++k) {
                double pr = rho[2 * (i * L + k)], pi = rho[2 * (i * L + k) + 1];
                double wr = x2[2 * k], wi = sg * x2[2 * k + 1];
                double mr = pr * wr - pi * wi, mi = pr * wi + pi * wr; /* (rho W2)_ik */
                double hr = (k == j ? 1.0 : 0.0) - rho[2 * (k * L + j)], hi = -rho[2 * (k * L + j) + 1];
                ar += mr * hr - mi * hi;
                ai += mr * hi + mi * hr;
            }
            double sr = tmp[2 * (i * L + j)] - ai, si = tmp[2 * (i * L + j) + 1] + ar; /* tmp + i (ar + i ai) */
            d[2 * (i * L + j)] = cr * sr - ci * si;
            d[2 * (i * L + j) + 1] = cr * si + ci * sr;
        }
}

static void gpsr_constant(const or_system* s, double* cr, double* ci)
{
    /* sqrt(i U0): U0 >= 0 gives sqrt(U0/2) (1 + i), U0 < 0 gives sqrt(-U0/2) (1 - i) */
    const double u0 = s->coulomb_type == 1 ? s->U[0] : 0.0;
    const double m = sqrt(fabs(u0) / 2.0);
    *cr = m;
    *ci = u0 >= 0.0 ? m : -m;
}

void oracle_gpsr_diffusion(const or_system* s, const double* rho_u, const double* rho_d, const double* xi,
                           double* d_u, double* d_d)
{
    const int L = s->L;
    double cr, ci;
    gpsr_constant(s, &cr, &ci);
    double* tmp = (double*)malloc(sizeof(double) * 2 * L * L);
    if (!tmp) return;
    gpsr_diffusion_one(L, rho_u, xi, xi + 2 * L, 0, cr, ci, d_u, tmp);
    gpsr_diffusion_one(L, rho_d, xi, xi + 2 * L, 1, cr, ci, d_d, tmp);
    free(tmp);
}

/* FHGPSRSolver::step (src/FHGPSRSolver.cpp:608-637): a trajectory holding a NaN or Inf becomes all NaN;
 * otherwise rho += RK4 "3/8" drift (:168-201) + diffusion, both from the step's initial state.  xi: the
 * step's noise [2][L][2] already scaled by sqrt(h); du, dd, tmp: scratch [L][L][2]. */
static void gpsr_step(const or_system* s, double t, double t_max, double h, double* ru, double* rd, const double* xi,
                      double* du, double* dd, double* tmp, work* w)
{
    const int n = 2 * s->L * s->L;
    int bad = 0;
    for (int i = 0; i < n; ++i)
        if (!isfinite(ru[i]) || !isfinite(rd[i])) bad = 1;
    if (bad) {
        for (int i = 0; i < n; ++i) {
            ru[i] = NAN;
            rd[i] = NAN;
        }
        return;
    }
    double cr, ci;
    gpsr_constant(s, &cr, &ci);
    gpsr_diffusion_one(s->L, ru, xi, xi + 2 * s->L, 0, cr, ci, du, tmp);
    gpsr_diffusion_one(s->L, rd, xi, xi + 2 * s->L, 1, cr, ci, dd, tmp);
    step_rk4_38(s, t, t_max, h, ru, rd, w);
    for (int i = 0; i < n; ++i) {
        ru[i] += du[i];
        rd[i] += dd[i];
    }
}

/* rho = R (diag(n - shift) + S.W) R^dagger   (src/FHfTWASolver.cpp:348-362, :187-198; shift 1/2).
 * GPSR: shift 0 and W = 0, i.e. the normally ordered R diag(n) R^dagger; the reference's GPSR starts from
 * diag(IC occupations) (src/FHGPSRSolver.cpp:77-83), the same state for the site-diagonal ICs. */
static void sample_state(int L, const double* nocc, const double* R, const double* wpairs /* [npairs][2] */,
                         double shift, double* rho, double* tmp)
{
    memset(rho, 0, sizeof(double) * 2 * L * L);
    for (int i = 0; i < L; ++i) rho[2 * (i * L + i)] = nocc[i] - shift;
    int64_t p = 0;
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j, ++p) {
            double sij = sqrt((nocc[i] + nocc[j] - 2.0 * nocc[i] * nocc[j]) / 2.0);
            double wr = wpairs[2 * p] * sij, wi = wpairs[2 * p + 1] * sij;
            rho[2 * (i * L + j)] = wr;
            rho[2 * (i * L + j) + 1] = wi;
            rho[2 * (j * L + i)] = wr;
            rho[2 * (j * L + i) + 1] = -wi;
        }
    if (!R) return;
    /* tmp = R * rho */
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double ar = 0.0, ai = 0.0;
            for (int k = 0; k < L; ++k) {
                const double* r = R + 2 * (i * L + k);
                const double* m = rho + 2 * (k * L + j);
                ar += r[0] * m[0] - r[1] * m[1];
                ai += r[0] * m[1] + r[1] * m[0];
            }
            tmp[2 * (i * L + j)] = ar;
            tmp[2 * (i * L + j) + 1] = ai;
        }
    /* rho = tmp * R^dagger */
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double ar = 0.0, ai = 0.0;
            for (int k = 0; k < L; ++k) {
                const double* m = tmp + 2 * (i * L + k);
                const double* r = R + 2 * (j * L + k); /* conj(R_jk) */
                ar += m[0] * r[0] + m[1] * r[1];
                ai += m[1] * r[0] - m[0] * r[1];
            }
            rho[2 * (i * L + j)] = ar;
            rho[2 * (i * L + j) + 1] = ai;
        }
}

/* Per time point accumulation (src/FHfTWASolver.cpp:366-418), real parts. */
static void record(const or_system* s, double t, double t_max, const double* ru, const double* rd, int64_t ti,
                   int64_t T, double* acc, work* w, int nb, const double* bases, int ce)
{
    const int L = s->L;
    double* su = acc;                      /* [L][T] */
    double* sd = acc + (size_t)L * T;
    double* s2u = acc + (size_t)2 * L * T;
    double* s2d = acc + (size_t)3 * L * T;
    double* sud = acc + (size_t)4 * L * T;
    double* se = acc + (size_t)5 * L * T;  /* [4][T] */
    double* sc = se + (size_t)4 * T;       /* [4][T] */
    for (int i = 0; i < L; ++i) {
        double ur = ru[2 * (i * L + i)], ui = ru[2 * (i * L + i) + 1];
        double dr = rd[2 * (i * L + i)], di = rd[2 * (i * L + i) + 1];
        su[(size_t)i * T + ti] += ur;
        sd[(size_t)i * T + ti] += dr;
        s2u[(size_t)i * T + ti] += ur * ur + ui * ui;
        s2d[(size_t)i * T + ti] += dr * dr + di * di;
        sud[(size_t)i * T + ti] += ur * dr - ui * di;
    }
    /* energies with N = rho + I/2 (:401-411); GPSR: N = rho (src/FHGPSRSolver.cpp:811-816), no E_V */
    const double hf = s->gpsr ? 0.0 : 0.5;
    double eth = 0.0;
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            double nr = ru[2 * (i * L + j)] + rd[2 * (i * L + j)] + (i == j ? 2.0 * hf : 0.0);
            eth += nr * s->tM[i * L + j];
        }
    double eu = 0.0, ev = 0.0;
    if (s->coulomb_type != 0) {
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                double nui = ru[2 * (i * L + i)] + hf, nuii = ru[2 * (i * L + i) + 1];
                double ndi = rd[2 * (i * L + i)] + hf, ndii = rd[2 * (i * L + i) + 1];
                double nuj = ru[2 * (j * L + j)] + hf, nujj = ru[2 * (j * L + j) + 1];
                double ndj = rd[2 * (j * L + j)] + hf, ndjj = rd[2 * (j * L + j) + 1];
                double u = s->U[i * L + j];
                /* Re(Nu_i U_ij Nd_j + Nd_i U_ij Nu_j) / 2 */
                eu += u * ((nui * ndj - nuii * ndjj) + (ndi * nuj - ndii * nujj)) / 2.0;
                if (s->coulomb_type == 2 && !s->gpsr) {
                    double v = s->V[i * L + j];
                    ev += v * ((nui * nuj - nuii * nujj) + (ndi * ndj - ndii * ndjj)) / 2.0;
                    if (s->fock) /* - Re sum V o N o N^T / 2 per spin */
                        for (int sg = 0; sg < 2; ++sg) {
                            const double* r = sg ? rd : ru;
                            double ar = r[2 * (i * L + j)] + (i == j ? 0.5 : 0.0), ai = r[2 * (i * L + j) + 1];
                            double br = r[2 * (j * L + i)] + (i == j ? 0.5 : 0.0), bi = r[2 * (j * L + i) + 1];
                            ev -= v * (ar * br - ai * bi) / 2.0;
                        }
                }
            }
    }
    double epe = 0.0;
    if (s->pot)
        for (int i = 0; i < L; ++i)
            epe += s->pot[i] * (ru[2 * (i * L + i)] + rd[2 * (i * L + i)] + 2.0 * hf);
    if (s->pulse_type != 0) {
        ext_pot(s, t, t_max, w->P);
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                double nr = ru[2 * (i * L + j)] + rd[2 * (i * L + j)] + (i == j ? 2.0 * hf : 0.0);
                double ni = ru[2 * (i * L + j) + 1] + rd[2 * (i * L + j) + 1];
                epe += nr * w->P[2 * (i * L + j)] - ni * w->P[2 * (i * L + j) + 1];
            }
    }
    se[0 * T + ti] += eth;
    se[1 * T + ti] += eu;
    se[2 * T + ti] += ev;
    se[3 * T + ti] += epe;
    sc[0 * T + ti] += ru[2 * ce];
    sc[1 * T + ti] += ru[2 * ce + 1];
    sc[2 * T + ti] += rd[2 * ce];
    sc[3 * T + ti] += rd[2 * ce + 1];
    /* diagvec(V^dagger rho V), real part (:385-394) */
    double* sb = sc + (size_t)4 * T; /* [nb][2][L][T] */
    for (int b = 0; b < nb; ++b)
        for (int sg = 0; sg < 2; ++sg) {
            const double* V = bases + (size_t)(b * 2 + sg) * L * L * 2;
            const double* rho = sg ? rd : ru;
            for (int k = 0; k < L; ++k) {
                double m = 0.0;
                for (int i = 0; i < L; ++i) {
                    double xr = 0.0, xi = 0.0; /* x = (rho V)_ik */
                    for (int j = 0; j < L; ++j) {
                        const double pr = rho[2 * (i * L + j)], pi = rho[2 * (i * L + j) + 1];
                        const double vr = V[2 * (j * L + k)], vi = V[2 * (j * L + k) + 1];
                        xr += pr * vr - pi * vi;
                        xi += pr * vi + pi * vr;
                    }
                    const double cr = V[2 * (i * L + k)], ci = -V[2 * (i * L + k) + 1]; /* conj(V_ik) */
                    m += cr * xr - ci * xi;
                }
                sb[((size_t)(b * 2 + sg) * L + k) * T + ti] += m;
            }
        }
}

static int run_impl(const or_system* s, const or_params* p, const double* noise, int use_seed, uint64_t seed,
                    int nthreads, or_out* o)
{
    const int L = s->L;
    if (L <= 0 || p->n_traj <= 0 || p->dt <= 0.0 || p->n_sub <= 0) return -1;
    const int64_t T = oracle_n_time_points(p->dt, p->t_max);
    const int64_t npairs = (int64_t)L * (L - 1) / 2;
    const int nb = p->bases ? p->n_bases : 0;
    const int ce = p->corner_j < 0 ? L - 1 : p->corner_i * L + p->corner_j; /* element index i*L + j */
    if (p->corner_j >= L || p->corner_i < 0 || p->corner_i >= L) return -1;
    const size_t acc_n = (size_t)(5 * L + 8) * T + (size_t)nb * 2 * L * T;
    const double h = p->dt / (double)p->n_sub;
    int nth = 1;
#ifdef _OPENMP
    nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    double* accs = (double*)calloc(acc_n * (size_t)nth, sizeof(double));
    if (!accs) return -2;
    int failed = 0;

#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* acc = accs + acc_n * (size_t)tid;
        work w = {0};
        double* ru = (double*)malloc(sizeof(double) * 2 * L * L * 6 + sizeof(double) * 4 * L);
        double* wbuf = (double*)calloc((size_t)2 * (npairs > 0 ? npairs : 1) * 2, sizeof(double));
        int ok = (work_alloc(&w, L) == 0) && ru && wbuf;
        if (!ok) {
#pragma omp atomic write
            failed = 1;
        }
        double* rd = ru + 2 * L * L;
        double* tmp = rd + 2 * L * L;
        double* du = tmp + 2 * L * L; /* GPSR diffusion increments and noise */
        double* dd = du + 2 * L * L;
        double* tmp2 = dd + 2 * L * L;
        double* xi = tmp2 + 2 * L * L;
        const double sh = sqrt(h);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < p->n_traj; ++b) {
            if (!ok) continue;
            const double* wb;
            if (s->gpsr) {
                wb = wbuf; /* zeros: GPSR starts from the state itself */
            } else if (use_seed) {
                oracle_noise(L, b, 1, seed, wbuf);
                wb = wbuf;
            } else {
                wb = noise + (size_t)b * 2 * npairs * 2;
            }
            const double shift = s->gpsr ? 0.0 : 0.5;
            sample_state(L, p->n_up, p->rot_up, wb, shift, ru, tmp);
            sample_state(L, p->n_dn, p->rot_dn, wb + 2 * npairs, shift, rd, tmp);
            double t = 0.0;
            for (int64_t ti = 0; ti < T; ++ti) {
                record(s, ti * p->dt, p->t_max, ru, rd, ti, T, acc, &w, nb, p->bases, ce);
                if (ti == T - 1) break; /* steps after the last record change no output */
                for (int ss = 0; ss < p->n_sub; ++ss) {
                    if (s->gpsr) {
                        /* noise of variance h per step (the reference scales by sqrt(dt) of the time grid,
                           src/FHGPSRSolver.cpp:219-220, :281, which equals h for stepsParTimePoints = 1) */
                        oracle_gpsr_noise(L, b, ti * p->n_sub + ss, s->gpsr_seed, xi);
                        for (int k = 0; k < 4 * L; ++k) xi[k] *= sh;
                        gpsr_step(s, t, p->t_max, h, ru, rd, xi, du, dd, tmp2, &w);
                    } else {
                        step_rk4_38(s, t, p->t_max, h, ru, rd, &w);
                    }
                    t += h;
                }
            }
            if (o->final_states) {
                double* fs = o->final_states + (size_t)b * 4 * L * L;
                memcpy(fs, ru, sizeof(double) * 2 * L * L);
                memcpy(fs + 2 * L * L, rd, sizeof(double) * 2 * L * L);
            }
        }
        if (ok) free(w.D);
        free(ru);
        free(wbuf);
    }
    if (failed) {
        free(accs);
        return -2;
    }
    /* thread-local -> global in thread order (src/FHfTWASolver.cpp:439-482) */
    for (int th = 1; th < nth; ++th)
        for (size_t i = 0; i < acc_n; ++i) accs[i] += accs[acc_n * (size_t)th + i];
    size_t LT = (size_t)L * T;
    if (o->sum_u) memcpy(o->sum_u, accs, sizeof(double) * LT);
    if (o->sum_d) memcpy(o->sum_d, accs + LT, sizeof(double) * LT);
    if (o->sum2_u) memcpy(o->sum2_u, accs + 2 * LT, sizeof(double) * LT);
    if (o->sum2_d) memcpy(o->sum2_d, accs + 3 * LT, sizeof(double) * LT);
    if (o->sum_ud) memcpy(o->sum_ud, accs + 4 * LT, sizeof(double) * LT);
    if (o->sum_e) memcpy(o->sum_e, accs + 5 * LT, sizeof(double) * 4 * T);
    if (o->sum_corner) memcpy(o->sum_corner, accs + 5 * LT + 4 * T, sizeof(double) * 4 * T);
    if (o->sum_basis && nb) memcpy(o->sum_basis, accs + 5 * LT + 8 * T, sizeof(double) * nb * 2 * LT);
    free(accs);
    return 0;
}

int oracle_run(const or_system* s, const or_params* p, const double* noise, int nthreads, or_out* o)
{
    return run_impl(s, p, noise, 0, 0, nthreads, o);
}

int oracle_run_seeded(const or_system* s, const or_params* p, uint64_t seed, int nthreads, or_out* o)
{
    return run_impl(s, p, NULL, 1, seed, nthreads, o);
}
