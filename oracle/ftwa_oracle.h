/*
 * ftwa_oracle.h -- CPU FP64 restatement of the reference fTWA algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fermiphasespace_amd/) links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do,
 * and only as the checker / the timed CPU baseline.
 *
 * It follows, function by function:
 *   oracle_philox4x32_10  Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11
 *                         (Philox4x32 with 10 rounds; pinned by the Random123 known-answer vectors)
 *   oracle_noise          the complex white noise of src/FHfTWASolver.cpp:343-346, drawn from a
 *                         counter-based stream instead of Armadillo's randomly seeded one
 *   oracle_run            FHfTWASolver::initiateSolver (src/FHfTWASolver.cpp:99-118, :187-198),
 *                         runSimulation (:222-556; tbStates / mfStates occupations :385-394),
 *                         derivative (:568-600), step_RK4_38 (:629-661); with or_system.gpsr set,
 *                         FHGPSRSolver::initiateSolver (src/FHGPSRSolver.cpp:70-101), derivative
 *                         (:110-133), step_Drift_RK4_38 (:168-201), step (:608-637), runSimulation
 *                         (:641-930)
 *   oracle_gpsr_diffusion the analytic diffusion increment FHGPSRSolver::step_Diffusion
 *                         (src/FHGPSRSolver.cpp:211-232), whose second moments are the diffusion matrix
 *                         D = [[0, A], [A^T, 0]] of the rsvd path the reference runs (:241-287, A from
 *                         quarterDiffusionMatrix :337-383)
 */
#ifndef FTWA_ORACLE_H
#define FTWA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_system {
    int L;
    const double* tM;    /* [L][L] real symmetric hopping                         */
    int coulomb_type;    /* 0 null, 1 onSite, 2 general (U and V)                 */
    const double* U;     /* [L][L]                                                */
    const double* V;     /* [L][L] (general only)                                 */
    const double* pot;   /* [L] constant diagonal potential or NULL               */
    int pulse_type;      /* 0 none, 1 gaussian, 2 knots                           */
    double A0, t0, td, w0, phi;
    double pol[3];
    int n_knots;
    const double* knots;
    const double* disp;  /* [L][L][3] r_i - r_j                                   */
    int fock;            /* 1: Hartree-Fock, Delta_s gains -N_s o V and E_V its exchange term
                            (FHHFSolver::derivative, src/FHHFSolver.cpp:575-608, energies :356-357) */
    int gpsr;            /* 1: Gaussian phase-space representation (FHGPSRSolver): rho is the
                            normally ordered density matrix (no 1/2 shifts), Delta_s = tM + diag(U rho_sbar_ii)
                            + extPot (no V), and every step adds the stochastic diffusion term        */
    uint64_t gpsr_seed;  /* key of the counter-based diffusion noise                                   */
} or_system;

typedef struct or_params {
    double dt, t_max;
    int n_sub;
    int64_t n_traj;
    const double* n_up;   /* [L] */
    const double* n_dn;   /* [L] */
    const double* rot_up; /* [L][L] complex interleaved or NULL (identity) */
    const double* rot_dn;
    int n_bases;          /* bases whose occupations diag(V^dagger rho V) are summed (tbStates / mfStates) */
    const double* bases;  /* [n_bases][2 spins][L][L] complex interleaved, columns the basis states */
    int corner_i, corner_j; /* element summed into sum_corner: rho(corner_i, corner_j); corner_j < 0 means
                               (0, L-1), the wholeDMatrix corner (src/FHSolver.cpp:196-197); a core-hole
                               system sums (site, L-1) (src/FHfTWASolver.cpp:413-417) */
} or_params;

/* Raw ensemble sums (real parts), each [rows][T], any pointer may be NULL. */
typedef struct or_out {
    double* sum_u;      /* [L][T] sum rho_up_ii            */
    double* sum_d;      /* [L][T] sum rho_dn_ii            */
    double* sum2_u;     /* [L][T] sum |rho_up_ii|^2        */
    double* sum2_d;     /* [L][T] sum |rho_dn_ii|^2        */
    double* sum_ud;     /* [L][T] sum rho_up_ii rho_dn_ii  */
    double* sum_e;      /* [4][T] tH, U, V, pe             */
    double* sum_corner; /* [4][T] re/im rho_up(ci,cj), re/im rho_dn(ci,cj) (default (0, L-1)) */
    double* final_states; /* [B][2][L][L][2]               */
    double* sum_basis;  /* [n_bases][2][L][T] sum Re diag(V^dagger rho V) (src/FHfTWASolver.cpp:385-394) */
} or_out;

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int64_t oracle_n_time_points(double dt, double t_max);
/* noise for trajectories [first, first+count): [count][2][L(L-1)/2][2] */
void oracle_noise(int L, int64_t first, int64_t count, uint64_t seed, double* out);
/* pulse amplitude a(t) (scalar multiplying pol) */
double oracle_pulse_amplitude(const or_system* s, double t, double t_max);
/* Runs B = p->n_traj trajectories from the given noise ([B][2][npairs][2]), nthreads OpenMP threads
 * (0 = all).  Returns 0 on success. */
int oracle_run(const or_system* s, const or_params* p, const double* noise, int nthreads, or_out* o);
/* Same, drawing the noise from the counter-based generator (seed) trajectory by trajectory,
 * which keeps memory O(threads) for large B (used for the CPU baseline). */
int oracle_run_seeded(const or_system* s, const or_params* p, uint64_t seed, int nthreads, or_out* o);
/* Diffusion noise of step `step` of trajectory `traj`: xi_1 = out[0][.], xi_2 = out[1][.], complex
 * [2][L][2], E|xi|^2 = 1 (Re and Im independent N(0, 1/2)); a step of length h scales it by sqrt(h). */
void oracle_gpsr_noise(int L, int64_t traj, int64_t step, uint64_t seed, double* out);
/* One GPSR diffusion increment at rho (complex [L][L][2] per spin) for the scaled noise xi ([2][L][2]):
 * d_up = c (rhoH W1 rho + i rho W2 rhoH), d_dn the same with conj(W1), conj(W2), c = sqrt(i U(0,0)),
 * rhoH = I - rho, W_k = diag(xi_k); zero unless the Coulomb interaction is on-site. */
void oracle_gpsr_diffusion(const or_system* s, const double* rho_u, const double* rho_d, const double* xi,
                           double* d_u, double* d_d);
/* One derivative evaluation f = -i(rho Delta - Delta rho) for both spins (for tests). */
void oracle_derivative(const or_system* s, double t, double t_max, const double* rho_u, const double* rho_d,
                       double* f_u, double* f_d);

#ifdef __cplusplus
}
#endif
#endif
