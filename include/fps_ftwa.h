/*
 * fps_ftwa.h -- C ABI of the MI355X (gfx950) fTWA trajectory engine.
 *
 * The reference has no FFI: its boundary is the C++ class FHfTWASolver, driven by
 * fps.cpp:192-207 (construct, five setters, initiateSolver(), runSimulation()).
 * Each entry point below replaces one piece of that class; the citation next to it
 * names the reference member it stands for.  Plain pointers and sizes only: no C++
 * or torch types cross this boundary, and no exception escapes it.
 *
 * Conventions
 *   - Every call returns FPSFTWA_OK (0) or a negative error code; the message of the
 *     last failure on the calling thread is fpsftwa_last_error().
 *   - Matrices are row-major, M[i*L + j].  Complex arrays are interleaved (re, im).
 *   - Input arrays are copied in: the caller keeps ownership.
 *   - A handle owns one device, one HIP stream and its device buffers.  Handles are
 *     not shared between threads; different handles may run concurrently.
 *   - There is no CPU fallback: without a usable gfx950 device, fpsftwa_create()
 *     fails with FPSFTWA_EDEVICE.
 */
#ifndef FPS_FTWA_H
#define FPS_FTWA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FPSFTWA_VERSION_MAJOR 0
#define FPSFTWA_VERSION_MINOR 1
#define FPSFTWA_VERSION_PATCH 0

/* ---- error codes ------------------------------------------------------------ */
#define FPSFTWA_OK 0
#define FPSFTWA_EINVAL -1       /* bad argument / inconsistent sizes            */
#define FPSFTWA_EDEVICE -2      /* no gfx950 device, or a HIP runtime failure   */
#define FPSFTWA_ENOMEM -3       /* device allocation failed                     */
#define FPSFTWA_ESTATE -4       /* call out of order (e.g. result before run)   */
#define FPSFTWA_EUNSUPPORTED -5 /* valid request this build does not implement  */
#define FPSFTWA_EIO -6          /* file could not be read / written             */

/* ---- Coulomb interaction type (FHCoulombInteractionType, include/FHCoulombInteraction.h:14-19) */
#define FPSFTWA_COULOMB_NULL 0    /* CEnull: no Hartree term at all                 */
#define FPSFTWA_COULOMB_ONSITE 1  /* CEonSite: U only (V skipped)                   */
#define FPSFTWA_COULOMB_GENERAL 2 /* CEdistanceInteraction / CEgiven: U and V       */

/* ---- time-dependent potential (FHExternalPotentialPulse, src/FHExternalPotentialPulse.cpp:86-110) */
#define FPSFTWA_PULSE_NONE 0
#define FPSFTWA_PULSE_GAUSSIAN 1 /* A(t) = A0 exp(-((t-t0)/td)^2/2) cos(w0 (t-t0) + phi) */
#define FPSFTWA_PULSE_KNOTS 2    /* A(t) piecewise linear through equally spaced knots   */

/* ---- trajectory engine (precision / kernel family) ----------------------------- */
#define FPSFTWA_ENGINE_FP64_PLAIN 0   /* dense Delta, reference operation order, one launch per time point */
#define FPSFTWA_ENGINE_FP64_BATCHED 1 /* LDS sparse-stencil kernel, FP64 trajectories                     */
#define FPSFTWA_ENGINE_FP32_BATCHED 2 /* LDS sparse-stencil kernel, FP32 trajectories, FP64 ensemble sums */

/* ---- dynamics (fpsftwa_set_method) ------------------------------------------------ */
#define FPSFTWA_METHOD_FTWA 0 /* fTWA: sampled initial states, Hartree mean field (FHfTWASolver)       */
#define FPSFTWA_METHOD_HF 1   /* time-dependent Hartree-Fock: one deterministic state, Fock exchange  */
                              /* with the V interaction (FHHFSolver, src/FHHFSolver.cpp:575-608)       */
#define FPSFTWA_METHOD_GPSR 2 /* Gaussian phase-space representation: stochastic trajectories of the   */
                              /* normally ordered density matrix (FHGPSRSolver, src/FHGPSRSolver.cpp)  */

/* ---- results to save (FHSolver::setResultsToSave, src/FHSolver.cpp:34-45) -------- */
#define FPSFTWA_RES_DIAG 0x01     /* "diagDMatrix"   -> Niiu, Niid          */
#define FPSFTWA_RES_DIAG2 0x02    /* "diag2DMatrix"  -> N2iiu, N2iid        */
#define FPSFTWA_RES_DOUBLONS 0x04 /* "onSiteDoublons"-> NuNd                */
#define FPSFTWA_RES_ENERGIES 0x08 /* "energies"      -> E {tH,U,V,pe,tot}   */
#define FPSFTWA_RES_WHOLE 0x10    /* "wholeDMatrix"  -> N1hu, N1hd (corner) */

/* ---- result arrays for fpsftwa_get_result ----------------------------------------- */
#define FPSFTWA_OUT_TIME 0      /* [T]            t = n*dt                                   */
#define FPSFTWA_OUT_NIIU 1      /* [L][T]         Niiu  (src/FHfTWASolver.cpp:509)           */
#define FPSFTWA_OUT_NIID 2      /* [L][T]         Niid  (:510)                               */
#define FPSFTWA_OUT_N2IIU 3     /* [L][T]         N2iiu (:520)                               */
#define FPSFTWA_OUT_N2IID 4     /* [L][T]         N2iid (:521)                               */
#define FPSFTWA_OUT_NUND 5      /* [L][T]         NuNd  (:530)                               */
#define FPSFTWA_OUT_ENERGY 6    /* [5][T]         tH, U, V, pe, tot (:399-412, :543-547)     */
#define FPSFTWA_OUT_N1HU 7      /* [2][T]         re, im of <N_{0,L-1}> up   (FHSolver.cpp:196), or of the  */
#define FPSFTWA_OUT_N1HD 8      /* [2][T]         ... down (:197); element set by fpsftwa_set_recorded_element */
#define FPSFTWA_OUT_NIIU_ERR 9  /* [L][T]         statistical error sigma/sqrt(B) of Niiu    */
#define FPSFTWA_OUT_NIID_ERR 10 /* [L][T]                                          of Niid    */
#define FPSFTWA_OUT_NUND_ERR 11 /* [L][T]                                          of NuNd    */
#define FPSFTWA_OUT_ENERGY_ERR 12 /* [4][T]                                        of E rows  */
/* raw ensemble sums (before the reference's normalisation), for parity checks */
#define FPSFTWA_OUT_SUM_RHO_U 13   /* [L][T] sum_b rho_up_ii                        */
#define FPSFTWA_OUT_SUM_RHO_D 14   /* [L][T] sum_b rho_dn_ii                        */
#define FPSFTWA_OUT_SUM_RHO2_U 15  /* [L][T] sum_b |rho_up_ii|^2                    */
#define FPSFTWA_OUT_SUM_RHO2_D 16  /* [L][T] sum_b |rho_dn_ii|^2                    */
#define FPSFTWA_OUT_SUM_RHOUD 17   /* [L][T] sum_b rho_up_ii rho_dn_ii               */
#define FPSFTWA_OUT_SUM_ENERGY 18  /* [4][T] sum_b E_k                               */
/* occupations of the bases given to fpsftwa_set_bases (tbStates -> Mkku/Mkkd, mfStates -> Mukku/Mukkd) */
#define FPSFTWA_OUT_BASIS 19       /* [n_bases][2][L][T] sum_b diag(V^T rho V) / B + 1/2 (src/FHfTWASolver.cpp:531-540) */
#define FPSFTWA_OUT_SUM_BASIS 20   /* [n_bases][2][L][T] sum_b diag(V^T rho V)                                     */
#define FPSFTWA_OUT_COUNT 21

/* System matrices, as FHSystem hands them to the solver through getUMatrix / gettMatrix /
 * getVMatrix / getExtPot (src/FHSystem.cpp:387-391).  All values are real: the reference's
 * JSON front end only ever builds real U, V, tM and a real diagonal extPot. */
typedef struct fpsftwa_system {
    int32_t n_sites;               /* L                                                       */
    const double* hopping;         /* [L][L] tM, real symmetric (FHTunnelingInteraction)     */
    int32_t coulomb_type;          /* FPSFTWA_COULOMB_*                                      */
    const double* coulomb_u;       /* [L][L] U (opposite spin); NULL only for COULOMB_NULL   */
    const double* coulomb_v;       /* [L][L] V (same spin); read only for COULOMB_GENERAL    */
    const double* potential;       /* [L] diagonal of a constant extPot, or NULL (EPnull)    */
    /* Peierls pulse on the hopping: extPot_ij(t) = tM_ij exp(i A(t).(r_i - r_j)), added to tM */
    int32_t pulse_type;            /* FPSFTWA_PULSE_*                                         */
    double pulse_A0, pulse_t0, pulse_td, pulse_w0, pulse_phi;
    double pulse_pol[3];           /* polarisation vector                                     */
    int32_t pulse_n_knots;         /* PULSE_KNOTS: amplitudes a_k at t_k = k*t_max/(n-1)      */
    const double* pulse_knots;     /* [n_knots]                                               */
    const double* displacement;    /* [L][L][3] r_i - r_j (minimum image when periodic), or NULL when no pulse */
} fpsftwa_system;

typedef struct fpsftwa_handle fpsftwa_handle;

/* Message of the last failed call on this thread ("" if none). */
const char* fpsftwa_last_error(void);

/* Library version and the device it runs on (name buffer may be NULL). */
int fpsftwa_version(int* major, int* minor, int* patch);
int fpsftwa_device_info(char* name, size_t name_len, int* cu_count, double* hbm_gib);

/* Peaks measured on the current device (no counterpart in the reference: the roofline denominators
 * re-measured on the machine, beside the published figures): a 1 GiB device copy (bytes read plus
 * written per second), FP32 and FP64 FMA loops (non-packed FP32, as the integrators are built) and a
 * 32x32x2 FP32 MFMA loop.  Allocates 2 GiB of device memory for the copy and takes about a second. */
typedef struct fpsftwa_peaks {
    double hbm_copy_gbps;     /* (read + written bytes) / s, in GB/s                          */
    double copy_bytes;        /* bytes moved per copy launch                                  */
    double fp32_fma_tflops;   /* v_fma_f32, 2 flop per lane and instruction                   */
    double fp64_fma_tflops;   /* v_fma_f64                                                    */
    double fp32_mfma_tflops;  /* v_mfma_f32_32x32x2_f32, 4,096 flop per wave and instruction  */
} fpsftwa_peaks;
int fpsftwa_measure_peaks(fpsftwa_peaks* out);

/* FHfTWASolver::FHfTWASolver(FHSystem*) -- include/FHfTWASolver.h:28, src/FHfTWASolver.cpp:4-16.
 * Copies the system in; fails with FPSFTWA_EDEVICE when no gfx950 device is usable. */
int fpsftwa_create(const fpsftwa_system* system, fpsftwa_handle** out);
int fpsftwa_destroy(fpsftwa_handle* h);

/* FHfTWASolver::setMethodParameters / setTimeParameters / setfTWAParameters
 * (src/FHfTWASolver.cpp:49-69, src/FHSolver.cpp:95-101).  nTimePoints = floor(t_max/dt + 1);
 * each time point is followed by n_sub RK4 "3/8" steps of dt/n_sub.  int_order is accepted and
 * ignored, as the reference forces RK4-3/8 in initiateSolver (src/FHfTWASolver.cpp:90).
 * seed keys the counter-based noise; engine is one of FPSFTWA_ENGINE_*. */
int fpsftwa_set_params(fpsftwa_handle* h, double dt, double t_max, int32_t n_sub, int64_t n_traj,
                       int32_t int_order, uint64_t seed, int32_t engine);

/* FHSolver::setResultsToSave (src/FHSolver.cpp:34-45); flags = FPSFTWA_RES_* bitmask. */
int fpsftwa_set_results(fpsftwa_handle* h, uint32_t flags);

/* Initial state, as initiateSolver builds it (src/FHfTWASolver.cpp:99-118, :187-198):
 * occupations n_up[L], n_dn[L] in [0,1] and optional unitary rotations R_up, R_dn
 * ([L][L] complex interleaved; NULL = identity, the givenDiagonal/spinWave/chargeWave case). */
int fpsftwa_set_initial_state(fpsftwa_handle* h, const double* n_up, const double* n_dn,
                              const double* rot_up, const double* rot_dn);

/* Fixed one-particle bases whose occupations diag(V^T rho V) are recorded at every time point, as
 * runSimulation does for "tbStates" (tbgs_V) and "mfStates" (mfgs_Vu, mfgs_Vd)
 * (src/FHfTWASolver.cpp:385-394, :531-540).  bases = [n_bases][2 spins][L][L] real, row-major, columns
 * the basis states (fpsftwa_ground_state's layout); n_bases in [0, 2], 0 (or NULL) clears them.  Read
 * the result with FPSFTWA_OUT_BASIS.  Not available with the adjoint. */
int fpsftwa_set_bases(fpsftwa_handle* h, int32_t n_bases, const double* bases);

/* FPSFTWA_METHOD_FTWA (default), FPSFTWA_METHOD_HF or FPSFTWA_METHOD_GPSR.  HF follows FHHFSolver
 * (src/FHHFSolver.cpp:190-420): the state is the density matrix N = R diag(n) R^dagger itself (no sampling;
 * every trajectory is the same, so n_traj 1 is the natural choice), Delta_sigma = tM + diag(U N_sbar) +
 * diag(V N_sigma) - N_sigma o V + extPot (:575-608), the same RK4 "3/8" steps, and E_V includes the
 * exchange energy -sum V o N o N^T / 2 (:356-357).  With a general (V) interaction the Fock term
 * needs the plain FP64 engine; with null or on-site interactions every engine runs HF.
 * GPSR follows FHGPSRSolver (src/FHGPSRSolver.cpp): every trajectory starts from rho = R diag(n) R^dagger
 * (the normally ordered density matrix, no 1/2), drifts with Delta_sigma = tM + diag(U rho_sbar_ii) +
 * extPot (no V; :110-133) by the RK4 "3/8" step and diffuses with c (rhoH W1 rho + i rho W2 rhoH), c =
 * sqrt(i U(0,0)), W_k diagonal complex noise of variance h per step from the handle's seed (:211-232; only
 * for an on-site interaction, :248); a trajectory that turns NaN or Inf stays NaN (:611-615).  The results
 * are plain ensemble means: Niiu = <Re rho_ii>, N2iiu = <|rho_ii|^2>, NuNd = <Re rho_up_ii rho_dn_ii>,
 * energies without V, Mkk = <Re diag(V^T rho V)> (:895-929).  The batched engines run GPSR for L <= 32, 4
 * or 8 neighbours and no pulse; the plain FP64 engine runs any system.  No adjoint. */
int fpsftwa_set_method(fpsftwa_handle* h, int32_t method);

/* The off-diagonal element rho_sigma(i, j) whose ensemble mean FPSFTWA_OUT_N1HU / N1HD return.
 * Default (0, L-1): the "wholeDMatrix" corner (src/FHSolver.cpp:196-197).  A core-hole system
 * records (site, L-1) instead, its "NiCh" observable (src/FHfTWASolver.cpp:413-417, :548-551). */
int fpsftwa_set_recorded_element(fpsftwa_handle* h, int32_t i, int32_t j);

/* Oracle mode: use the caller's noise instead of the counter-based generator.
 * noise = [n_traj][2 spins][L(L-1)/2 pairs][re, im], pair p enumerating i<j row by row;
 * the sampled state is rho_sigma = R (diag(n - 1/2) + S.W) R^dagger with
 * W_ij = noise, W_ji = conj(noise), S_ij = sqrt((n_i + n_j - 2 n_i n_j)/2)
 * (src/FHfTWASolver.cpp:342-362).  Pass NULL to return to the generator. */
int fpsftwa_set_noise(fpsftwa_handle* h, const double* noise, size_t n_values);

/* The generator's noise for trajectories [first, first+count), same layout as above. */
int fpsftwa_sample_noise(fpsftwa_handle* h, int64_t first, int64_t count, double* out, size_t n_values);

/* The HIP stream (a hipStream_t, passed as void*) the handle enqueues its work on; NULL returns to the
 * stream the handle created.  The reference has no streams; this is the stream argument of a run
 * (the caller's stream, e.g. a framework's current stream).  Waits for the handle's work on the
 * previous stream.  Runs stay synchronous: fpsftwa_run returns after its results reached the host. */
int fpsftwa_set_stream(fpsftwa_handle* h, void* stream);

/* No counterpart in the reference: with on != 0, a batched run captures its integrator and ensemble-
 * reduction launches (every time chunk) on the handle's stream into one HIP graph and launches that graph
 * instead of the individual kernels; runs that use the GEMM library (the dense path, basis occupations)
 * launch directly as before.  Results are bitwise the same either way.  Off by default. */
int fpsftwa_set_graph(fpsftwa_handle* h, int32_t on);

/* FHfTWASolver::runSimulation (src/FHfTWASolver.cpp:222-556) without the file write:
 * samples, integrates every trajectory, reduces the ensemble.  Synchronous. */
int fpsftwa_run(fpsftwa_handle* h);

/* Result arrays (FPSFTWA_OUT_*).  With out == NULL only *needed (element count) is set. */
int fpsftwa_get_result(fpsftwa_handle* h, int32_t which, double* out, size_t n_values, size_t* needed);

/* Final per-trajectory states rho_up, rho_dn ([count][2][L][L] complex interleaved, FP64)
 * after the last recorded time point. */
int fpsftwa_get_states(fpsftwa_handle* h, int64_t first, int64_t count, double* out, size_t n_values);

/* Device time of the dominant trajectory kernel over the last fpsftwa_run: launches and
 * summed milliseconds measured with HIP events on the handle's stream, plus whole-run ms. */
int fpsftwa_get_timing(fpsftwa_handle* h, int64_t* launches, double* kernel_ms, double* run_ms);

/* Engine tuning (no reference counterpart; the autotuner's knobs): threads per trajectory
 * workgroup for the full-matrix batched kernel (512 or 1024; 0 = automatic), time points per launch
 * (0 = automatic) and the FP32 kernel variant (FPSFTWA_KERNEL_*). */
#define FPSFTWA_KERNEL_AUTO 0      /* FP32: spin-sequential where it fits (at L = 64 with four
                                    * neighbours and one hopping value in its direct-commutator
                                    * form), else upper-triangle, else full-matrix; FP64 likewise
                                    * without the FP32-only ones                                    */
#define FPSFTWA_KERNEL_FULL 1      /* full-matrix stencil kernel                                   */
#define FPSFTWA_KERNEL_HERMITIAN 2 /* upper-triangle kernel (L <= 64; in FP64 the spin-sequential
                                    * one, no pulse, null or on-site Coulomb)                         */
#define FPSFTWA_KERNEL_DENSE 3     /* X = Y T per RK stage for the batch (any precision, L <= 1024)
                                    * with the library GEMM; AUTO takes the dense path for L > 64 or
                                    * more than 8 neighbours (in FP32 as DENSE_MFMA where L allows)   */
#define FPSFTWA_KERNEL_PIPELINED 4 /* upper-triangle kernel with the two spins half a stage apart
                                    * (FP32, L <= 64, no pulse, null or on-site Coulomb)              */
#define FPSFTWA_KERNEL_SEQUENTIAL 5 /* upper-triangle kernel with one spin's rows in LDS at a time,
                                    * two workgroups per CU in FP32 (L <= 64, X = T Y in LDS; FP32
                                    * also 64 < L <= 107, each element's commutator formed directly
                                    * from the rows; no pulse, null or on-site Coulomb, no basis
                                    * occupations)                                                    */
#define FPSFTWA_KERNEL_DENSE_MFMA 6 /* FP32 dense path with the product X = Y T and the RK update fused in
                                    * one hand-written matrix-core kernel per stage (v_mfma_f32_16x16x4f32,
                                    * 8 trajectory-spin slices per workgroup sharing the hopping operand
                                    * in LDS; L even with a pulse, L % 4 == 0 without; else as DENSE).
                                    * What AUTO runs on the dense path in FP32; DENSE keeps the library
                                    * GEMM followed by the update kernel                                 */
int fpsftwa_set_tuning(fpsftwa_handle* h, int32_t block_threads, int32_t time_chunk, int32_t kernel);

/* Autotuner (no reference counterpart): times every batched kernel family and shape that can run
 * this system (upper-triangle both-spin, spin-pipelined and spin-sequential, full-matrix stencil at
 * 512/1024 threads, dense with the library GEMM and, in FP32, fused on the matrix cores; two time-chunk
 * lengths) on a shortened copy of the configured run -- trial_steps steps of at most 16,384
 * trajectories, each candidate run twice and the second timed -- and keeps the fastest as the
 * handle's tuning (fpsftwa_set_tuning). Outputs may be NULL. Requires a batched engine. */
int fpsftwa_autotune(fpsftwa_handle* h, int32_t trial_steps, int32_t* kernel, int32_t* block_threads,
                     int32_t* time_chunk, double* traj_steps_per_s);

/* Runs n handles at once (no reference counterpart: the reference runs one system per process,
 * fps.cpp:192-207). Each handle runs fpsftwa_run from its own host thread on its own stream, so
 * systems of different size and lattice share the GPU; results are per handle, as from fpsftwa_run.
 * Returns the first failure (its message names the system). */
int fpsftwa_run_many(fpsftwa_handle** handles, int32_t n);

/* No counterpart in the reference: fpsftwa_run for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It makes no device
 * allocation: the caller passes the trajectory state and the ensemble sums as device buffers, and the
 * library's work memory (system tables, occupations, the trajectory records of a time chunk and their
 * partial sums) is a fixed device array of FPSFTWA_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends.  Same samples, kernels and sums as fpsftwa_run with the same handle settings.
 *   state  device, >= n_traj * 2 * L * L complex of the engine's precision (FP32 engine: float, FP64
 *          batched: double; interleaved re, im; [traj][spin][i][j]), state_bytes its size: receives the
 *          sampled initial states and holds the final states on return;
 *   sums   device, n_sums >= T * (6L + 12) * 2 doubles, T = floor(t_max/dt + 1): the ensemble sums per
 *          time point as (sum, compensation) pairs [T][Q][2], Q = 6L + 12; for site i, q = 6i + m with
 *          m = 0..5: sum rho_up_ii, sum rho_dn_ii, sum |rho_up_ii|^2, sum |rho_dn_ii|^2,
 *          sum rho_up_ii rho_dn_ii, sum ((rho_up_ii + 1/2)(rho_dn_ii + 1/2))^2; q = 6L + k (k < 4): sum E_k
 *          (tH, U, V, pe);
 *          q = 6L + 4 + k: sum E_k^2; q = 6L + 8 + c: the recorded element's sums (re, im up; re, im down).
 *          Every entry is written (nothing needs clearing); fpsftwa_get_result does not see this run.
 * Limits: the batched FP32 and FP64 engines on the stencil-sized kernels (not the dense path), the fTWA
 * method, the generator's noise, diagonal initial states, no basis occupations, and one time point's
 * records n_traj * (2L + 8) reals plus their partial sums within the scratch (FPSFTWA_EINVAL beyond).
 * Synchronous.  Calls that overlap on one GPU share its scratch and race: serialise them. */
#define FPSFTWA_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int fpsftwa_run_noalloc(fpsftwa_handle* h, void* state, size_t state_bytes, double* sums, size_t n_sums);

/* Discrete adjoint of a run (no reference counterpart; it differentiates what runSimulation
 * computes, src/FHfTWASolver.cpp:342-426 with the derivative :568-600 and the step :629-661).
 * Objective: the ensemble double occupancy at the last time point,
 *   J = (1/L) sum_i NuNd_i(t_max) = (1/B) sum_b (1/L) sum_i (rho_up,ii + 1/2)(rho_dn,ii + 1/2)
 * (NuNd as in :395-398, :524-531). The same samples as fpsftwa_run (seed, initial state, rotations)
 * are integrated with the textbook 3/8 tableau, the state is checkpointed every checkpoint_every
 * steps, and each segment is recomputed and reversed. Precision follows the engine set by
 * fpsftwa_set_params (FP32 batched: FP32 trajectories, FP64 gradient sums; otherwise FP64).
 * Outputs (each may be NULL):
 *   objective   J;
 *   grad_u[L]   dJ/dU_ii (on-site Coulomb);
 *   grad_t[L*L] dJ/dtM_kl for the entries of the hopping pattern (tM_kl != 0) and the diagonal,
 *               each entry treated as independent (a symmetric bond change is the sum of kl and lk);
 *               0 elsewhere;
 *   grad_pulse  knot pulse: dJ/dknot_k [n_knots]; Gaussian pulse: dJ/dA0 [1]; n_pulse its length.
 * Unsupported: general (distance) Coulomb interactions, L > 256. */
#define FPSFTWA_OBJ_DOUBLE_OCCUPANCY 0
typedef struct fpsftwa_adjoint_spec {
    int32_t objective;        /* FPSFTWA_OBJ_DOUBLE_OCCUPANCY */
    int32_t checkpoint_every; /* steps between checkpoints, 0 = 50 */
} fpsftwa_adjoint_spec;
int fpsftwa_run_adjoint(fpsftwa_handle* h, const fpsftwa_adjoint_spec* spec, double* objective, double* grad_u,
                        double* grad_t, double* grad_pulse, size_t n_pulse);

/* fps.cpp:88-220 for "method":"fTWA" (and "HF", :88-101, "GPSR", :208-220): read <path>, build the system,
 * run, and write <folder><sim_name>_fTWA_res.json (_HF_res.json, _GPSR_res.json) in the reference's results
 * format (src/FHSolver.cpp:171-307).  out_path overrides the output file name when not NULL. */
int fpsftwa_run_json_file(const char* path, const char* out_path);

/* The system and initial occupations the JSON front end builds from a simulation file
 * (FHSystem(Json, folder), src/FHSystem.cpp:123-326).  Arrays may be NULL; *n_sites is set
 * first, so call once with NULL arrays to size them. */
int fpsftwa_load_system_json(const char* path, int32_t* n_sites, int32_t* coulomb_type,
                             double* hopping, double* coulomb_u, double* coulomb_v, double* potential,
                             double* n_up, double* n_dn, double* positions);

/* The initial state the JSON front end builds: occupations n_up[L], n_dn[L] and the rotations
 * R_up, R_dn ([L][L] complex interleaved; the identity unless IC type is tbGroundState or
 * mfGroundState), as initiateSolver sets them (src/FHfTWASolver.cpp:99-153). */
int fpsftwa_load_initial_state_json(const char* path, int32_t* n_sites, double* n_up, double* n_dn,
                                    double* rot_up, double* rot_dn);

/* Eigenbasis of FHSystem::computeTBGroundState (mean_field 0, src/FHSystem.cpp:909-937: eigenvectors
 * of the real symmetric hopping[L][L], rot_dn = rot_up) or computeMFGroundState (mean_field 1,
 * :940-1019: eigenvectors of hopping + diag(U n_other) per spin, iterated by plain substitution until
 * the density residue < 1e-4 L, at most 100 iterations).  rot_* are real [L][L] row-major, columns
 * the eigenvectors in ascending eigenvalue order, each column's largest-magnitude entry positive.
 * energy_up / energy_dn ([L], may be NULL) receive the eigenvalues (tbgs_E for both spins, or the last
 * iteration's mfgs_Eu / mfgs_Ed), which the results file echoes as "energyVector" / "energyVectorUp" /
 * "energyVectorDown" (src/FHSystem.cpp:1143-1230).  Host-only (no GPU call). */
int fpsftwa_ground_state(int32_t n_sites, const double* hopping, const double* coulomb_u, int32_t n_up,
                         int32_t n_dn, int32_t mean_field, double* rot_up, double* rot_dn,
                         int32_t* iterations, double* residue, double* energy_up, double* energy_dn);

#ifdef __cplusplus
}
#endif
#endif /* FPS_FTWA_H */
