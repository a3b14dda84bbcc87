// ftwa_kernels.hpp -- host-side interface to the gfx950 fTWA kernels (ftwa_kernels.hip).
//
// Data layout in HBM (one handle):
//   state  [B][2 spins][L][L] complex<Real>, row-major, interleaved (re, im)
//   rec    [Tc][B][R] Real, R = 2L + 8 per-trajectory records of one time point:
//            [0,L) rho_up_ii  [L,2L) rho_dn_ii  2L+0..3 E_tH, E_U, E_V, E_pe  2L+4..7 re/im of rho_up(ci, cj), rho_dn(ci, cj)
//            and, when basis occupations are requested, R = 2L + 8 + L(L+1): [2L+8, R) Re rho_sigma(i, j),
//            j >= i, spin-major, rows packed (dev::packed_row); GPSR records end with 2L more values,
//            Im rho_up_ii, Im rho_dn_ii (GPSR's rho is not Hermitian; the density part then holds the
//            symmetric part (Re rho_ij + Re rho_ji) / 2)
//   part   [Tc][nchunk][Q][2] double   partial ensemble sums (sum, compensation)
//   acc    [T][Q][2] double            ensemble sums (sum, compensation), Q = 6L + 12:
//            per site i: S_u, S_d, S_uu, S_dd, S_ud, S_(nu nd)^2 at [6i .. 6i+5];
//            6L + k (k<4): S_Ek;  6L + 4 + k: S_Ek^2;  6L + 8 + k (k<4): corner sums
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpsamd {

// Neighbour list of one site as one aligned record, so a wave-uniform site index reads it with a
// single scalar load: k[n] the neighbour, t[n] the hopping tM_ik (unused entries: the site, 0).
template <typename Real, int Z>
struct alignas(Z * (sizeof(int) + sizeof(Real)) > 32 ? 64 : 32) RowNbrs {
    int k[Z];
    Real t[Z];
};

// Device view of the static system.  Neighbour lists are ELL with z entries per site,
// stored entry-major ([n][L]) so consecutive sites are consecutive addresses; unused entries
// point at the site itself with zero hopping.
template <typename Real>
struct DevSystem {
    int L = 0;
    int z = 0;
    int coulomb_type = 0;          // 0 null, 1 onSite, 2 general
    const int* nbr_k = nullptr;    // [z][L]  neighbour k of site i (tM_ik != 0, k != i)
    const Real* nbr_t = nullptr;   // [z][L]  tM_ik
    const int* nbr_ks = nullptr;   // [z][L]  the same neighbours, entry n at the n-th most common shift
    const Real* nbr_ts = nullptr;  // [z][L]  k - i mod 32 where the site has one (the direct-commutator kernel)
    const Real* nbr_pd = nullptr;  // [z][L]  pol . (r_i - r_k)  (pulse phase per unit amplitude)
    const Real* u_diag = nullptr;  // [L]     U_ii
    const Real* U = nullptr;       // [L][L]  dense U (general)
    const Real* V = nullptr;       // [L][L]  dense V (general)
    const Real* pdiag = nullptr;   // [L]     diagonal of Delta besides the Hartree term: pot + tM_ii
                                   //         (+ tM_ii again from the pulse potential when present)
    const Real* tdiag = nullptr;   // [L]     tM_ii (energy bookkeeping)
    const Real* pot = nullptr;     // [L]     constant diagonal potential (plain path: h and E_pe)
    const Real* epot = nullptr;    // [L]     diagonal of the potential in E_pe (batched paths): pot, plus
                                   //         tM_ii when a pulse is present
    const double* amp = nullptr;   // [T*nsub*4] pulse amplitude a(tau) at each RK stage time, or null
    const Real* tM = nullptr;      // [L][L]  dense hopping (plain path)
    const Real* pd = nullptr;      // [L][L]  pol . (r_i - r_j) dense (plain path, pulse)
    const void* rows = nullptr;    // [L] RowNbrs<Real, z> (batched path)
    int hop_uniform = 0;           // 1: every ELL entry (padding included) holds the hopping hop_value
    Real hop_value = 0;
    int fock = 0;                  // 1: Hartree-Fock dynamics, Delta_sigma gains -V o N_sigma (plain path)
    int gpsr = 0;                  // 1: GPSR dynamics: normally ordered rho (no 1/2), Delta_sigma = tM +
                                   //    diag(U rho_sbar_ii) + potential (no V), diffusion every step
    double gc_re = 0.0, gc_im = 0.0;  // GPSR diffusion constant sqrt(i U(0,0)) (0: no diffusion)
};

struct RunShape {
    int64_t B = 0;        // trajectories in this launch
    int64_t b0 = 0;       // global index of the first trajectory (noise counter)
    int T = 0;            // time points of the whole run
    int nsub = 1;
    double h = 0.0;       // dt / nsub
    int R = 0;            // record width
    int dens = 0;         // 1: records also carry the packed real density matrices (R = 2L + 8 + L(L+1))
    int ci = 0, cj = 0;   // the off-diagonal element rho(ci, cj) recorded per time point: the wholeDMatrix
                          // corner (0, L-1), or the core-hole element (site, L-1)
    uint64_t seed = 0;    // the run's noise key (initial-state noise; GPSR diffusion noise)
    const double* sample_occ = nullptr;  // [2][L] occupations: when set, the launch starting at time point 0
                                         // samples the initial state itself instead of reading it (kernels
                                         // that support it: herm_seq; launch_sample otherwise)
};

// Sampling: rho_sigma = diag(n - 1/2) + S.W (rotation applied separately; GPSR: diag(n), no 1/2).  noise
// (optional, FP64 [B][2][npairs][2]) replaces the counter-based generator.
template <typename Real>
hipError_t launch_sample(const DevSystem<Real>& s, const double* n_up, const double* n_dn, const double* noise,
                         uint64_t seed, int64_t b0, int64_t B, Real* state, hipStream_t st);

// Export of the generator's noise (FP64, layout as above).
hipError_t launch_noise(int L, uint64_t seed, int64_t b0, int64_t B, double* out, hipStream_t st);

// rho <- R rho R^dagger per trajectory and spin (R FP64 complex [2][L][L]).
template <typename Real>
hipError_t launch_rotate(int L, const double* rot, int64_t B, Real* state, Real* scratch, hipStream_t st);

// Batched LDS stencil integrator: time points [t0, t1) for B trajectories, recording each.
// Returns hipErrorInvalidValue when the shape does not fit the kernel family.
template <typename Real>
hipError_t launch_stencil(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                          int nt_hint, hipStream_t st);
// Neighbour-list width the batched kernel is compiled for (4 or 8); 0 when z is too large.
inline int stencil_padded_z(int z) { return z <= 4 ? 4 : (z <= 8 ? 8 : 0); }
// Whether the stencil kernel supports this L (LDS capacity) and its chosen block size.
template <typename Real>
bool stencil_fits(int L, int z, int* nt);

// FP32 batched integrator on the upper triangle (X = T rho exchanged through LDS); L <= 64, z <= 8.
bool herm_fits(int L, int z);
hipError_t launch_herm(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                       int nt_hint, hipStream_t st);
// Spin-pipelined variant of the upper-triangle kernel (no pulse; null or on-site Coulomb; L <= 64).
bool herm_pipe_fits(const DevSystem<float>& s);
hipError_t launch_herm_pipe(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                            hipStream_t st);

// Spin-sequential upper-triangle integrator (ftwa_herm_seq.hip), FP64 and FP32: L <= 64, z 4 or 8, no
// pulse, null or on-site Coulomb, no density records.  nt_hint 1024 selects 16 waves per workgroup,
// otherwise 8.
template <typename Real>
bool herm_seq_fits(const DevSystem<Real>& s);
template <typename Real>
hipError_t launch_herm_seq(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                           int nt_hint, hipStream_t st);

// Its FP32 form for 64 < L <= 107 (ftwa_herm_wide.hip): only Y in LDS, each element's commutator formed
// directly from its rows; z 4 or 8, no pulse, null or on-site Coulomb, no density records; samples the
// initial state in a run's first launch like the spin-sequential kernel.
bool herm_wide_fits(const DevSystem<float>& s);
// The same kernel where it beats the spin-sequential kernel for L <= 64 (L = 64, z = 4, one hopping value).
bool herm_wide_direct_preferred(const DevSystem<float>& s);
hipError_t launch_herm_wide(const DevSystem<float>& s, const RunShape& r, int t0, int t1, float* state, float* rec,
                            hipStream_t st);

// Batched GPSR integrator (ftwa_gpsr.hip), FP64 and FP32: full (non-Hermitian) matrices in LDS, RK4 "3/8"
// drift plus the diffusion increment every step; L <= 32, z 4 or 8, no pulse.  noise holds
// gpsr_noise_elems(L, B, nsub, t1 - t0) Reals: the chunk's diffusion noise, drawn by the launch itself.
template <typename Real>
bool gpsr_fits(const DevSystem<Real>& s);
template <typename Real>
size_t gpsr_noise_elems(int L, int64_t B, int nsub, int nt);
template <typename Real>
hipError_t launch_gpsr(const DevSystem<Real>& s, const RunShape& r, int t0, int t1, Real* state, Real* rec,
                       Real* noise, hipStream_t st);

// Dense batched path (ftwa_dense.hip): X = Y T per RK stage for the whole batch -- a library GEMM followed
// by the hand-written commutator/RK update kernel, or (FP32, on request, where L allows) one hand-written
// kernel fusing product and update on the matrix cores -- plus hand-written Hartree and record kernels.
// L <= 1024, any neighbour count.
struct DenseCtx;
DenseCtx* dense_create(hipStream_t st);
void dense_set_stream(DenseCtx* c, hipStream_t st);
void dense_destroy(DenseCtx* c);
// FP32: fuse product and update on the matrix cores (where dense_mfma_fits) instead of the library GEMM
void dense_set_fused_mfma(DenseCtx* c, bool on);
// whether the fused FP32 kernel runs this system (L % 4 == 0 with a pulse, L % 8 == 0 without)
bool dense_mfma_fits(int L, bool pulse);
// elements of Real the dense path needs for B trajectories (one buffer)
template <typename Real>
size_t dense_buffer_elems(int L, int64_t B);
// time points [t0, t1), records into rec; first packs `state` into the buffer's planes, last unpacks
// the final state back to `state`
template <typename Real>
hipError_t dense_run(DenseCtx* c, const DevSystem<Real>& s, const RunShape& r, Real* buf, Real* state, Real* rec,
                     int t0, int t1, bool first, bool last);

// Plain FP64 path: one launch per time point, dense Delta, global-memory products.
hipError_t launch_plain(const DevSystem<double>& s, const RunShape& r, int t, double* state, double* work,
                        double* rec, hipStream_t st);
// Per-trajectory FP64 scratch of the plain path: stage input, k1..k4, Delta, the GPSR diffusion
// increment (2 spins each), h and the GPSR step noise.
__host__ __device__ inline size_t plain_work_doubles(int L)
{
    const size_t LL = static_cast<size_t>(L) * L;
    return 7 * 2 * LL * 2 + 8 * static_cast<size_t>(L);
}

// Discrete adjoint (ftwa_adjoint.hip).  Work area per trajectory: 10 + seg_steps + 1 complex matrices
// of both spins; gradient accumulators per trajectory: [U_ii (L)][tM_kk (L)][ELL tM (z L)][dJ/da per
// RK stage (n_stage, 0 without a pulse)].
template <typename Real>
size_t adjoint_work_bytes(int L, int64_t B, int64_t seg_steps);
// Pulse hopping of every RK stage, [n_stage][z][L] complex (ttab is null without a pulse).
template <typename Real>
hipError_t launch_adjoint_table(const DevSystem<Real>& s, int64_t n_stage, Real* ttab, hipStream_t st);
template <typename Real>
hipError_t launch_adjoint_forward(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int64_t n0,
                                  int64_t n1, Real* state, Real* work, hipStream_t st);
template <typename Real>
hipError_t launch_adjoint_terminal(const DevSystem<Real>& s, int64_t B, const Real* state, Real* work, double* jb,
                                   hipStream_t st);
template <typename Real>
hipError_t launch_adjoint_backward(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int64_t n0,
                                   int64_t n1, const Real* ckpt, Real* segbuf, Real* work, double* gacc,
                                   int64_t n_stage, hipStream_t st);
hipError_t launch_adjoint_reduce(const double* src, int64_t B, int64_t n, double* out, hipStream_t st);
// LDS-resident form of the adjoint: threads per workgroup (0 when it does not apply: then the form above
// runs) and slots per thread; want_nt forces a workgroup size (0: the smallest that fits).  Per-trajectory
// buffers hold nt * slots complex values each: ckpt [nseg][B], seg [B][K], lam [B].  mode 0 runs the
// forward pass (checkpoints every K steps, the objective into jb), mode 1 reverses the segment [n0, n1).
template <typename Real>
int adjoint_lds_threads(int L, int z, int* slots, int want_nt = 0);
template <typename Real>
hipError_t launch_adjoint_lds(const DevSystem<Real>& s, const Real* ttab, const RunShape& r, int want_nt, int mode,
                              int64_t n0, int64_t n1, int64_t nst, int64_t K, int64_t n_stage, const Real* state,
                              Real* ckpt, Real* seg, Real* lam, double* gacc, double* jb, hipStream_t st);

// Ensemble reduction of records for time points [t0, t0+nt) into acc.  im_off >= 0: GPSR records, whose
// imaginary diagonals start at im_off (the sums then take the complex diagonals, see reduce_partial_kernel).
template <typename Real>
hipError_t launch_reduce(int L, int64_t B, int R, int nt, int t0, const Real* rec, double* part, int nchunk,
                         double* acc, hipStream_t st, int im_off = -1);
int reduce_chunks(int64_t B);

// Ensemble sums S[t][L(L+1)] (FP64) of the records' density part for nt time points (records with
// RunShape::dens); part holds nt * nchunk * L(L+1) * 2 doubles.
template <typename Real>
hipError_t launch_density_reduce(int L, int64_t B, int R, int nt, const Real* rec, double* part, int nchunk,
                                 double* S, hipStream_t st);

// Whether dense_run takes the long-range form (one GEMM per stage) rather than the LDS stencil form.
template <typename Real>
bool dense_uses_gemm(const DevSystem<Real>& s);

// Times every library solution of the dense path's stage GEMM for this batch (buffers of a previous
// dense_run) and keeps the fastest for the context's later runs; uses the long-range (GEMM) form.
template <typename Real>
hipError_t dense_tune_gemm(DenseCtx* c, const DevSystem<Real>& s, int64_t B, Real* buf, int* best_solution,
                           double* best_ms, int* n_tried);

// Occupations of nb fixed bases, M[b][sg][k] = sum_ij V_ik V_jk S_sg(i, j) = diag(V^T S V) with V =
// bases[b][sg] (real [L][L], columns the basis states), for nt time points of density sums S (as
// above), written to out[((b * 2 + sg) * L + k) * T + t0 + t].  work holds basis_work_doubles(L, nt).
hipError_t basis_project(DenseCtx* c, int L, int nt, int t0, int T, int nb, const double* bases, const double* S,
                         double* work, double* out);
size_t basis_work_doubles(int L, int nt);

}  // namespace fpsamd
