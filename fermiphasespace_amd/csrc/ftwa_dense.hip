// ftwa_dense.hip -- batched fTWA integrator for large or long-range lattices (L up to 1024).
//
// Same physics and records as the stencil engines (src/FHfTWASolver.cpp:366-426, :568-600, :629-661),
// organised around one matrix product per RK stage for the whole batch:
//
//   X = Y T  for every (trajectory, spin) at once,  T = hopping (+ pulse Peierls term), shared
//   k_ij = -i [ X_ij - conj(X_ji) + (h_j - h_i) Y_ij ]          (Y T - T Y = X - X^dagger)
//
// The product is a plain GEMM with the batch stacked into the rows, so it goes to the vendor library
// (rocBLAS, MFMA): rows are (trajectory, spin, i, plane) with the real and imaginary parts of row i
// stored as two planes, [b][s][i][plane][L].  Without a pulse T is real and X = Z T with Z the
// [4 B L x L] matrix of planes; with a pulse T = Tr + i Ti and [X_re X_im] = [Y_re Y_im] [[Tr, Ti],
// [-Ti, Tr]] with K = N = 2L.  Everything else is hand-written:
//   hartree kernel   h_s,i = pdiag_i + sum_j U_ij (Re rho_s',jj + 1/2) [+ V_ij (Re rho_s,jj + 1/2)]
//   update kernel    per 32x32 tile pair {(I,J), (J,I)}, I <= J: both X tiles through LDS, the RK4
//                    "3/8" registers (y, A, p) of the upper-triangle elements, next stage input
//                    written to both triangles (exactly Hermitian)
//   record kernel    per time point: diagonal, corner, energies (plain-path formulas)
#include <hip/hip_runtime.h>
#define ROCBLAS_BETA_FEATURES_API  // rocblas_gemm_ex_get_solutions: the GEMM solutions for the autotuner
#include <rocblas/rocblas.h>

#include <cstdint>
#include <type_traits>
#include <vector>

#include "ftwa_device.hpp"
#include "ftwa_kernels.hpp"

namespace fpsamd {

namespace {

using dev::wave_sum;

constexpr int kTile = 32;

template <typename Real>
__device__ __forceinline__ size_t row_off(int64_t b, int sg, int i, int L)  // start of row i (plane 0)
{
    return ((static_cast<size_t>(b) * 2 + sg) * L + i) * 2 * static_cast<size_t>(L);
}

// The stage input of the matrix-core kernel (L even): per slice, row blocks of 16 rows (the last one padded),
// each row [Re Y_i0 .. Re Y_i,L-1, Im Y_i0 .. Im Y_i,L-1] split into groups of 4 k, and within a row block
// the groups in k order, each group 16 rows x 4 k.  Element (i, k) sits at ki4(i, k, L); a 16-row by 16-k
// block is 1 KB contiguous, so a tile's reads and writes are whole cache lines.
__host__ __device__ inline size_t ki4_slice(int L) { return 2 * static_cast<size_t>(L) * ((L + 15) & ~15); }
__host__ __device__ inline uint32_t ki4(int i, int k, int L)
{
    return static_cast<uint32_t>(((i >> 4) * (L >> 1) + (k >> 2)) * 64 + (i & 15) * 4 + (k & 3));
}

// element (i, k) of slice (b, sg) of the stage input (k < L: Re Y_ik, L + k: Im Y_ik): rows of planes, or
// (kil) the layout of the matrix-core kernel
template <typename Real>
__device__ __forceinline__ size_t y_off(int64_t b, int sg, int i, int k, int L, int kil)
{
    return kil ? (static_cast<size_t>(b) * 2 + sg) * ki4_slice(L) + ki4(i, k, L) : row_off<Real>(b, sg, i, L) + k;
}

// interleaved [b][s][i][j] complex <-> planes [b][s][i][plane][j]; the pack also sets Ycur = y (y may be null:
// Ycur only)
template <typename Real>
__global__ void dense_pack_kernel(int L, int64_t B, const Real* inter, Real* y, Real* ycur)
{
    const int64_t n = B * 2 * static_cast<int64_t>(L) * L;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t bs = e / (static_cast<int64_t>(L) * L);
        const int ij = static_cast<int>(e - bs * L * L), i = ij / L, j = ij - i * L;
        const size_t o = (static_cast<size_t>(bs) * L + i) * 2 * L + j;
        const Real re = inter[2 * e], im = inter[2 * e + 1];
        if (y) {
            y[o] = re;
            y[o + L] = im;
        }
        ycur[o] = re;
        ycur[o + L] = im;
    }
}

template <typename Real>
__global__ void dense_unpack_kernel(int L, int64_t B, const Real* y, Real* inter)
{
    const int64_t n = B * 2 * static_cast<int64_t>(L) * L;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t bs = e / (static_cast<int64_t>(L) * L);
        const int ij = static_cast<int>(e - bs * L * L), i = ij / L, j = ij - i * L;
        // the upper triangle carries the state; the lower is its mirror
        const int a = i <= j ? i : j, c = i <= j ? j : i;
        const size_t o = (static_cast<size_t>(bs) * L + a) * 2 * L + c;
        inter[2 * e] = y[o];
        inter[2 * e + 1] = (i <= j ? Real(1) : Real(-1)) * y[o + L];
        if (i == j) inter[2 * e + 1] = Real(0);
    }
}

// the GEMM's right factor for stage amplitude a: real T [L][L] (no pulse) or the real 2L x 2L block
// form of T = Tr + i Ti (pulse); the diagonal of the hopping is carried by h
template <typename Real>
__global__ void dense_hop_kernel(DevSystem<Real> s, int64_t stage, Real* M)
{
    const int L = s.L;
    const bool pulse = s.amp != nullptr;
    const double a = pulse ? s.amp[stage] : 0.0;
    const int64_t n = static_cast<int64_t>(L) * L;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int k = static_cast<int>(e / L), l = static_cast<int>(e - static_cast<int64_t>(k) * L);
        const double t = k == l ? 0.0 : static_cast<double>(s.tM[e]);
        if (!pulse) {
            M[e] = static_cast<Real>(t);
            continue;
        }
        double sn = 0.0, cs = 0.0;
        if (t != 0.0) sincos(a * static_cast<double>(s.pd[e]), &sn, &cs);
        const Real tr = static_cast<Real>(t * (1.0 + cs)), ti = static_cast<Real>(t * sn);
        const int64_t L2 = 2 * static_cast<int64_t>(L);
        M[k * L2 + l] = tr;
        M[k * L2 + L + l] = ti;
        M[(L + k) * L2 + l] = -ti;
        M[(L + k) * L2 + L + l] = tr;
    }
}

// h of both spins of TB trajectories per workgroup (b0 = TB blockIdx), one thread per site i:
// h_up,i = pdiag_i + sum_j U_ij n_dn,j + V_ij n_up,j and h_dn,i with the spins exchanged (general Coulomb),
// pdiag_i + U_ii n_sbar,i (on-site), pdiag_i (none).  UT, VT: U and V transposed, so a wave reads a row
// of each per j, coalesced; the occupations n = Re rho_jj + 1/2 of the TB trajectories are LDS
// broadcasts.  The diagonal comes from dg [b][s][L] when given (written with the stage input by the
// matrix-core kernel), else from Ycur (layout kil as for y_off).  With ecl (general Coulomb; the stage input is a time
// point) also the energies E_U = sum_ij U_ij (n_up,i n_dn,j + n_dn,i n_up,j) / 2 and
// E_V = sum_ij V_ij (n_up,i n_up,j + n_dn,i n_dn,j) / 2 of the same sums, into ecl [b][2] (double).
template <typename Real, int TB, bool ECL>
__global__ void __launch_bounds__(256)
    dense_hartree_kernel(DevSystem<Real> s, int64_t B, const Real* UT, const Real* VT, const Real* dg,
                         const Real* ycur, int kil, Real* hv, double* ecl)
{
    __shared__ double red[4][TB][2];
    extern __shared__ unsigned char dsm[];
    Real* nd = reinterpret_cast<Real*>(dsm);  // [L][TB][2]
    const int L = s.L;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * TB;
    const int nb = static_cast<int>(B - b0 < TB ? B - b0 : TB);
    for (int e = threadIdx.x; e < TB * 2 * L; e += blockDim.x) {
        const int t = e / (2 * L), r = e - t * 2 * L, sg = r >= L ? 1 : 0, i = r - sg * L;
        Real v = 0;
        if (t < nb) v = (dg ? dg[(b0 + t) * 2 * L + r] : ycur[y_off<Real>(b0 + t, sg, i, i, L, kil)]) + Real(0.5);
        nd[(i * TB + t) * 2 + sg] = v;
    }
    __syncthreads();
    double eu[TB], ev[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) eu[t] = ev[t] = 0.0;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        Real hu[TB], hd[TB];
        const Real pd = s.pdiag[i];
#pragma unroll
        for (int t = 0; t < TB; ++t) hu[t] = hd[t] = 0;
        if (s.coulomb_type == 2 && !ECL) {
#pragma unroll 4
            for (int j = 0; j < L; ++j) {
                const Real u = UT[static_cast<size_t>(j) * L + i], v = VT[static_cast<size_t>(j) * L + i];
                const Real* nj = nd + j * TB * 2;
#pragma unroll
                for (int t = 0; t < TB; ++t) {
                    hu[t] += u * nj[2 * t + 1] + v * nj[2 * t];
                    hd[t] += u * nj[2 * t] + v * nj[2 * t + 1];
                }
            }
        } else if (s.coulomb_type == 2) {
            Real ud[TB], uu[TB], vu[TB], vd[TB];  // sum_j U_ij n_dn,j, U_ij n_up,j, V_ij n_up,j, V_ij n_dn,j
#pragma unroll
            for (int t = 0; t < TB; ++t) ud[t] = uu[t] = vu[t] = vd[t] = 0;
#pragma unroll 4
            for (int j = 0; j < L; ++j) {
                const Real u = UT[static_cast<size_t>(j) * L + i], v = VT[static_cast<size_t>(j) * L + i];
                const Real* nj = nd + j * TB * 2;
#pragma unroll
                for (int t = 0; t < TB; ++t) {
                    ud[t] += u * nj[2 * t + 1];
                    uu[t] += u * nj[2 * t];
                    vu[t] += v * nj[2 * t];
                    vd[t] += v * nj[2 * t + 1];
                }
            }
            const Real* ni = nd + i * TB * 2;
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                hu[t] = ud[t] + vu[t];
                hd[t] = uu[t] + vd[t];
                const double nu = static_cast<double>(ni[2 * t]), ndn = static_cast<double>(ni[2 * t + 1]);
                eu[t] += (nu * static_cast<double>(ud[t]) + ndn * static_cast<double>(uu[t])) / 2.0;
                ev[t] += (nu * static_cast<double>(vu[t]) + ndn * static_cast<double>(vd[t])) / 2.0;
            }
        } else if (s.coulomb_type == 1) {
            const Real u = s.u_diag[i];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                hu[t] = u * nd[(i * TB + t) * 2 + 1];
                hd[t] = u * nd[(i * TB + t) * 2];
            }
        }
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            if (t < nb) {
                hv[((b0 + t) * 2) * L + i] = pd + hu[t];
                hv[((b0 + t) * 2 + 1) * L + i] = pd + hd[t];
            }
        }
    }
    if (ECL && s.coulomb_type == 2) {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            const double su = wave_sum(eu[t]), sv = wave_sum(ev[t]);
            if (lane == 0) {
                red[w][t][0] = su;
                red[w][t][1] = sv;
            }
        }
        __syncthreads();
        if (threadIdx.x < 2 * nb) {
            const int t = threadIdx.x >> 1, k = threadIdx.x & 1;
            ecl[(b0 + t) * 2 + k] = red[0][t][k] + red[1][t][k] + red[2][t][k] + red[3][t][k];
        }
    }
}

// T = transpose of S, both [L][L]
template <typename Real>
__global__ void dense_transpose_kernel(int L, const Real* S, Real* T)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L * L; e += gridDim.x * blockDim.x) {
        const int i = e / L, j = e - i * L;
        T[static_cast<size_t>(j) * L + i] = S[e];
    }
}

// RK stage ST on the upper triangle.  Grid: (tile pairs x B, 2 spins); 256 threads.  The new stage input
// of tile (I, J) goes back through LDS (over the X tile (I, J), which only its own thread reads) so
// that the mirrored tile (J, I) is written row by row, coalesced, instead of one row per lane.
template <typename Real, int ST>
__global__ void __launch_bounds__(256)
    dense_update_kernel(int L, int npair, double hstep, const int2* pairs, const Real* X, const Real* hv, Real* y,
                        Real* A, Real* p, Real* ycur)
{
    __shared__ Real xa[2][kTile][kTile + 1];  // X tile (I, J), planes re/im; then the new stage input
    __shared__ Real xb[2][kTile][kTile + 1];  // X tile (J, I)
    const int64_t b = blockIdx.x / npair;
    const int2 pr = pairs[blockIdx.x - b * npair];
    const int sg = blockIdx.y;
    const int I0 = pr.x * kTile, J0 = pr.y * kTile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 per pass
    for (int r = ty; r < kTile; r += 8) {
        const int i = I0 + r, j = J0 + tx;
        if (i < L && j < L) {
            const size_t o = row_off<Real>(b, sg, i, L) + j;
            xa[0][r][tx] = X[o];
            xa[1][r][tx] = X[o + L];
        }
        const int i2 = J0 + r, j2 = I0 + tx;
        if (i2 < L && j2 < L) {
            const size_t o = row_off<Real>(b, sg, i2, L) + j2;
            xb[0][r][tx] = X[o];
            xb[1][r][tx] = X[o + L];
        }
    }
    __syncthreads();
    const Real h = static_cast<Real>(hstep);
    const Real h3 = static_cast<Real>(hstep / 3.0), h6 = static_cast<Real>(hstep / 6.0),
               h8 = static_cast<Real>(hstep / 8.0);
    const Real* hs = hv + (b * 2 + sg) * L;
    for (int r = ty; r < kTile; r += 8) {
        const int i = I0 + r, j = J0 + tx;
        if (i >= L || j >= L || i > j) continue;
        const size_t o = row_off<Real>(b, sg, i, L) + j;
        const Real xr = xa[0][r][tx], xi = xa[1][r][tx];    // X_ij
        const Real tr = xb[0][tx][r], ti = xb[1][tx][r];    // X_ji
        const Real yr = y[o], yi = y[o + L];
        Real Ar = 0, Ai = 0, pr_ = 0, pi = 0;
        if (ST > 0) {
            Ar = A[o];
            Ai = A[o + L];
            pr_ = p[o];
            pi = p[o + L];
        }
        Real cr, ci;  // current stage input
        if (ST == 0) {
            cr = yr;
            ci = yi;
        } else if (ST == 1) {
            cr = yr + h3 * pr_;
            ci = yi + h3 * pi;
        } else if (ST == 2) {
            cr = yr + h6 * (Ar - Real(3) * pr_);
            ci = yi + h6 * (Ai - Real(3) * pi);
        } else {
            cr = yr + h * pr_;
            ci = yi + h * pi;
        }
        // c = X_ij - conj(X_ji) + (h_j - h_i) Y_ij;  k = -i c
        const Real dh = hs[j] - hs[i];
        const Real cx = xr - tr + dh * cr, cy = xi + ti + dh * ci;
        Real kx = cy, ky = -cx;
        if (i == j) ky = 0;  // the diagonal slope is real
        Real nr, ni;
        if (ST == 0) {
            Ar = kx;
            Ai = ky;
            pr_ = kx;
            pi = ky;
            nr = yr + h3 * kx;
            ni = yi + h3 * ky;
        } else if (ST == 1) {
            Ar += Real(3) * kx;
            Ai += Real(3) * ky;
            pr_ -= kx;
            pi -= ky;
            nr = yr + h6 * (Ar - Real(3) * pr_);
            ni = yi + h6 * (Ai - Real(3) * pi);
        } else if (ST == 2) {
            Ar += Real(3) * kx;
            Ai += Real(3) * ky;
            pr_ += kx;
            pi += ky;
            nr = yr + h * pr_;
            ni = yi + h * pi;
        } else {
            Ar += kx;
            Ai += ky;
            nr = yr + h8 * Ar;
            ni = yi + h8 * Ai;
            y[o] = nr;
            y[o + L] = ni;
        }
        if (ST < 3) {
            A[o] = Ar;
            A[o + L] = Ai;
            p[o] = pr_;
            p[o + L] = pi;
        }
        ycur[o] = nr;
        ycur[o + L] = ni;
        xa[0][r][tx] = nr;
        xa[1][r][tx] = ni;
    }
    __syncthreads();
    // mirrored tile: Y_ji = conj(Y_ij) for i < j, row j = J0 + r, column i = I0 + tx
    for (int r = ty; r < kTile; r += 8) {
        const int jr = J0 + r, ic = I0 + tx;
        if (jr >= L || ic >= L || ic >= jr) continue;
        const size_t om = row_off<Real>(b, sg, jr, L) + ic;
        ycur[om] = xa[0][tx][r];
        ycur[om + L] = -xa[1][tx][r];
    }
}

// ---- fused product and update on the matrix cores, several (trajectory, spin) slices per workgroup ----
//
// One workgroup per (tile pair {(I, J), (J, I)} of 16 x 16 complex tiles, group of kGrpSlices slices),
// 4 waves of kWaveSlices slices each.  The stage's hopping operand M is the same for every slice, so the
// workgroup stages the columns of M it needs (those of tiles J and I, Re and Im with a pulse) through LDS in
// chunks of kKc k, double-buffered (one barrier per chunk), and every wave reads them back for each of its
// slices; the rows of the stage input are per slice and come straight from memory, in the row-block layout
// of ki4 (a 16-k block of a tile's 16 rows is one 1 KB load per wave; the pairs of one group run on one
// XCD, so after the first pair they are served from its L2).  v_mfma_f32_16x16x4f32: a 16 x 16
// output block per 4 k, lane l supplying row/column l % 16 at k = k0 + 4 (l / 16) + m to instruction m of a
// 16-k block from one 16-byte load (the same permutation of k in both factors leaves the product
// unchanged).  The accumulators of both tiles stay in registers; X_ji is brought to the lane holding X_ij
// through a small per-wave LDS transpose, the RK update follows, and the new stage input goes to the other
// buffer (its mirror through the same transpose, so rows are written contiguously).
// The update is the 3/8 rule of dense_update_kernel written around the stage input c (read from the stage
// input itself, which the products have just brought into the caches) so that each stage touches fewer
// registers: stage 0 keeps A = k1; stage 1 writes Y3 = c + h (k2 - 2/3 k1), p = 4/3 k1 - 2 k2 and
// A = k1 + 3 k2; stage 2 writes Y4 = c + h (p + k3) and A += 3 k3; stage 3 writes y + h/8 (A + k4).  The
// step result has the same single rounding at the size of y as dense_update_kernel's; only the stage
// inputs are formed from the previous one.
// With a pulse K = 2L: row i of the stage input is [Y_re Y_im] and M = [[Tr, Ti], [-Ti, Tr]] is symmetric
// (T is Hermitian), so Re X_ij = row i . M row j and Im X_ij = row i . M row L + j.  Without one K = L and
// M = T real symmetric: Re X = Y_re T, Im X = Y_im T.
constexpr int kMt = 16;           // complex tile edge
constexpr int kKc = 64;           // k per LDS chunk
constexpr int kKcPad = kKc + 4;   // row stride of the LDS chunk (17 16-byte units: conflict-free b128 reads)
constexpr int kWaveSlices = 2;    // (trajectory, spin) slices per wave
constexpr int kGrpSlices = 4 * kWaveSlices;
// y, A and p of the grouped kernel are tile-major: per slice and 16 x 16 tile pair, a Re plane then an Im
// plane of 256 floats, element (row 4 q + r, column c) of the tile at 64 r + 16 q + c.  One wave's access
// for one r is then 256 contiguous bytes instead of 64-byte pieces of four rows.
constexpr int kTileElems = 2 * kMt * kMt;

typedef float fps_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ fps_f32x4 ld4(const float* p) { return *reinterpret_cast<const fps_f32x4*>(p); }

__device__ __forceinline__ float ld1b(const float* base, uint32_t byte_off)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ void st1b(float* base, uint32_t byte_off, float v)
{
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// 16 bytes at a 32-bit byte offset from a wave-uniform base (the scalar-base form of the global load)
__device__ __forceinline__ fps_f32x4 ld4b(const float* base, uint32_t byte_off)
{
    return *reinterpret_cast<const fps_f32x4*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ fps_f32x4 keep4(bool on, fps_f32x4 v)  // v or zeros, per component (no branch)
{
    v.x = on ? v.x : 0.f;
    v.y = on ? v.y : 0.f;
    v.z = on ? v.z : 0.f;
    v.w = on ? v.w : 0.f;
    return v;
}

template <int ST, bool PULSE>
__global__ void __launch_bounds__(256)
    dense_mfma_group_kernel(int L, int64_t nbs, int npair, double hstep, const int2* pairs, const float* M,
                            const float* hv, const float* yin, float* y, float* A, float* p, float* yout, float* dg,
                            const float4* ew, double* epart)
{
    constexpr int NCOL = PULSE ? 4 * kMt : 2 * kMt;  // staged columns of M: J (Re, Im), I (Re, Im) | J, I
    constexpr int NA = PULSE ? 2 : 4;                // row streams per slice: I, J | I re, I im, J re, J im
    constexpr int NB = PULSE ? 4 : 2;                // staged column blocks
    constexpr int F4 = NCOL * kKc / 4;               // float4 per chunk of M
    static_assert(F4 % 256 == 0, "whole float4 per thread");
    constexpr int NLD = F4 / 256;
    static_assert(kKc == 64, "four 16-k blocks per chunk");
    __shared__ float ms[2][NCOL][kKcPad];
    __shared__ float xt[4][2][kMt][kMt + 1];  // per wave: transpose of one tile (re, im)
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: slice bases live in scalar registers
    const int K = PULSE ? 2 * L : L;
    const int q = lane >> 4, c16 = lane & 15;
    const int kq = 4 * (tid & (kKc / 4 - 1));  // k offset of this thread's float4 within a chunk of M
    const float h = static_cast<float>(hstep);
    const float h3 = static_cast<float>(hstep / 3.0), h8 = static_cast<float>(hstep / 8.0);
    const float c43 = 4.f / 3.f, c23 = 2.f / 3.f;
    float* ms0 = &ms[0][0][0];
    int mst[NLD];  // LDS float offset of this thread's float4 m within a chunk buffer
#pragma unroll
    for (int m = 0; m < NLD; ++m) mst[m] = ((tid + 256 * m) / (kKc / 4)) * kKcPad + kq;

    // ---- the item of this workgroup: (group of kGrpSlices slices, tile pair), group-major.  With a whole
    // multiple of 8 items, XCD x (workgroups b = x mod 8; dispatch is round-robin over the 8 XCDs) runs items
    // [x n/8, (x+1) n/8), so the pairs of a group run on one XCD and its stage input is served from that L2.
    const int64_t nitem = ((nbs + kGrpSlices - 1) / kGrpSlices) * npair;
    int64_t it = blockIdx.x;
    if ((nitem & 7) == 0) it = (it & 7) * (nitem >> 3) + (it >> 3);

    // ---- per-item state.  Off-diagonal pairs run blocks 0, 1 (rows I, columns J) and 2, 3 (rows J, columns I)
    // over the whole K.  A diagonal pair needs only X(I, I): blocks 0, 1 take k < kh and blocks 2, 3 the rest
    // (rows J = rows I, columns I = columns J, read from k + kh), and the halves are added at the end.  Stream
    // g (0: blocks 0, 1; 1: blocks 2, 3) is read at k + kofs[g] while k < kend[g].  Loads are unconditional from
    // clamped addresses, and what lies outside is replaced by zeros where the value is used (a conditional
    // 16-byte load compiles to a branch around a flat load).
    int I0, J0, kend0, kend1, klen;
    bool diag;
    uint32_t mp[NLD];  // byte offset in M of the staged row (= column, M symmetric) at k = kq (+ kofs)
    bool mon[NLD];     // column inside L (else row 0 is read and zeros are stored)
    int mend[NLD];     // loop k below which the column's values are inside K
    const float* ys[kWaveSlices];  // wave-uniform start of each slice's stage input
    int64_t bsl[kWaveSlices];
    bool son[kWaveSlices];
    uint32_t ap[NA];   // per-lane byte offsets of the row streams within a slice
    uint32_t eo[4], eoi[4];  // byte offsets of Re, Im of this lane's element r (row I0 + 4 q + r, column J0 + c16)
    uint32_t tp;       // byte offset of this lane's element r = 0 in the slice's tile-major RK registers
    int pix;           // tile pair index
    bool eon[4];       // element r in the upper triangle
#define FPS_GROUP_SETUP(ITEM)                                                                                 \
    {                                                                                                         \
        const int pair_ = static_cast<int>((ITEM) % npair);                                                   \
        const int64_t grp_ = (ITEM) / npair;                                                                  \
        const int2 pr_ = pairs[pair_];                                                                        \
        I0 = pr_.x * kMt;                                                                                     \
        J0 = pr_.y * kMt;                                                                                     \
        diag = pr_.x == pr_.y;                                                                                \
        const bool diag_ = diag;                                                                              \
        const int kh_ = min(((K / 2) + 15) & ~15, K); /* K < 16: the whole K on blocks 0, 1 */                \
        const int kofs1_ = diag_ ? kh_ : 0;                                                                   \
        kend0 = diag_ ? kh_ : K;                                                                              \
        kend1 = diag_ ? K - kh_ : K;                                                                          \
        klen = kend0;                                                                                         \
        _Pragma("unroll") for (int m = 0; m < NLD; ++m) {                                                     \
            const int f = tid + 256 * m, cc = f / (kKc / 4);                                                  \
            const int blkc = cc / kMt, c = cc % kMt;                                                          \
            const bool second = PULSE ? (blkc & 2) != 0 : blkc != 0; /* columns I: blocks 2, 3 */             \
            const int col = (second ? I0 : J0) + c;                                                           \
            mon[m] = col < L;                                                                                 \
            const int mr = PULSE && (blkc & 1) ? col + L : col;                                               \
            mp[m] = 4u * static_cast<uint32_t>((mon[m] ? mr : 0) * K + kq + (second ? kofs1_ : 0));           \
            mend[m] = second ? kend1 : kend0;                                                                 \
        }                                                                                                     \
        _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s) {                                             \
            const int64_t bs = grp_ * kGrpSlices + w * kWaveSlices + s;                                       \
            son[s] = bs < nbs;                                                                                \
            bsl[s] = son[s] ? bs : 0;                                                                         \
            ys[s] = yin + static_cast<size_t>(bsl[s]) * ki4_slice(L);                                          \
        }                                                                                                     \
        /* rows past L lie in the padded row block; they only feed outputs that are not stored */            \
        ap[0] = 4u * ki4(I0 + c16, 4 * q, L);                                                                 \
        if (PULSE) {                                                                                          \
            ap[1] = 4u * ki4(J0 + c16, 4 * q + kofs1_, L);                                                    \
        } else {                                                                                              \
            ap[1] = 4u * ki4(I0 + c16, L + 4 * q, L);                                                         \
            ap[NA - 2] = 4u * ki4(J0 + c16, 4 * q + kofs1_, L);                                               \
            ap[NA - 1] = 4u * ki4(J0 + c16, L + 4 * q + kofs1_, L);                                           \
        }                                                                                                     \
        tp = 4u * static_cast<uint32_t>(pair_ * kTileElems + lane);                                           \
        pix = pair_;                                                                                          \
        const int jc_ = J0 + c16;                                                                             \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                                       \
            const int i = I0 + 4 * q + r;                                                                     \
            eon[r] = i < L && jc_ < L && i <= jc_;                                                            \
            eo[r] = eon[r] ? 4u * ki4(i, jc_, L) : 0u;                                                        \
            eoi[r] = eon[r] ? 4u * ki4(i, L + jc_, L) : 0u;                                                   \
        }                                                                                                     \
    }

    // the first chunk of M and the first 16-k block of rows of the current item
    fps_f32x4 mreg[NLD];
    int mokb = 0;  // bit m: float4 m of the chunk in mreg lies inside its column
    fps_f32x4 an[kWaveSlices][NA];
    bool aok0, aok1;  // the prefetched rows of streams 0 and 1 lie inside K
#define FPS_GROUP_FIRST_LOADS()                                                                               \
    {                                                                                                         \
        mokb = 0;                                                                                             \
        _Pragma("unroll") for (int m = 0; m < NLD; ++m) {                                                     \
            const bool ok = kq < mend[m];                                                                     \
            mokb |= ok ? 1 << m : 0;                                                                          \
            mreg[m] = ld4b(M, mp[m] - (ok ? 0u : 4u * kq));                                                   \
        }                                                                                                     \
        aok0 = 4 * q < kend0;                                                                                 \
        aok1 = 4 * q < kend1;                                                                                 \
        _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s)                                               \
            _Pragma("unroll") for (int n = 0; n < NA; ++n)                                                    \
                an[s][n] = ld4b(ys[s], ap[n] - ((n < NA / 2 ? aok0 : aok1) ? 0u : 256u * q));                 \
    }

    // the RK registers and Hartree fields of this lane's elements of the current item, read ahead of its
    // products (elements outside the upper triangle read element (0, 0) and are not used)
    float rc[kWaveSlices][4][2], ry[kWaveSlices][4][2], rA[kWaveSlices][4][2], rp[kWaveSlices][4][2];
    float rhi[kWaveSlices][4], rhj[kWaveSlices];
    float4 rw[4];  // stage 3 with epart: the energy weights of this lane's elements (shared by the slices)
#define FPS_GROUP_PREFETCH_RK()                                                                               \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) rw[r] = float4{0.f, 0.f, 0.f, 0.f};                         \
    if (ST == 3 && epart) {                                                                                   \
        _Pragma("unroll") for (int r = 0; r < 4; ++r)                                                         \
            rw[r] = ew[eon[r] ? static_cast<size_t>(I0 + 4 * q + r) * L + J0 + c16 : 0];                      \
    }                                                                                                         \
    _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s) {                                                 \
        const size_t to = static_cast<size_t>(bsl[s]) * npair * kTileElems;                                   \
        const float* hs = hv + bsl[s] * L;                                                                    \
        rhj[s] = hs[J0 + c16 < L ? J0 + c16 : L - 1];                                                         \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                                       \
            const int i = I0 + 4 * q + r;                                                                     \
            rhi[s][r] = hs[i < L ? i : L - 1];                                                                \
            const uint32_t te = tp + 256u * r;                                                                \
            rc[s][r][0] = ld1b(ys[s], eo[r]);                                                                 \
            rc[s][r][1] = ld1b(ys[s], eoi[r]);                                                                \
            ry[s][r][0] = ry[s][r][1] = rA[s][r][0] = rA[s][r][1] = rp[s][r][0] = rp[s][r][1] = 0.f;          \
            if (ST == 3) {                                                                                    \
                ry[s][r][0] = ld1b(y + to, te);                                                               \
                ry[s][r][1] = ld1b(y + to, te + 4u * kMt * kMt);                                              \
            }                                                                                                 \
            if (ST > 0) {                                                                                     \
                rA[s][r][0] = ld1b(A + to, te);                                                               \
                rA[s][r][1] = ld1b(A + to, te + 4u * kMt * kMt);                                              \
            }                                                                                                 \
            if (ST == 2) {                                                                                    \
                rp[s][r][0] = ld1b(p + to, te);                                                               \
                rp[s][r][1] = ld1b(p + to, te + 4u * kMt * kMt);                                              \
            }                                                                                                 \
        }                                                                                                     \
    }

    // One 16-k block: take the prefetched rows (waiting for them here, a block after their loads), on the
    // first block of a chunk start the next chunk of M, prefetch the next block's rows, read the staged
    // columns, then the products.
#define FPS_GROUP_BLOCK(KB, MLOAD)                                                                          \
    {                                                                                                        \
        const int k0 = kc0 + 16 * (KB);                                                                      \
        __builtin_amdgcn_sched_barrier(0); /* this block's waits stay behind the previous block's products */ \
        fps_f32x4 a[kWaveSlices][NA];                                                                        \
        _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s)                                              \
            _Pragma("unroll") for (int n = 0; n < NA; ++n) a[s][n] = keep4(n < NA / 2 ? aok0 : aok1, an[s][n]); \
        if (MLOAD) {                                                                                         \
            mokb = 0;                                                                                        \
            _Pragma("unroll") for (int m = 0; m < NLD; ++m) {                                                \
                const bool ok = more && kc0 + kKc + kq < mend[m];                                            \
                mokb |= ok ? 1 << m : 0;                                                                     \
                mreg[m] = ld4b(M, mp[m] + (ok ? 4u * (kc0 + kKc) : -4u * kq));                              \
            }                                                                                                \
        }                                                                                                    \
        const int kn = k0 + 16 + 4 * q;                                                                      \
        aok0 = kn < kend0;                                                                                   \
        aok1 = kn < kend1;                                                                                   \
        _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s)                                              \
            _Pragma("unroll") for (int n = 0; n < NA; ++n)                                                   \
                an[s][n] = ld4b(ys[s], ap[n] + ((n < NA / 2 ? aok0 : aok1) ? 64u * (k0 + 16) : -256u * q));   \
        fps_f32x4 bm[NB];                                                                                    \
        _Pragma("unroll") for (int n = 0; n < NB; ++n) bm[n] = ld4(mb + n * kMt * kKcPad + 16 * (KB));      \
        __builtin_amdgcn_sched_barrier(0); /* the prefetches stay ahead of this block's products */          \
        _Pragma("unroll") for (int s = 0; s < kWaveSlices; ++s) {                                            \
            _Pragma("unroll") for (int n = 0; n < 4; ++n) {                                                  \
                const fps_f32x4 av = PULSE ? a[s][n >> 1] : a[s][n];                                         \
                const fps_f32x4 bv = PULSE ? bm[n] : bm[n >> 1];                                             \
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc[s][n], 0, 0, 0);            \
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc[s][n], 0, 0, 0);            \
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc[s][n], 0, 0, 0);            \
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc[s][n], 0, 0, 0);            \
            }                                                                                                \
        }                                                                                                    \
    }

    FPS_GROUP_SETUP(it)
    FPS_GROUP_FIRST_LOADS()
    FPS_GROUP_PREFETCH_RK()
    float(*tw)[kMt][kMt + 1] = xt[w];
    {
        // ---- the first chunk of M to LDS
#pragma unroll
        for (int m = 0; m < NLD; ++m)
            *reinterpret_cast<fps_f32x4*>(ms0 + mst[m]) = keep4(mon[m] && ((mokb >> m) & 1), mreg[m]);
        __syncthreads();
        fps_f32x4 acc[kWaveSlices][4];
#pragma unroll
        for (int s = 0; s < kWaveSlices; ++s)
#pragma unroll
            for (int n = 0; n < 4; ++n) acc[s][n] = fps_f32x4{0.f, 0.f, 0.f, 0.f};
        const int nch = (klen + kKc - 1) / kKc;
        for (int ch = 0; ch < nch; ++ch) {
            const int kc0 = ch * kKc;
            const bool more = ch + 1 < nch;
            const float* mb = ms0 + (ch & 1) * (NCOL * kKcPad) + c16 * kKcPad + 4 * q;
            const int nkb = klen - kc0 >= kKc ? 4 : (klen - kc0 + 15) / 16;  // fewer only in the last chunk
            FPS_GROUP_BLOCK(0, true)
            if (nkb > 1) FPS_GROUP_BLOCK(1, false)
            if (nkb > 2) FPS_GROUP_BLOCK(2, false)
            if (nkb > 3) FPS_GROUP_BLOCK(3, false)
            if (more) {
                float* dst = ms0 + ((ch + 1) & 1) * (NCOL * kKcPad);
#pragma unroll
                for (int m = 0; m < NLD; ++m)
                    *reinterpret_cast<fps_f32x4*>(dst + mst[m]) = keep4(mon[m] && ((mokb >> m) & 1), mreg[m]);
            }
            __syncthreads();
        }
        if (diag) {  // X(I, I) from the two halves of K, added in a fixed order
#pragma unroll
            for (int s = 0; s < kWaveSlices; ++s) {
                acc[s][0] += acc[s][2];
                acc[s][1] += acc[s][3];
                acc[s][2] = acc[s][0];
                acc[s][3] = acc[s][1];
            }
        }

        // ---- RK update.  Accumulator register r of a lane holds row 4 (lane / 16) + r, column lane % 16.
#pragma unroll
        for (int s = 0; s < kWaveSlices; ++s) {
            if (!son[s]) continue;
            const int64_t bs = bsl[s];
#pragma unroll
            for (int r = 0; r < 4; ++r) {  // blocks 2, 3 hold X(J, I) (X(I, I) for a diagonal pair)
                tw[0][4 * q + r][c16] = acc[s][2][r];
                tw[1][4 * q + r][c16] = acc[s][3][r];
            }
            __builtin_amdgcn_wave_barrier();
            float xjr[4], xji[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                xjr[r] = tw[0][c16][4 * q + r];
                xji[r] = tw[1][c16][4 * q + r];
            }
            __builtin_amdgcn_wave_barrier();
            const int j = J0 + c16;
            const float hj = rhj[s];
            const size_t so = static_cast<size_t>(bs) * ki4_slice(L);
            const size_t to = static_cast<size_t>(bs) * npair * kTileElems;
            double eks = 0.0, eps = 0.0;  // stage 3 with epart: this lane's share of the hopping energies
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = I0 + 4 * q + r;
                float nr = 0.f, ni = 0.f;
                if (eon[r]) {
                    const uint32_t o = eo[r], oi = eoi[r];
                    const uint32_t te = tp + 256u * r, tei = te + 4u * kMt * kMt;
                    const float xr = acc[s][0][r], xi = acc[s][1][r];  // X_ij
                    const float tr = xjr[r], ti = xji[r];              // X_ji
                    const float cr = rc[s][r][0], ci = rc[s][r][1];  // the stage input
                    float Ar = rA[s][r][0], Ai = rA[s][r][1], pr_ = rp[s][r][0], pi = rp[s][r][1];
                    // c = X_ij - conj(X_ji) + (h_j - h_i) Y_ij;  k = -i c
                    const float dh = hj - rhi[s][r];
                    const float cx = xr - tr + dh * cr, cy = xi + ti + dh * ci;
                    float kx = cy, ky = -cx;
                    if (i == j) ky = 0;  // the diagonal slope is real
                    if (ST == 0) {  // Y2 = y + h/3 k1
                        Ar = kx;
                        Ai = ky;
                        nr = cr + h3 * kx;
                        ni = ci + h3 * ky;
                    } else if (ST == 1) {  // Y3 = Y2 + h (k2 - 2/3 k1); p = 4/3 k1 - 2 k2 = (Y4 - Y3)/h - k3
                        pr_ = c43 * Ar - 2.f * kx;
                        pi = c43 * Ai - 2.f * ky;
                        nr = cr + h * (kx - c23 * Ar);
                        ni = ci + h * (ky - c23 * Ai);
                        Ar += 3.f * kx;
                        Ai += 3.f * ky;
                    } else if (ST == 2) {  // Y4 = Y3 + h (p + k3)
                        nr = cr + h * (pr_ + kx);
                        ni = ci + h * (pi + ky);
                        Ar += 3.f * kx;
                        Ai += 3.f * ky;
                    } else {  // y + h/8 (k1 + 3 k2 + 3 k3 + k4)
                        Ar += kx;
                        Ai += ky;
                        nr = ry[s][r][0] + h8 * Ar;
                        ni = ry[s][r][1] + h8 * Ai;
                        st1b(y + to, te, nr);
                        st1b(y + to, tei, ni);
                        if (epart) {
                            eks += static_cast<double>(rw[r].x) * nr;
                            eps += static_cast<double>(rw[r].y) * nr - static_cast<double>(rw[r].z) * ni;
                        }
                    }
                    if (ST < 3) {
                        st1b(A + to, te, Ar);
                        st1b(A + to, tei, Ai);
                    }
                    if (ST == 1) {
                        st1b(p + to, te, pr_);
                        st1b(p + to, tei, pi);
                    }
                    st1b(yout + so, o, nr);
                    st1b(yout + so, oi, ni);
                    if (i == j) dg[bs * L + i] = nr;  // the diagonal again, compact, for the Hartree fields
                }
                tw[0][4 * q + r][c16] = nr;
                tw[1][4 * q + r][c16] = ni;
            }
            if (ST == 3 && epart) {  // one partial sum per (slice, tile pair), summed by the record in order
                eks = wave_sum(eks);
                eps = wave_sum(eps);
                if (lane == 0) {
                    epart[(bs * npair + pix) * 2] = eks;
                    epart[(bs * npair + pix) * 2 + 1] = eps;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // mirrored tile: Y_ji = conj(Y_ij) for i < j; this lane writes row j = J0 + 4 q + r, column i = I0 + c16
            const int ic = I0 + c16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int jr = J0 + 4 * q + r;
                const float mr = tw[0][c16][4 * q + r], mi = tw[1][c16][4 * q + r];
                if (jr < L && ic < L && ic < jr) {
                    st1b(yout + so, 4u * ki4(jr, ic, L), mr);
                    st1b(yout + so, 4u * ki4(jr, L + ic, L), -mi);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
#undef FPS_GROUP_BLOCK
#undef FPS_GROUP_PREFETCH_RK
#undef FPS_GROUP_FIRST_LOADS
#undef FPS_GROUP_SETUP
}

// y of the grouped kernel from the interleaved state [b][s][i][j]: every element of the tile-major form,
// zero outside the upper triangle.  Grid-stride over nbs * npair * 256 complex elements.
__global__ void dense_tile_pack_kernel(int L, int64_t nbs, int npair, const int2* pairs, const float* inter,
                                       float* yt)
{
    const int64_t n = nbs * npair * (kMt * kMt);
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t t = e / (kMt * kMt);  // slice * npair + pair
        const int k = static_cast<int>(e - t * (kMt * kMt)), r = k >> 6, lane = k & 63;
        const int2 pr = pairs[t % npair];
        const int i = pr.x * kMt + 4 * (lane >> 4) + r, j = pr.y * kMt + (lane & 15);
        float re = 0.f, im = 0.f;
        if (i < L && j < L && i <= j) {
            const int64_t o = ((t / npair) * L + i) * static_cast<int64_t>(L) + j;
            re = inter[2 * o];
            im = inter[2 * o + 1];
        }
        yt[t * kTileElems + k] = re;
        yt[t * kTileElems + kMt * kMt + k] = im;
    }
}

// the stage input of the matrix-core kernel from the interleaved state [b][s][i][j] (zeros in the padding
// rows of the last row block)
__global__ void dense_pack_ki4_kernel(int L, int64_t nbs, const float* inter, float* yk)
{
    const int L16 = (L + 15) & ~15;
    const int64_t n = nbs * L16 * 2 * static_cast<int64_t>(L);
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t bs = e / (static_cast<int64_t>(L16) * 2 * L);
        const int ik = static_cast<int>(e - bs * L16 * 2 * L), i = ik / (2 * L), k = ik - i * 2 * L;
        const int j = k < L ? k : k - L;
        const float v = i < L ? inter[2 * ((bs * L + i) * static_cast<int64_t>(L) + j) + (k < L ? 0 : 1)] : 0.f;
        yk[bs * ki4_slice(L) + ki4(i, k, L)] = v;
    }
}

// the interleaved state from the upper triangle of that stage input (the lower is its mirror)
__global__ void dense_unpack_ki4_kernel(int L, int64_t nbs, const float* yk, float* inter)
{
    const int64_t n = nbs * static_cast<int64_t>(L) * L;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t bs = e / (static_cast<int64_t>(L) * L);
        const int ij = static_cast<int>(e - bs * L * L), i = ij / L, j = ij - i * L;
        const int a = i <= j ? i : j, c = i <= j ? j : i;
        const float* ys = yk + bs * ki4_slice(L);
        inter[2 * e] = ys[ki4(a, c, L)];
        inter[2 * e + 1] = i == j ? 0.f : (i < j ? 1.f : -1.f) * ys[ki4(a, L + c, L)];
    }
}

// Hopping of the stage, T_ik for the ELL entries [z][L] (k = nb_n(i)): tM_ik (1 + e^{i a pd_ik}) with a
// pulse, tM_ik without.
template <typename Real>
__global__ void dense_ell_table_kernel(DevSystem<Real> s, int64_t stage, Real* tt)
{
    const int n = s.z * s.L;
    const bool pulse = s.amp != nullptr;
    const double a = pulse ? s.amp[stage] : 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const double t = static_cast<double>(s.nbr_t[e]);
        double sn = 0.0, cs = 0.0;
        if (pulse) sincos(a * static_cast<double>(s.nbr_pd[e]), &sn, &cs);
        tt[2 * e] = static_cast<Real>(pulse ? t * (1.0 + cs) : t);
        tt[2 * e + 1] = static_cast<Real>(pulse ? t * sn : 0.0);
    }
}

// RK stage ST for short-range lattices without the product: X_ij = sum_n Y_{i,k_n(j)} conj(T_{j,k_n(j)})
// from full rows of the stage input staged in LDS (rows of tile I and of tile J), then the same update
// as dense_update_kernel.  The stage input is double-buffered (yin -> yout) because other workgroups
// still read rows this one rewrites.  Grid: (tile pairs x B, 2 spins); 256 threads.
template <typename Real, int ST>
__global__ void __launch_bounds__(256)
    dense_stencil_update_kernel(int L, int z, int npair, double hstep, const int2* pairs, const int* nbk,
                                const Real* ttab, const Real* hv, const Real* yin, Real* y, Real* A, Real* p,
                                Real* yout)
{
    extern __shared__ unsigned char dsm[];
    Real* rowsI = reinterpret_cast<Real*>(dsm);           // [32][2][L]
    Real* rowsJ = rowsI + kTile * 2 * L;                   // [32][2][L]
    Real* Ts = rowsJ + kTile * 2 * L;                      // [z][L][2]
    int* ks = reinterpret_cast<int*>(Ts + 2 * z * L);      // [z][L]
    const int64_t b = blockIdx.x / npair;
    const int2 pr = pairs[blockIdx.x - b * npair];
    const int sg = blockIdx.y;
    const int I0 = pr.x * kTile, J0 = pr.y * kTile;
    const int nI = min(kTile, L - I0), nJ = min(kTile, L - J0);
    for (int e = threadIdx.x; e < nI * 2 * L; e += blockDim.x) rowsI[e] = yin[row_off<Real>(b, sg, I0, L) + e];
    for (int e = threadIdx.x; e < nJ * 2 * L; e += blockDim.x) rowsJ[e] = yin[row_off<Real>(b, sg, J0, L) + e];
    for (int e = threadIdx.x; e < z * L; e += blockDim.x) {
        ks[e] = nbk[e];
        Ts[2 * e] = ttab[2 * e];
        Ts[2 * e + 1] = ttab[2 * e + 1];
    }
    __syncthreads();
    const Real h = static_cast<Real>(hstep);
    const Real h3 = static_cast<Real>(hstep / 3.0), h6 = static_cast<Real>(hstep / 6.0),
               h8 = static_cast<Real>(hstep / 8.0);
    const Real* hs = hv + (b * 2 + sg) * L;
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
        const int r = e >> 5, c = e & 31;
        const int i = I0 + r, j = J0 + c;
        if (i >= L || j >= L || i > j) continue;
        // X_ij from row i (tile I), X_ji from row j (tile J)
        const Real* ri = rowsI + r * 2 * L;
        const Real* rj = rowsJ + c * 2 * L;
        Real xr = 0, xi = 0, tr = 0, ti = 0;
        for (int n = 0; n < z; ++n) {
            const int kj = ks[n * L + j], ki = ks[n * L + i];
            const Real Tjr = Ts[2 * (n * L + j)], Tji = Ts[2 * (n * L + j) + 1];  // T_{j,kj}; T_{kj,j} = conj
            const Real Tir = Ts[2 * (n * L + i)], Tii = Ts[2 * (n * L + i) + 1];
            const Real ar = ri[kj], ai = ri[L + kj];   // Y_{i,kj}
            const Real br = rj[ki], bi = rj[L + ki];   // Y_{j,ki}
            xr += ar * Tjr + ai * Tji;                  // Y_{i,kj} conj(T_{j,kj})
            xi += ai * Tjr - ar * Tji;
            tr += br * Tir + bi * Tii;
            ti += bi * Tir - br * Tii;
        }
        const size_t o = row_off<Real>(b, sg, i, L) + j;
        const Real yr = y[o], yi = y[o + L];
        Real Ar = 0, Ai = 0, pr_ = 0, pi = 0;
        if (ST > 0) {
            Ar = A[o];
            Ai = A[o + L];
            pr_ = p[o];
            pi = p[o + L];
        }
        const Real cr = ri[j], ci = ri[L + j];  // current stage input Y_ij
        const Real dh = hs[j] - hs[i];
        const Real cx = xr - tr + dh * cr, cy = xi + ti + dh * ci;
        Real kx = cy, ky = -cx;
        if (i == j) ky = 0;
        Real nr, ni;
        if (ST == 0) {
            Ar = kx;
            Ai = ky;
            pr_ = kx;
            pi = ky;
            nr = yr + h3 * kx;
            ni = yi + h3 * ky;
        } else if (ST == 1) {
            Ar += Real(3) * kx;
            Ai += Real(3) * ky;
            pr_ -= kx;
            pi -= ky;
            nr = yr + h6 * (Ar - Real(3) * pr_);
            ni = yi + h6 * (Ai - Real(3) * pi);
        } else if (ST == 2) {
            Ar += Real(3) * kx;
            Ai += Real(3) * ky;
            pr_ += kx;
            pi += ky;
            nr = yr + h * pr_;
            ni = yi + h * pi;
        } else {
            Ar += kx;
            Ai += ky;
            nr = yr + h8 * Ar;
            ni = yi + h8 * Ai;
            y[o] = nr;
            y[o + L] = ni;
        }
        if (ST < 3) {
            A[o] = Ar;
            A[o + L] = Ai;
            p[o] = pr_;
            p[o + L] = pi;
        }
        yout[o] = nr;
        yout[o + L] = ni;
        if (i != j) {
            const size_t om = row_off<Real>(b, sg, j, L) + i;
            yout[om] = nr;
            yout[om + L] = -ni;
        }
    }
}

// dst[0, n) = src[0, n) by the whole workgroup, 16-byte accesses where both are 16-byte aligned
template <typename Real>
__device__ __forceinline__ void copy_block(Real* dst, const Real* src, int n)
{
    constexpr int V = 16 / sizeof(Real);
    int done = 0;
    if ((reinterpret_cast<uintptr_t>(dst) % 16) == 0 && (reinterpret_cast<uintptr_t>(src) % 16) == 0) {
        const int nv = n / V;
        for (int e = threadIdx.x; e < nv; e += blockDim.x)
            reinterpret_cast<float4*>(dst)[e] = reinterpret_cast<const float4*>(src)[e];
        done = nv * V;
    }
    for (int e = done + threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
}

// RK stage ST with one workgroup per (trajectory, spin) holding that spin's whole stage input in LDS
// (short-range lattices up to L ~ 140 in FP32): read once, X_ij and X_ji gathered from LDS for the
// thread's upper-triangle elements (at most QMAX each), a barrier, the RK update against HBM
// registers, the next stage input written in place (both triangles) and streamed out coalesced.
// TAB: neighbour tables staged in LDS (false: read through the caches, so two workgroups fit a CU)
template <typename Real, int ST, int QMAX, bool TAB>
__global__ void __launch_bounds__(1024)
    dense_wg_update_kernel(int L, int z, double hstep, const int* __restrict__ nbk, const Real* __restrict__ ttab,
                           const Real* hv, const Real* yin, Real* y, Real* A, Real* p, Real* yout)
{
    extern __shared__ unsigned char dsm[];
    Real* Ys = reinterpret_cast<Real*>(dsm);              // [L][2][L] as in HBM
    const Real* Ts = ttab;                                // [z][L][2]
    const int* ks = nbk;                                  // [z][L]
    const int64_t b = blockIdx.x >> 1;
    const int sg = blockIdx.x & 1;
    const size_t base = row_off<Real>(b, sg, 0, L);
    const int n2 = 2 * L * L;
    copy_block(Ys, yin + base, n2);
    if constexpr (TAB) {
        Real* Tl = Ys + 2 * L * L;
        int* kl = reinterpret_cast<int*>(Tl + 2 * z * L);
        for (int e = threadIdx.x; e < z * L; e += blockDim.x) {
            kl[e] = nbk[e];
            Tl[2 * e] = ttab[2 * e];
            Tl[2 * e + 1] = ttab[2 * e + 1];
        }
        Ts = Tl;
        ks = kl;
    }
    __syncthreads();
    const int NU = L * (L + 1) / 2;
    const Real* hs = hv + (b * 2 + sg) * L;
    int ii[QMAX], jj[QMAX];
    Real kx[QMAX], ky[QMAX];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int u = threadIdx.x + q * blockDim.x;
        ii[q] = -1;
        if (u >= NU) continue;
        // row i of the upper triangle (with diagonal): the largest i with i*(2L - i + 1)/2 <= u
        const float c = static_cast<float>(2 * L + 1);
        int i = static_cast<int>(0.5f * (c - sqrtf(fmaxf(c * c - 8.0f * static_cast<float>(u), 0.0f))));
        i = max(0, min(i, L - 1));
        while (i > 0 && i * (2 * L - i + 1) / 2 > u) --i;
        while (i + 1 < L && (i + 1) * (2 * L - i) / 2 <= u) ++i;
        const int j = i + (u - i * (2 * L - i + 1) / 2);
        ii[q] = i;
        jj[q] = j;
        const Real* ri = Ys + i * 2 * L;
        const Real* rj = Ys + j * 2 * L;
        Real xr = 0, xi = 0, tr = 0, ti = 0;
        for (int n = 0; n < z; ++n) {
            const int kj = ks[n * L + j], ki = ks[n * L + i];
            const Real Tjr = Ts[2 * (n * L + j)], Tji = Ts[2 * (n * L + j) + 1];
            const Real Tir = Ts[2 * (n * L + i)], Tii = Ts[2 * (n * L + i) + 1];
            const Real ar = ri[kj], ai = ri[L + kj];
            const Real br = rj[ki], bi = rj[L + ki];
            xr += ar * Tjr + ai * Tji;
            xi += ai * Tjr - ar * Tji;
            tr += br * Tir + bi * Tii;
            ti += bi * Tir - br * Tii;
        }
        const Real dh = hs[j] - hs[i];
        const Real cx = xr - tr + dh * ri[j], cy = xi + ti + dh * ri[L + j];
        kx[q] = cy;
        ky[q] = i == j ? Real(0) : -cx;
    }
    __syncthreads();  // every gather of this stage input is done before it is overwritten
    const Real h = static_cast<Real>(hstep);
    const Real h3 = static_cast<Real>(hstep / 3.0), h6 = static_cast<Real>(hstep / 6.0),
               h8 = static_cast<Real>(hstep / 8.0);
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        if (ii[q] < 0) continue;
        const int i = ii[q], j = jj[q];
        const size_t o = base + static_cast<size_t>(i) * 2 * L + j;
        const Real yr = y[o], yi = y[o + L];
        Real Ar = 0, Ai = 0, pr_ = 0, pi = 0;
        if (ST > 0) {
            Ar = A[o];
            Ai = A[o + L];
            pr_ = p[o];
            pi = p[o + L];
        }
        Real nr, ni;
        if (ST == 0) {
            Ar = kx[q];
            Ai = ky[q];
            pr_ = kx[q];
            pi = ky[q];
            nr = yr + h3 * kx[q];
            ni = yi + h3 * ky[q];
        } else if (ST == 1) {
            Ar += Real(3) * kx[q];
            Ai += Real(3) * ky[q];
            pr_ -= kx[q];
            pi -= ky[q];
            nr = yr + h6 * (Ar - Real(3) * pr_);
            ni = yi + h6 * (Ai - Real(3) * pi);
        } else if (ST == 2) {
            Ar += Real(3) * kx[q];
            Ai += Real(3) * ky[q];
            pr_ += kx[q];
            pi += ky[q];
            nr = yr + h * pr_;
            ni = yi + h * pi;
        } else {
            Ar += kx[q];
            Ai += ky[q];
            nr = yr + h8 * Ar;
            ni = yi + h8 * Ai;
            y[o] = nr;
            y[o + L] = ni;
        }
        if (ST < 3) {
            A[o] = Ar;
            A[o + L] = Ai;
            p[o] = pr_;
            p[o + L] = pi;
        }
        Ys[i * 2 * L + j] = nr;
        Ys[i * 2 * L + L + j] = ni;
        Ys[j * 2 * L + i] = nr;
        Ys[j * 2 * L + L + i] = -ni;
    }
    __syncthreads();
    copy_block(yout + base, Ys, n2);
}

// The pulse factors of the long-range energy at one time point, shared by every trajectory:
// et[i L + j] = (tM_ij cos(a pd_ij), tM_ij sin(a pd_ij)) (cos 1, sin 0 on the diagonal).
template <typename Real>
__global__ void dense_energy_table_kernel(DevSystem<Real> s, int64_t stage, double2* et)
{
    const int L = s.L;
    const double a = s.amp[stage];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L * L; e += gridDim.x * blockDim.x) {
        const int i = e / L, j = e - i * L;
        const double tt = static_cast<double>(s.tM[e]);
        double sn = 0.0, cs = 1.0;
        if (i != j) sincos(a * static_cast<double>(s.pd[e]), &sn, &cs);
        et[e] = double2{tt * cs, tt * sn};
    }
}

// The hopping energies' weights over the upper triangle at one time point, for the partial sums of the
// matrix-core kernel (rho is Hermitian, so (i, j) and (j, i) fold into one term):
// ew[i L + j] = (tM_ij + tM_ji, f_ij + f_ji, g_ij - g_ji) for i < j and (tM_ii, tM_ii, 0) on the diagonal,
// f + i g = tM e^{i a pd} (a = 0 without a pulse); the lower triangle is not read.
template <typename Real>
__global__ void dense_energy_weights_kernel(DevSystem<Real> s, int64_t stage, float4* ew)
{
    const int L = s.L;
    const double a = s.amp ? s.amp[stage] : 0.0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L * L; e += gridDim.x * blockDim.x) {
        const int i = e / L, j = e - i * L;
        if (i > j) continue;
        const double tij = static_cast<double>(s.tM[e]);
        if (i == j) {
            ew[e] = float4{static_cast<float>(tij), static_cast<float>(tij), 0.f, 0.f};
            continue;
        }
        const size_t et_ = static_cast<size_t>(j) * L + i;
        const double tji = static_cast<double>(s.tM[et_]);
        double sij = 0.0, cij = 1.0, sji = 0.0, cji = 1.0;
        if (s.amp) {
            sincos(a * static_cast<double>(s.pd[e]), &sij, &cij);
            sincos(a * static_cast<double>(s.pd[et_]), &sji, &cji);
        }
        ew[e] = float4{static_cast<float>(tij + tji), static_cast<float>(tij * cij + tji * cji),
                       static_cast<float>(tij * sij - tji * sji), 0.f};
    }
}

// records of one time point, one trajectory per workgroup (formulas of the plain path).  et: the table of
// dense_energy_table_kernel for this time point (long-range hopping with a pulse).  dgp (may be null): the
// diagonal [b][s][L] of Ycur, compact.  epart (may be null): the hopping energies' partial sums of the
// matrix-core kernel, [b][s][npart] x (kinetic, with the pulse), used instead of the rows of Ycur.  ecl: the
// general Coulomb energies [b][2] of dense_hartree_kernel for this time point (general Coulomb only).  kil: the
// layout of Ycur, as for y_off.
template <typename Real>
__global__ void dense_record_kernel(DevSystem<Real> s, int64_t stage, const Real* ycur, Real* rec, int R, int dens,
                                    int ci, int cj, const double2* et, const Real* dgp, const double* epart, int npart,
                                    const double* ecl, int kil)
{
    __shared__ double ew[4 * 4];
    extern __shared__ double dg[];  // [2][L] n_sigma,j = rho_sigma,jj + 1/2
    const int L = s.L;
    const int64_t b = blockIdx.x;
    Real* rr = rec + b * R;
    for (int e = threadIdx.x; e < 2 * L; e += blockDim.x) {
        const int sg = e >= L ? 1 : 0, i = e - sg * L;
        const Real d = dgp ? dgp[b * 2 * L + e] : ycur[y_off<Real>(b, sg, i, i, L, kil)];
        rr[e] = d;
        dg[e] = static_cast<double>(d) + 0.5;
    }
    __syncthreads();
    auto Y = [&](int sg, int i, int k) { return ycur[y_off<Real>(b, sg, i, k, L, kil)]; };
    if (threadIdx.x == 0) {
        rr[2 * L + 4] = Y(0, ci, cj);
        rr[2 * L + 5] = Y(0, ci, L + cj);
        rr[2 * L + 6] = Y(1, ci, cj);
        rr[2 * L + 7] = Y(1, ci, L + cj);
    }
    if (dens)  // block of 256 threads: 4 waves
        dev::record_density(rr + 2 * L + 8, L, threadIdx.x & 63, threadIdx.x >> 6, 4, [&](int row, int col) {
            const int sg = row >= L ? 1 : 0;
            return Y(sg, row - sg * L, col);
        });
    double eth = 0, eu = 0, ev = 0, epe = 0;
    const bool pulse = s.amp != nullptr;
    const double a = pulse ? s.amp[stage] : 0.0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) / 64;
    if (epart) {
        // the tiles' partial sums of both spins in a fixed order, and the 1 of n = rho + 1/2 on the diagonal
        const double* ep = epart + b * 2 * npart * 2;
        for (int e = threadIdx.x; e < 2 * npart; e += blockDim.x) {
            eth += ep[2 * e];
            if (pulse) epe += ep[2 * e + 1];
        }
        for (int i = threadIdx.x; i < L; i += blockDim.x) {
            const double tii = static_cast<double>(s.tM[static_cast<size_t>(i) * L + i]);
            eth += tii;
            if (pulse) epe += tii;
        }
    } else if (4 * s.z >= L) {
        // long-range hopping: rows of the dense tM, one wave per row and lanes along it, so the rows of
        // the stage input are read coalesced (one lane per row touched a cache line per neighbour and
        // fetched 17.9 GB per time point on the 196-site bilayer)
        for (int i = wv; i < L; i += nwv) {
            for (int j = lane; j < L; j += 64) {
                const double tt = static_cast<double>(s.tM[static_cast<size_t>(i) * L + j]);
                if (tt == 0.0) continue;
                const double nr =
                    static_cast<double>(Y(0, i, j)) + static_cast<double>(Y(1, i, j)) + (i == j ? 1.0 : 0.0);
                const double ni = static_cast<double>(Y(0, i, L + j)) + static_cast<double>(Y(1, i, L + j));
                eth += nr * tt;
                if (pulse) {
                    const double2 f = et[static_cast<size_t>(i) * L + j];
                    epe += nr * f.x - ni * f.y;
                }
            }
        }
    } else {
        for (int i = threadIdx.x; i < L; i += blockDim.x) {
            for (int n = 0; n <= s.z; ++n) {  // the neighbour list of i, then the diagonal entry
                const int j = n < s.z ? s.nbr_k[n * L + i] : i;
                const double tt = n < s.z ? static_cast<double>(s.nbr_t[n * L + i]) : static_cast<double>(s.tdiag[i]);
                if (tt == 0.0) continue;
                const double nr =
                    static_cast<double>(Y(0, i, j)) + static_cast<double>(Y(1, i, j)) + (i == j ? 1.0 : 0.0);
                const double ni = static_cast<double>(Y(0, i, L + j)) + static_cast<double>(Y(1, i, L + j));
                eth += nr * tt;
                if (pulse) {
                    double sn = 0.0, cs = 1.0;
                    if (n < s.z) sincos(a * static_cast<double>(s.nbr_pd[n * L + i]), &sn, &cs);
                    epe += nr * tt * cs - ni * tt * sn;
                }
            }
        }
    }
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        const double nu = dg[i], ndn = dg[L + i];
        epe += static_cast<double>(s.pot[i]) * (nu + ndn);
        if (s.coulomb_type == 1) eu += static_cast<double>(s.u_diag[i]) * nu * ndn;
    }
    if (s.coulomb_type == 2 && threadIdx.x == 0) {  // general Coulomb: from the Hartree kernel's sums
        eu += ecl[2 * b];
        ev += ecl[2 * b + 1];
    }
    eth = wave_sum(eth);
    eu = wave_sum(eu);
    ev = wave_sum(ev);
    epe = wave_sum(epe);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) / 64;
    if ((threadIdx.x & 63) == 0) {
        ew[w * 4 + 0] = eth;
        ew[w * 4 + 1] = eu;
        ew[w * 4 + 2] = ev;
        ew[w * 4 + 3] = epe;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0;
        for (int k = 0; k < nw; ++k) t += ew[k * 4 + threadIdx.x];
        rr[2 * L + threadIdx.x] = static_cast<Real>(t);
    }
}

unsigned grid_for(int64_t n, int bs)
{
    int64_t g = (n + bs - 1) / bs;
    if (g > 65535 * 16) g = 65535 * 16;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

rocblas_status gemm(rocblas_handle hb, int m, int n, int k, const float* A, int lda, const float* B, int ldb,
                    float* C, int ldc)
{
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm(hb, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero, C,
                         ldc);
}
rocblas_status gemm(rocblas_handle hb, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc)
{
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero, C,
                         ldc);
}

template <typename Real>
constexpr rocblas_datatype blas_type()
{
    return std::is_same<Real, float>::value ? rocblas_datatype_f32_r : rocblas_datatype_f64_r;
}

// C = A B (column-major, no transposes) with a chosen library solution (0: the library's default);
// with solution != 0 and list != nullptr only the solution list is returned (list_size in/out).
template <typename Real>
rocblas_status gemm_solution(rocblas_handle hb, int solution, int m, int n, int k, const Real* A, int lda,
                             const Real* B, int ldb, Real* C, int ldc, rocblas_int* list = nullptr,
                             rocblas_int* list_size = nullptr)
{
    const Real one = 1, zero = 0;
    const rocblas_datatype t = blas_type<Real>();
    if (list_size)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"  // a beta entry point of the library
        return rocblas_gemm_ex_get_solutions(hb, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, t,
                                             lda, B, t, ldb, &zero, C, t, ldc, C, t, ldc, t,
                                             rocblas_gemm_algo_solution_index, rocblas_gemm_flags_none, list,
                                             list_size);
#pragma clang diagnostic pop
    if (solution == 0) return gemm(hb, m, n, k, A, lda, B, ldb, C, ldc);
    return rocblas_gemm_ex(hb, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, t, lda, B, t, ldb,
                           &zero, C, t, ldc, C, t, ldc, t, rocblas_gemm_algo_solution_index, solution,
                           rocblas_gemm_flags_none);
}

}  // namespace

struct DenseCtx {
    rocblas_handle blas = nullptr;
    hipStream_t stream = nullptr;
    int solution = 0;        // GEMM solution index chosen by dense_tune_gemm (0: the library's default)
    bool fused_mfma = false;  // FP32: product and update fused on the matrix cores instead of the library GEMM
};

void dense_set_fused_mfma(DenseCtx* c, bool on)
{
    if (c) c->fused_mfma = on;
}

bool dense_mfma_fits(int L, bool pulse)
{
    // rows of the stage input and of M 16-byte aligned and K = 2L (pulse) or L a multiple of 4: whole 16-byte
    // loads of the row streams (without a pulse the imaginary plane starts L floats into a row)
    return L >= 4 && L <= 1024 && (pulse ? L % 2 == 0 : L % 4 == 0);
}

DenseCtx* dense_create(hipStream_t st)
{
    DenseCtx* c = new DenseCtx;
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
        delete c;
        return nullptr;
    }
    rocblas_set_stream(c->blas, st);
    c->stream = st;
    return c;
}

void dense_set_stream(DenseCtx* c, hipStream_t st)
{
    if (!c) return;
    rocblas_set_stream(c->blas, st);
    c->stream = st;
}

void dense_destroy(DenseCtx* c)
{
    if (!c) return;
    if (c->blas) rocblas_destroy_handle(c->blas);
    delete c;
}

template <typename Real>
bool dense_uses_gemm(const DevSystem<Real>& s)
{
    // short-range lattices skip the product: X comes from rows of the stage input staged in LDS
    const int L = s.L;
    const size_t lds_st = 2 * kTile * 2 * static_cast<size_t>(L) * sizeof(Real) +
                          static_cast<size_t>(s.z) * L * (2 * sizeof(Real) + sizeof(int));
    return !(4 * s.z < L && lds_st <= 160 * 1024);
}

// values in each of y, A, p, Ycur and X: 4 L^2 per trajectory as planes; the matrix-core kernel's stage
// input (row blocks padded to 16 rows) and its tile-major y, A and p can take more
static size_t dense_per(int L, int64_t B)
{
    const size_t nt16 = static_cast<size_t>((L + kMt - 1) / kMt);
    return std::max(4 * static_cast<size_t>(L) * (nt16 * kMt), nt16 * (nt16 + 1) / 2 * 2 * kTileElems) *
           static_cast<size_t>(B);
}

// Byte offsets of the dense path's buffers: y, A, p, Ycur, X (dense_per values each) and the GEMM factor
// M ((2L)^2) back to back, then 256-byte aligned: the energy tables (L^2 double2 and L^2 float4), the
// hopping energies' partial sums ([2B][tile pairs] x 2 doubles), the Coulomb energies ([B][2] doubles),
// h (2 L per trajectory), U and V
// transposed, the compact diagonal of the stage input (2 L per trajectory) and the tile pairs.
struct DenseLayout {
    size_t per, et, ew, ep, ecl, hv, ut, vt, dg, pairs, bytes;
};

template <typename Real>
DenseLayout dense_layout(int L, int64_t B)
{
    DenseLayout d;
    const size_t LL = static_cast<size_t>(L) * L;
    const int nt = (L + kTile - 1) / kTile, nt16 = (L + kMt - 1) / kMt;
    const size_t np = static_cast<size_t>(nt) * (nt + 1) / 2, np16 = static_cast<size_t>(nt16) * (nt16 + 1) / 2;
    d.per = dense_per(L, B);
    size_t o = (5 * d.per + 4 * LL) * sizeof(Real);
    auto take = [&](size_t bytes) {
        const size_t r = (o + 255) / 256 * 256;
        o = r + bytes;
        return r;
    };
    d.et = take(LL * sizeof(double2));
    d.ew = take(LL * sizeof(float4));
    d.ep = take(2 * static_cast<size_t>(B) * np16 * 2 * sizeof(double));
    d.ecl = take(2 * static_cast<size_t>(B) * sizeof(double));
    d.hv = take(2 * static_cast<size_t>(L) * B * sizeof(Real));
    d.ut = take(LL * sizeof(Real));
    d.vt = take(LL * sizeof(Real));
    d.dg = take(2 * static_cast<size_t>(L) * B * sizeof(Real));
    d.pairs = take((np + np16) * sizeof(int2));
    d.bytes = o;
    return d;
}

template <typename Real>
size_t dense_buffer_elems(int L, int64_t B)
{
    return (dense_layout<Real>(L, B).bytes + sizeof(Real) - 1) / sizeof(Real);
}

template <typename Real>
hipError_t dense_run(DenseCtx* c, const DevSystem<Real>& s, const RunShape& r, Real* buf, Real* state, Real* rec,
                     int t0, int t1, bool first, bool last)
{
    const int L = s.L;
    const int64_t B = r.B;
    const size_t per = dense_per(L, B);
    Real* y = buf;
    Real* A = y + per;
    Real* p = A + per;
    Real* ycur = p + per;
    Real* X = ycur + per;
    Real* M = X + per;
    const DenseLayout lay = dense_layout<Real>(L, B);
    char* bb = reinterpret_cast<char*>(buf);
    double2* et = reinterpret_cast<double2*>(bb + lay.et);
    float4* ewt = reinterpret_cast<float4*>(bb + lay.ew);
    double* epart = reinterpret_cast<double*>(bb + lay.ep);
    double* ecl = reinterpret_cast<double*>(bb + lay.ecl);
    Real* hv = reinterpret_cast<Real*>(bb + lay.hv);
    Real* UT = reinterpret_cast<Real*>(bb + lay.ut);
    Real* VT = reinterpret_cast<Real*>(bb + lay.vt);
    Real* dg = reinterpret_cast<Real*>(bb + lay.dg);
    int2* pairs = reinterpret_cast<int2*>(bb + lay.pairs);
    const int nt = (L + kTile - 1) / kTile, nt16 = (L + kMt - 1) / kMt;
    const int npair = nt * (nt + 1) / 2, npair16 = nt16 * (nt16 + 1) / 2;
    int2* pairs16 = pairs + npair;  // the 16 x 16 tile pairs of the grouped matrix-core kernel
    hipStream_t st = c->stream;
    // FP32 long-range form with the product fused on the matrix cores: y, A, p tile-major (the state is read
    // back from Ycur, which holds it in planes after every whole step)
    const bool tiled = std::is_same<Real, float>::value && c->fused_mfma && dense_mfma_fits(L, s.amp != nullptr) &&
                       dense_uses_gemm(s);
    if (first) {
        if (s.coulomb_type == 2) {
            hipLaunchKernelGGL(dense_transpose_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)),
                               dim3(256), 0, st, L, s.U, UT);
            hipLaunchKernelGGL(dense_transpose_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)),
                               dim3(256), 0, st, L, s.V, VT);
        }
        std::vector<int2> hp;
        hp.reserve(static_cast<size_t>(npair + npair16));
        for (int I = 0; I < nt; ++I)
            for (int J = I; J < nt; ++J) hp.push_back(int2{I, J});
        for (int I = 0; I < nt16; ++I)
            for (int J = I; J < nt16; ++J) hp.push_back(int2{I, J});
        hipError_t e = hipMemcpyAsync(pairs, hp.data(), hp.size() * sizeof(int2), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);  // hp goes out of scope
        if (e != hipSuccess) return e;
        if constexpr (std::is_same<Real, float>::value) {
            if (tiled) {
                hipLaunchKernelGGL(dense_pack_ki4_kernel, dim3(grid_for(per / 2, 256)), dim3(256), 0, st, L, 2 * B,
                                   state, ycur);
                hipLaunchKernelGGL(dense_tile_pack_kernel, dim3(grid_for(2 * B * npair16 * kMt * kMt, 256)), dim3(256),
                                   0, st, L, 2 * B, npair16, pairs16, state, y);
            }
        }
        if (!tiled)
            hipLaunchKernelGGL(dense_pack_kernel<Real>, dim3(grid_for(per / 2, 256)), dim3(256), 0, st, L, B, state, y,
                               ycur);
    }
    const bool pulse = s.amp != nullptr;
    const size_t lds_st = 2 * kTile * 2 * static_cast<size_t>(L) * sizeof(Real) +
                          static_cast<size_t>(s.z) * L * (2 * sizeof(Real) + sizeof(int));
    const bool stencil = !dense_uses_gemm(s);  // the LDS stencil form, whose tiles take lds_st bytes
    // the whole spin block in one workgroup's LDS when it fits (upper elements <= 8 per thread)
    const size_t lds_wg = 2 * static_cast<size_t>(L) * L * sizeof(Real) +
                          static_cast<size_t>(s.z) * L * (2 * sizeof(Real) + sizeof(int));
    const bool whole = stencil && lds_wg <= 160 * 1024 && L * (L + 1) / 2 <= 8 * 1024;
    const size_t lds_half = 2 * static_cast<size_t>(L) * L * sizeof(Real);
    const bool two = whole && 2 * lds_half <= 160 * 1024;  // tables through the caches, two workgroups per CU
    // the block alone fits (FP64 up to L = 101): tables through the caches, one workgroup per CU
    const bool alone = stencil && !whole && lds_half <= 160 * 1024 && L * (L + 1) / 2 <= 8 * 1024;
    const dim3 grid(static_cast<unsigned>(npair * B), 2);
    if (first && !pulse) {
        if (stencil)
            hipLaunchKernelGGL(dense_ell_table_kernel<Real>, dim3(grid_for(static_cast<int64_t>(s.z) * L, 256)),
                               dim3(256), 0, st, s, 0, M);
        else
            hipLaunchKernelGGL(dense_hop_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)), dim3(256), 0,
                               st, s, 0, M);
    }
    if (stencil) {
        static bool attr = false;  // one-time opt-in to more than 64 KB of dynamic LDS
        if (!attr) {
            const void* ks[12] = {
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 0, 8, false>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 1, 8, false>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 2, 8, false>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 3, 8, false>),
                                 reinterpret_cast<const void*>(dense_stencil_update_kernel<Real, 0>),
                                 reinterpret_cast<const void*>(dense_stencil_update_kernel<Real, 1>),
                                 reinterpret_cast<const void*>(dense_stencil_update_kernel<Real, 2>),
                                 reinterpret_cast<const void*>(dense_stencil_update_kernel<Real, 3>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 0, 8, true>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 1, 8, true>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 2, 8, true>),
                                 reinterpret_cast<const void*>(dense_wg_update_kernel<Real, 3, 8, true>)};
            for (const void* k : ks) {
                hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                if (e != hipSuccess) return e;
            }
            attr = true;
        }
    }
    // the stage input alternates between ycur and X in the stencil form (four swaps per step, so every
    // time point finds it in ycur)
    Real* yin = ycur;
    Real* yout = X;
    // trajectories per Hartree workgroup: occupations of up to 8 in 48 KB of LDS
    // Hartree fields: 8 trajectories per workgroup while their occupations fit 48 KB of LDS; the diagonal
    // from dg once the matrix-core kernel has written it
    const bool h8 = 8 * 2 * static_cast<size_t>(L) * sizeof(Real) <= 48 * 1024;
    bool dg_ok = false;
    auto hartree = [&](const Real* yst, double* ec) {
        const Real* d = dg_ok ? dg : nullptr;
        const int kil = tiled ? 1 : 0;
        const unsigned g8 = static_cast<unsigned>((B + 7) / 8), g1 = static_cast<unsigned>(B);
        const size_t l8 = 8 * 2 * L * sizeof(Real), l1 = 2 * L * sizeof(Real);
        if (h8 && ec)
            hipLaunchKernelGGL((dense_hartree_kernel<Real, 8, true>), g8, dim3(256), l8, st, s, B, UT, VT, d, yst, kil,
                               hv, ec);
        else if (h8)
            hipLaunchKernelGGL((dense_hartree_kernel<Real, 8, false>), g8, dim3(256), l8, st, s, B, UT, VT, d, yst,
                               kil, hv, ec);
        else if (ec)
            hipLaunchKernelGGL((dense_hartree_kernel<Real, 1, true>), g1, dim3(256), l1, st, s, B, UT, VT, d, yst, kil,
                               hv, ec);
        else
            hipLaunchKernelGGL((dense_hartree_kernel<Real, 1, false>), g1, dim3(256), l1, st, s, B, UT, VT, d, yst,
                               kil, hv, ec);
    };
    bool ep_ok = false;  // epart holds the hopping energies of this time point
    for (int t = t0; t < t1; ++t) {
        if (pulse && 4 * s.z >= L && !ep_ok)
            hipLaunchKernelGGL(dense_energy_table_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)),
                               dim3(256), 0, st, s, static_cast<int64_t>(t) * r.nsub * 4, et);
        // general Coulomb: the fields of the next stage (the time point's state) and the time point's
        // Coulomb energies from the same sums
        bool hv_ready = false;
        if (s.coulomb_type == 2) {
            hartree(ycur, ecl);
            hv_ready = true;
        }
        hipLaunchKernelGGL(dense_record_kernel<Real>, dim3(static_cast<unsigned>(B)), dim3(256),
                           2 * L * sizeof(double), st, s, static_cast<int64_t>(t) * r.nsub * 4, ycur,
                           rec + static_cast<int64_t>(t - t0) * B * r.R, r.R, r.dens, r.ci, r.cj, et,
                           dg_ok ? dg : nullptr, ep_ok ? epart : nullptr, npair16,
                           s.coulomb_type == 2 ? ecl : nullptr, tiled ? 1 : 0);
        ep_ok = false;
        if (t == r.T - 1) break;
        // the next time point's hopping energies as partial sums of the matrix-core kernel's last stage
        const bool want_ep = tiled && 4 * s.z >= L && t + 1 < t1;
        if (want_ep)
            hipLaunchKernelGGL(dense_energy_weights_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)),
                               dim3(256), 0, st, s, static_cast<int64_t>(t + 1) * r.nsub * 4, ewt);
        for (int sub = 0; sub < r.nsub; ++sub) {
            for (int stg = 0; stg < 4; ++stg) {
                const int64_t g = (static_cast<int64_t>(t) * r.nsub + sub) * 4 + stg;
                const bool erec = want_ep && sub == r.nsub - 1 && stg == 3;
                if (!(hv_ready && sub == 0 && stg == 0)) hartree(yin, nullptr);
                if (stencil) {
                    if (pulse)
                        hipLaunchKernelGGL(dense_ell_table_kernel<Real>,
                                           dim3(grid_for(static_cast<int64_t>(s.z) * L, 256)), dim3(256), 0, st, s, g,
                                           M);
#define FPS_STENCIL_STAGE(STG)                                                                                   \
    if (two || alone)                                                                                              \
        hipLaunchKernelGGL((dense_wg_update_kernel<Real, STG, 8, false>), dim3(static_cast<unsigned>(2 * B)),       \
                           dim3(1024), lds_half, st, L, s.z, r.h, s.nbr_k, M, hv, yin, y, A, p, yout);             \
    else if (whole)                                                                                                \
        hipLaunchKernelGGL((dense_wg_update_kernel<Real, STG, 8, true>), dim3(static_cast<unsigned>(2 * B)),        \
                           dim3(1024), lds_wg, st, L, s.z, r.h, s.nbr_k, M, hv, yin, y, A, p, yout);               \
    else                                                                                                           \
        hipLaunchKernelGGL((dense_stencil_update_kernel<Real, STG>), grid, dim3(256), lds_st, st, L, s.z, npair,    \
                           r.h, pairs, s.nbr_k, M, hv, yin, y, A, p, yout)
                    switch (stg) {
                        case 0: FPS_STENCIL_STAGE(0); break;
                        case 1: FPS_STENCIL_STAGE(1); break;
                        case 2: FPS_STENCIL_STAGE(2); break;
                        default: FPS_STENCIL_STAGE(3); break;
                    }
#undef FPS_STENCIL_STAGE
                    Real* tmp = yin;
                    yin = yout;
                    yout = tmp;
                    continue;
                }
                if (pulse)
                    hipLaunchKernelGGL(dense_hop_kernel<Real>, dim3(grid_for(static_cast<int64_t>(L) * L, 256)),
                                       dim3(256), 0, st, s, g, M);
                if constexpr (std::is_same<Real, float>::value) {
                    if (tiled) {
                        // product and update fused on the matrix cores; the stage input alternates
                        // between ycur and X (four swaps per step)
                        const int64_t ngrp = (2 * B + kGrpSlices - 1) / kGrpSlices;
                        const dim3 g2(static_cast<unsigned>(ngrp * npair16));
#define FPS_MFMA_STAGE(STG)                                                                                    \
    if (pulse)                                                                                                   \
        hipLaunchKernelGGL((dense_mfma_group_kernel<STG, true>), g2, dim3(256), 0, st, L, 2 * B, npair16, r.h,   \
                           pairs16, M, hv, yin, y, A, p, yout, dg, ewt, erec ? epart : nullptr);                \
    else                                                                                                         \
        hipLaunchKernelGGL((dense_mfma_group_kernel<STG, false>), g2, dim3(256), 0, st, L, 2 * B, npair16, r.h,  \
                           pairs16, M, hv, yin, y, A, p, yout, dg, ewt, erec ? epart : nullptr)
                        switch (stg) {
                            case 0: FPS_MFMA_STAGE(0); break;
                            case 1: FPS_MFMA_STAGE(1); break;
                            case 2: FPS_MFMA_STAGE(2); break;
                            default: FPS_MFMA_STAGE(3); break;
                        }
#undef FPS_MFMA_STAGE
                        Real* tmp = yin;
                        yin = yout;
                        yout = tmp;
                        dg_ok = true;
                        ep_ok = ep_ok || erec;
                        continue;
                    }
                }
                rocblas_status bs;
                if (!pulse) {
                    // rows (b, s, i, plane) of length L:  X = Z T, row-major -> column-major X^T = T^T Z^T
                    bs = gemm_solution<Real>(c->blas, c->solution, L, static_cast<int>(4 * B * L), L, M, L, ycur, L, X,
                                             L);
                } else {
                    bs = gemm_solution<Real>(c->blas, c->solution, 2 * L, static_cast<int>(2 * B * L), 2 * L, M, 2 * L,
                                             ycur, 2 * L, X, 2 * L);
                }
                if (bs != rocblas_status_success) return hipErrorLaunchFailure;
                switch (stg) {
                    case 0:
                        hipLaunchKernelGGL((dense_update_kernel<Real, 0>), grid, dim3(256), 0, st, L, npair, r.h,
                                           pairs, X, hv, y, A, p, ycur);
                        break;
                    case 1:
                        hipLaunchKernelGGL((dense_update_kernel<Real, 1>), grid, dim3(256), 0, st, L, npair, r.h,
                                           pairs, X, hv, y, A, p, ycur);
                        break;
                    case 2:
                        hipLaunchKernelGGL((dense_update_kernel<Real, 2>), grid, dim3(256), 0, st, L, npair, r.h,
                                           pairs, X, hv, y, A, p, ycur);
                        break;
                    default:
                        hipLaunchKernelGGL((dense_update_kernel<Real, 3>), grid, dim3(256), 0, st, L, npair, r.h,
                                           pairs, X, hv, y, A, p, ycur);
                        break;
                }
            }
        }
    }
    if (last) {
        if constexpr (std::is_same<Real, float>::value) {
            if (tiled)
                hipLaunchKernelGGL(dense_unpack_ki4_kernel, dim3(grid_for(per / 2, 256)), dim3(256), 0, st, L, 2 * B,
                                   ycur, state);
        }
        if (!tiled)
            hipLaunchKernelGGL(dense_unpack_kernel<Real>, dim3(grid_for(per / 2, 256)), dim3(256), 0, st, L, B, y,
                               state);
    }
    return hipGetLastError();
}

// ---- occupations of fixed bases (tbStates / mfStates) ----
//
// The reference sums diag(V^dagger rho V) over trajectories at each time point
// (src/FHfTWASolver.cpp:385-394).  The bases here are real (eigenvectors of real symmetric
// matrices) and rho is Hermitian, so (V^T rho V)_kk = sum_ij V_ik V_jk Re rho_ij: linear in rho, and
// the ensemble sum is diag(V^T S V) with S the ensemble sum of Re rho.  The records carry Re rho's
// upper triangle, launch_density_reduce forms S per time point, and here one FP64 GEMM per (basis,
// spin) takes every time point of a chunk at once: P = S V with the time points stacked into the
// rows, then M_k = sum_i V_ik P_ik.

namespace {

// full[t][i][j] = S_sg(min(i, j), max(i, j)) for nt time points
__global__ void basis_unpack_kernel(int L, int nt, int sg, const double* __restrict__ S, double* __restrict__ full)
{
    const int64_t LL = static_cast<int64_t>(L) * L;
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= nt * LL) return;
    const int t = static_cast<int>(idx / LL);
    const int e = static_cast<int>(idx - t * LL);
    const int i = e / L, j = e - (e / L) * L;
    const int a = i < j ? i : j, b = i < j ? j : i;
    const int P = L * (L + 1) / 2;
    full[idx] = S[static_cast<int64_t>(t) * 2 * P + sg * P + dev::packed_row(L, a) + b];
}

// out[k][t0 + t] (row stride T) = sum_i V[i][k] Pm[t][i][k]
__global__ void basis_diag_kernel(int L, int nt, int t0, int T, const double* __restrict__ V,
                                  const double* __restrict__ Pm, double* __restrict__ out)
{
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= static_cast<int64_t>(nt) * L) return;
    const int t = static_cast<int>(idx / L);
    const int k = static_cast<int>(idx - static_cast<int64_t>(t) * L);
    const double* pt = Pm + static_cast<int64_t>(t) * L * L;
    double m = 0.0;
    for (int i = 0; i < L; ++i) m += V[static_cast<int64_t>(i) * L + k] * pt[static_cast<int64_t>(i) * L + k];
    out[static_cast<int64_t>(k) * T + t0 + t] = m;
}

}  // namespace

size_t basis_work_doubles(int L, int nt) { return 2 * static_cast<size_t>(nt) * L * L; }

hipError_t basis_project(DenseCtx* c, int L, int nt, int t0, int T, int nb, const double* bases, const double* S,
                         double* work, double* out)
{
    const int64_t n = static_cast<int64_t>(nt) * L * L;
    double* full = work;
    double* prod = work + n;
    for (int b = 0; b < nb; ++b) {
        for (int sg = 0; sg < 2; ++sg) {
            const double* V = bases + (static_cast<int64_t>(b) * 2 + sg) * L * L;
            hipLaunchKernelGGL(basis_unpack_kernel, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, L, nt, sg, S,
                               full);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            // row-major P[nt L][L] = full[nt L][L] V[L][L], i.e. column-major P^T = V^T full^T
            if (gemm(c->blas, L, nt * L, L, V, L, full, L, prod, L) != rocblas_status_success)
                return hipErrorLaunchFailure;
            hipLaunchKernelGGL(basis_diag_kernel, dim3(grid_for(static_cast<int64_t>(nt) * L, 256)), dim3(256), 0,
                               c->stream, L, nt, t0, T, V, prod, out + (static_cast<int64_t>(b) * 2 + sg) * L * T);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace fpsamd

namespace fpsamd {

template <typename Real>
hipError_t dense_tune_gemm(DenseCtx* c, const DevSystem<Real>& s, int64_t B, Real* buf, int* best_solution,
                           double* best_ms, int* n_tried)
{
    // the stage product of dense_run for this batch, on the buffers a previous run left: every
    // solution the library offers for the shape is timed (one warm-up call, then three), and the
    // fastest is kept for later runs of this handle
    const int L = s.L;
    const size_t per = dense_per(L, B);
    Real* ycur = buf + 3 * per;
    Real* X = ycur + per;
    Real* M = X + per;
    const bool pulse = s.amp != nullptr;
    const int m = pulse ? 2 * L : L, k = m;
    const int n = static_cast<int>(pulse ? 2 * B * L : 4 * B * L);
    rocblas_int count = 0;
    if (gemm_solution<Real>(c->blas, 1, m, n, k, M, m, ycur, k, X, m, nullptr, &count) != rocblas_status_success)
        count = 0;
    std::vector<rocblas_int> list(static_cast<size_t>(count));
    if (count > 0 && gemm_solution<Real>(c->blas, 1, m, n, k, M, m, ycur, k, X, m, list.data(), &count) !=
                         rocblas_status_success)
        count = 0;
    list.resize(static_cast<size_t>(count));
    list.insert(list.begin(), 0);  // the default first: a solution must beat it
    hipEvent_t e0, e1;
    hipError_t he = hipEventCreate(&e0);
    if (he != hipSuccess) return he;
    he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return he;
    }
    int best = 0, tried = 0;
    double best_t = -1.0;
    for (rocblas_int sol : list) {
        if (gemm_solution<Real>(c->blas, sol, m, n, k, M, m, ycur, k, X, m) != rocblas_status_success) continue;
        (void)hipEventRecord(e0, c->stream);
        bool ok = true;
        for (int rep = 0; rep < 3 && ok; ++rep)
            ok = gemm_solution<Real>(c->blas, sol, m, n, k, M, m, ycur, k, X, m) == rocblas_status_success;
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess || !ok) continue;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ++tried;
        if (best_t < 0.0 || ms / 3.0 < best_t) {
            best_t = ms / 3.0;
            best = sol;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->solution = best;
    if (best_solution) *best_solution = best;
    if (best_ms) *best_ms = best_t;
    if (n_tried) *n_tried = tried;
    return hipGetLastError();
}

template size_t dense_buffer_elems<float>(int, int64_t);
template size_t dense_buffer_elems<double>(int, int64_t);
template hipError_t dense_run<float>(DenseCtx*, const DevSystem<float>&, const RunShape&, float*, float*, float*, int,
                                     int, bool, bool);
template hipError_t dense_run<double>(DenseCtx*, const DevSystem<double>&, const RunShape&, double*, double*, double*,
                                      int, int, bool, bool);
template hipError_t dense_tune_gemm<float>(DenseCtx*, const DevSystem<float>&, int64_t, float*, int*, double*, int*);
template bool dense_uses_gemm<float>(const DevSystem<float>&);
template bool dense_uses_gemm<double>(const DevSystem<double>&);
template hipError_t dense_tune_gemm<double>(DenseCtx*, const DevSystem<double>&, int64_t, double*, int*, double*, int*);
}  // namespace fpsamd
