// ftwa_api.cpp -- the C ABI of include/fps_ftwa.h: handle, device buffers, run orchestration.
//
// One handle = one device + one HIP stream.  fpsftwa_run restates FHfTWASolver::runSimulation
// (src/FHfTWASolver.cpp:222-556): sample every trajectory (:342-362), integrate the time grid
// recording observables at each time point (:366-426), reduce the ensemble (:438-482) and keep the
// sums for the reference's normalisation (:489-551), which fpsftwa_get_result applies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/fps_ftwa.h"
#include "ftwa_kernels.hpp"
#include "ftwa_internal.hpp"

using namespace fpsamd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(FPSFTWA_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Device buffer owned by a handle.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n ? n : 1);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

bool is_gfx950(int dev, std::string* name)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    if (name) *name = prop.gcnArchName;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

struct fpsftwa_handle {
    int device = 0;
    hipStream_t stream = nullptr;      // the stream every call enqueues on: own_stream or the caller's
    hipStream_t own_stream = nullptr;  // created with the handle
    // system (host copy)
    int L = 0;
    int coulomb_type = 0;
    std::vector<double> tM, U, V, pot, disp;
    int pulse_type = 0;
    double A0 = 0, t0 = 0, td = 1, w0 = 0, phi = 0;
    double pol[3] = {0, 0, 0};
    std::vector<double> knots;
    // parameters
    double dt = 0.01, t_max = 2.0;
    int nsub = 1;
    int64_t B = 10;
    uint64_t seed = 0;
    int engine = FPSFTWA_ENGINE_FP64_BATCHED;
    uint32_t flags = FPSFTWA_RES_DIAG | FPSFTWA_RES_ENERGIES;
    std::vector<double> n_up, n_dn, rot_up, rot_dn;
    int n_bases = 0;
    std::vector<double> bases;  // [n_bases][2][L][L] real
    int rec_i = 0, rec_j = -1;  // recorded off-diagonal element (j < 0: L - 1)
    int method = FPSFTWA_METHOD_FTWA;
    int method_run = FPSFTWA_METHOD_FTWA;  // the method of the completed run (result normalisation)
    std::vector<double> noise;
    int nt_hint = 0;
    int time_chunk = 0;
    int kernel = FPSFTWA_KERNEL_AUTO;
    bool used_herm = false;
    bool used_dense = false;
    DenseCtx* dense = nullptr;  // GEMM library context of the dense path, created on first use
    // results
    bool has_result = false;
    int T = 0;
    int64_t B_run = 0;
    std::vector<double> acc;  // [T][Q][2]
    std::vector<double> basis_acc;  // [n_bases][2][L][T] ensemble sums of diag(V^T rho V)
    int n_bases_run = 0;
    int64_t launches = 0;
    double kernel_ms = 0.0, run_ms = 0.0;
    // device buffers
    DevBuf sys_buf, state, rec, part, acc_d, work, noise_d, rot_d, scratch, occ_d, dense_buf;
    DevBuf bases_d, dens_part, dens_S, basis_work, basis_out;  // basis occupations
    DevBuf gnoise;                                              // GPSR diffusion noise of a time chunk
    bool state_double = true;
    bool graph = false;           // capture a run's integrator and reduction launches in one HIP graph
    bool graph_used = false;      // the last run launched its chunks as a graph
    hipGraph_t graph_g = nullptr;  // the last run's graph and executable, released after the run
    hipGraphExec_t graph_x = nullptr;
};

// Work memory of fpsftwa_run_noalloc: system tables, occupations, the trajectory records of a time chunk
// and their partial ensemble sums.  One copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or copy (on the host it is a stand-in, not a device address): it
// uses the per-device address that scratch_address() obtained from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[FPSFTWA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on device dev, looked up once per device with dev current (the caller
// has made it current).
int scratch_address(int dev, void** out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(FPSFTWA_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FPSFTWA_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return FPSFTWA_OK;
}

int check_handle(fpsftwa_handle* h)
{
    if (!h) return fail(FPSFTWA_EINVAL, "null handle");
    return FPSFTWA_OK;
}

// The neighbour entries of every site (ELL form of tM's off-diagonal non-zeros), z entries per site with
// z the batched kernels' compiled width (4 or 8 where it applies); -1 marks an empty entry.
// In index order (by_shift false) the entries of a site are its neighbours in column order, empty ones
// last.  By shift, entry n of every site holds, where the site has one, the neighbour at the n-th most
// common shift k - i modulo 32 (ties: the smaller shift first); the others fill the remaining entries in
// index order.  The direct-commutator kernel, whose rows are 1 mod 32 complex long, then reads entry n of
// every lane's site at one bank offset from the lane's own element, so a wave's neighbour reads are
// conflict-free except where a neighbour wraps around a periodic boundary (on the 8x8 square: entries by
// direction, the x wraps excepted; in index order the directions of sites on the boundary rows and
// columns land in different entries).
std::vector<std::vector<int>> neighbour_entries(const fpsftwa_handle* h, int& z, bool by_shift)
{
    const int L = h->L;
    std::vector<std::vector<int>> nb(L);
    z = 0;
    for (int i = 0; i < L; ++i) {
        for (int k = 0; k < L; ++k)
            if (k != i && h->tM[(size_t)i * L + k] != 0.0) nb[i].push_back(k);
        z = std::max<int>(z, (int)nb[i].size());
    }
    z = std::max(z, 1);
    if (const int zp = stencil_padded_z(z)) z = zp;  // the batched kernel's compiled width
    if (!by_shift) {
        for (int i = 0; i < L; ++i) nb[i].resize(z, -1);
        return nb;
    }
    std::array<int, 32> hist{};
    for (int i = 0; i < L; ++i)
        for (int k : nb[i]) ++hist[((k - i) % 32 + 32) % 32];
    std::vector<int> shifts(32);
    for (int d = 0; d < 32; ++d) shifts[d] = d;
    std::stable_sort(shifts.begin(), shifts.end(), [&](int a, int b) { return hist[a] > hist[b]; });
    for (int i = 0; i < L; ++i) {
        std::vector<int> left = nb[i], slot(z, -1);
        for (int n = 0; n < z && n < 32 && hist[shifts[n]] > 0; ++n)
            for (size_t m = 0; m < left.size(); ++m)
                if (((left[m] - i) % 32 + 32) % 32 == shifts[n]) {
                    slot[n] = left[m];
                    left.erase(left.begin() + (std::ptrdiff_t)m);
                    break;
                }
        size_t m = 0;
        for (int n = 0; n < z && m < left.size(); ++n)
            if (slot[n] < 0) slot[n] = left[m++];
        nb[i] = slot;
    }
    return nb;
}

// Upload the static system in precision Real into one device allocation and fill the view.
// dst (capacity cap bytes), when given, receives the tables instead of the handle's own buffer; *used is
// set to the bytes taken there.
template <typename Real>
int upload_system(fpsftwa_handle* h, DevSystem<Real>& ds, int T, char* dst = nullptr, size_t cap = 0,
                  size_t* used = nullptr)
{
    const int L = h->L;
    int z = 0;
    const std::vector<std::vector<int>> nb = neighbour_entries(h, z, false), nbs = neighbour_entries(h, z, true);
    const double* pol = h->pol;
    auto pdot = [&](int i, int k) {
        if (h->disp.empty()) return 0.0;
        const double* d = &h->disp[((size_t)i * L + k) * 3];
        return pol[0] * d[0] + pol[1] * d[1] + pol[2] * d[2];
    };
    std::vector<int> nk((size_t)z * L), nks((size_t)z * L);
    std::vector<Real> nt((size_t)z * L), npd((size_t)z * L), nts((size_t)z * L);
    for (int i = 0; i < L; ++i)
        for (int n = 0; n < z; ++n) {
            const bool on = nb[i][n] >= 0;
            const int k = on ? nb[i][n] : i;
            nk[(size_t)n * L + i] = k;
            nt[(size_t)n * L + i] = on ? (Real)h->tM[(size_t)i * L + k] : Real(0);
            npd[(size_t)n * L + i] = on ? (Real)pdot(i, k) : Real(0);
            const int ks = nbs[i][n] >= 0 ? nbs[i][n] : i;
            nks[(size_t)n * L + i] = ks;
            nts[(size_t)n * L + i] = nbs[i][n] >= 0 ? (Real)h->tM[(size_t)i * L + ks] : Real(0);
        }
    // per-site records for the batched kernel (z is 4 or 8 there; other widths use the plain path)
    std::vector<unsigned char> rows;
    auto pack_rows = [&](auto rec_tag) {
        using RecT = decltype(rec_tag);
        constexpr int ZR = sizeof(RecT{}.k) / sizeof(int);
        rows.assign(sizeof(RecT) * L, 0);
        for (int i = 0; i < L; ++i) {
            RecT rr{};
            for (int n = 0; n < ZR; ++n) {
                rr.k[n] = nk[(size_t)n * L + i];
                rr.t[n] = nt[(size_t)n * L + i];
            }
            std::memcpy(rows.data() + sizeof(RecT) * i, &rr, sizeof(RecT));
        }
    };
    if (z == 4) pack_rows(RowNbrs<Real, 4>{});
    else if (z == 8) pack_rows(RowNbrs<Real, 8>{});
    std::vector<Real> udiag(L), pdiag(L), tdiag(L), potv(L), epot(L), Ud((size_t)L * L), Vd((size_t)L * L), tMd((size_t)L * L),
        pdd((size_t)L * L);
    for (int i = 0; i < L; ++i) {
        udiag[i] = h->coulomb_type ? (Real)h->U[(size_t)i * L + i] : Real(0);
        const double p = h->pot.empty() ? 0.0 : h->pot[i];
        // with a pulse the diagonal hopping also enters the pulse potential (t_ii e^{i A.0} = t_ii,
        // src/FHExternalPotentialPulse.cpp:96-100): twice in h, once more in E_pe
        const double tii = h->tM[(size_t)i * L + i];
        potv[i] = (Real)p;
        epot[i] = (Real)(p + (h->pulse_type != 0 ? tii : 0.0));
        tdiag[i] = (Real)tii;
        pdiag[i] = (Real)(p + (h->pulse_type != 0 ? 2.0 : 1.0) * tii);
    }
    for (size_t e = 0; e < (size_t)L * L; ++e) {
        Ud[e] = h->U.empty() ? Real(0) : (Real)h->U[e];
        Vd[e] = h->V.empty() ? Real(0) : (Real)h->V[e];
        tMd[e] = (Real)h->tM[e];
        pdd[e] = (Real)pdot((int)(e / L), (int)(e % L));
    }
    // pulse amplitude at every RK stage time: tau = g*hstep + c_s*hstep, c = (0, 1/3, 2/3, 1)
    std::vector<double> amp;
    if (h->pulse_type != 0) {
        const double hs = h->dt / h->nsub;
        const int64_t n = (int64_t)T * h->nsub;
        amp.resize((size_t)n * 4);
        PulseSpec pp{h->pulse_type, h->A0, h->t0, h->td, h->w0, h->phi, h->knots.data(), (int)h->knots.size(), h->t_max};
        const double cs[4] = {0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0};
        for (int64_t g = 0; g < n; ++g)
            for (int s = 0; s < 4; ++s) amp[(size_t)g * 4 + s] = pulse_amplitude(pp, (double)g * hs + cs[s] * hs);
    }
    // one allocation, 256-byte aligned sections
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_nk = take(nk.size() * sizeof(int)), o_nt = take(nt.size() * sizeof(Real)),
                 o_nks = take(nks.size() * sizeof(int)), o_nts = take(nts.size() * sizeof(Real)),
                 o_npd = take(npd.size() * sizeof(Real)), o_ud = take(L * sizeof(Real)),
                 o_pd = take(L * sizeof(Real)), o_td = take(L * sizeof(Real)), o_pot = take(L * sizeof(Real)), o_epot = take(L * sizeof(Real)),
                 o_U = take(Ud.size() * sizeof(Real)), o_V = take(Vd.size() * sizeof(Real)),
                 o_tM = take(tMd.size() * sizeof(Real)), o_pdd = take(pdd.size() * sizeof(Real)),
                 o_amp = take(std::max<size_t>(amp.size(), 1) * sizeof(double)),
                 o_rows = take(std::max<size_t>(rows.size(), 1));
    char* base = dst;
    if (dst) {
        if (off > cap)
            return fail(FPSFTWA_EINVAL, "the system tables (" + std::to_string(off) + " B) exceed the " +
                                            std::to_string(cap) + " B left in the fixed scratch");
        if (used) *used = off;
    } else {
        HIP_TRY(h->sys_buf.ensure(off));
        base = h->sys_buf.as<char>();
    }
    auto up = [&](size_t o, const void* src, size_t bytes) {
        return hipMemcpyAsync(base + o, src, bytes, hipMemcpyHostToDevice, h->stream);
    };
    HIP_TRY(up(o_nk, nk.data(), nk.size() * sizeof(int)));
    HIP_TRY(up(o_nt, nt.data(), nt.size() * sizeof(Real)));
    HIP_TRY(up(o_nks, nks.data(), nks.size() * sizeof(int)));
    HIP_TRY(up(o_nts, nts.data(), nts.size() * sizeof(Real)));
    HIP_TRY(up(o_npd, npd.data(), npd.size() * sizeof(Real)));
    HIP_TRY(up(o_ud, udiag.data(), L * sizeof(Real)));
    HIP_TRY(up(o_pd, pdiag.data(), L * sizeof(Real)));
    HIP_TRY(up(o_td, tdiag.data(), L * sizeof(Real)));
    HIP_TRY(up(o_pot, potv.data(), L * sizeof(Real)));
    HIP_TRY(up(o_epot, epot.data(), L * sizeof(Real)));
    HIP_TRY(up(o_U, Ud.data(), Ud.size() * sizeof(Real)));
    HIP_TRY(up(o_V, Vd.data(), Vd.size() * sizeof(Real)));
    HIP_TRY(up(o_tM, tMd.data(), tMd.size() * sizeof(Real)));
    HIP_TRY(up(o_pdd, pdd.data(), pdd.size() * sizeof(Real)));
    if (!amp.empty()) HIP_TRY(up(o_amp, amp.data(), amp.size() * sizeof(double)));
    if (!rows.empty()) HIP_TRY(up(o_rows, rows.data(), rows.size()));
    HIP_TRY(hipStreamSynchronize(h->stream));  // host vectors go out of scope
    ds.L = L;
    ds.z = z;
    ds.coulomb_type = h->coulomb_type;
    ds.nbr_k = reinterpret_cast<const int*>(base + o_nk);
    ds.nbr_t = reinterpret_cast<const Real*>(base + o_nt);
    ds.nbr_ks = reinterpret_cast<const int*>(base + o_nks);
    ds.nbr_ts = reinterpret_cast<const Real*>(base + o_nts);
    ds.nbr_pd = reinterpret_cast<const Real*>(base + o_npd);
    ds.u_diag = reinterpret_cast<const Real*>(base + o_ud);
    ds.pdiag = reinterpret_cast<const Real*>(base + o_pd);
    ds.tdiag = reinterpret_cast<const Real*>(base + o_td);
    ds.pot = reinterpret_cast<const Real*>(base + o_pot);
    ds.epot = reinterpret_cast<const Real*>(base + o_epot);
    ds.U = reinterpret_cast<const Real*>(base + o_U);
    ds.V = reinterpret_cast<const Real*>(base + o_V);
    ds.tM = reinterpret_cast<const Real*>(base + o_tM);
    ds.pd = reinterpret_cast<const Real*>(base + o_pdd);
    ds.amp = amp.empty() ? nullptr : reinterpret_cast<const double*>(base + o_amp);
    ds.rows = rows.empty() ? nullptr : base + o_rows;
    ds.hop_uniform = 1;
    for (const Real t : nt) ds.hop_uniform &= t == nt[0] ? 1 : 0;
    ds.hop_value = nt.empty() ? Real(0) : nt[0];
    ds.fock = (h->method == FPSFTWA_METHOD_HF && h->coulomb_type == FPSFTWA_COULOMB_GENERAL) ? 1 : 0;
    ds.gpsr = h->method == FPSFTWA_METHOD_GPSR ? 1 : 0;
    if (ds.gpsr && h->coulomb_type == FPSFTWA_COULOMB_ONSITE) {
        // diffusion constant sqrt(i U0), U0 = Re U(0,0) (src/FHGPSRSolver.cpp:217, :630); diffusion only
        // for an on-site interaction (:248)
        const double u0 = h->U[0];
        const double m = std::sqrt(std::fabs(u0) / 2.0);
        ds.gc_re = m;
        ds.gc_im = u0 >= 0.0 ? m : -m;
    }
    return FPSFTWA_OK;
}

int prepare_initial(fpsftwa_handle* h)
{
    const int L = h->L;
    std::vector<double> occ(2 * (size_t)L);
    for (int i = 0; i < L; ++i) {
        occ[i] = h->n_up.empty() ? 0.0 : h->n_up[i];
        occ[L + i] = h->n_dn.empty() ? 0.0 : h->n_dn[i];
    }
    HIP_TRY(h->occ_d.ensure(occ.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(h->occ_d.p, occ.data(), occ.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (!h->noise.empty()) {
        HIP_TRY(h->noise_d.ensure(h->noise.size() * sizeof(double)));
        HIP_TRY(hipMemcpyAsync(h->noise_d.p, h->noise.data(), h->noise.size() * sizeof(double), hipMemcpyHostToDevice,
                               h->stream));
    }
    if (!h->rot_up.empty()) {
        const size_t LL2 = (size_t)L * L * 2;
        std::vector<double> rot(2 * LL2);
        std::memcpy(rot.data(), h->rot_up.data(), LL2 * sizeof(double));
        std::memcpy(rot.data() + LL2, h->rot_dn.data(), LL2 * sizeof(double));
        HIP_TRY(h->rot_d.ensure(rot.size() * sizeof(double)));
        HIP_TRY(hipMemcpyAsync(h->rot_d.p, rot.data(), rot.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FPSFTWA_OK;
}

// The state buffer, and the sampled (and rotated) initial states in it unless the integrator samples
// them in its first launch (in_kernel).
template <typename Real>
int sample_states(fpsftwa_handle* h, const DevSystem<Real>& ds, bool in_kernel = false)
{
    const int L = h->L;
    const size_t NE = 2 * (size_t)L * L;
    HIP_TRY(h->state.ensure((size_t)h->B * NE * 2 * sizeof(Real)));
    if (in_kernel) return FPSFTWA_OK;
    const double* occ = h->occ_d.as<double>();
    const double* noise = h->noise.empty() ? nullptr : h->noise_d.as<double>();
    if (h->method == FPSFTWA_METHOD_HF || h->method == FPSFTWA_METHOD_GPSR) {  // the state itself: zero noise
        const size_t nz = (size_t)h->B * 2 * ((size_t)L * (L - 1) / 2) * 2;
        HIP_TRY(h->noise_d.ensure(std::max<size_t>(nz, 1) * sizeof(double)));
        HIP_TRY(hipMemsetAsync(h->noise_d.p, 0, std::max<size_t>(nz, 1) * sizeof(double), h->stream));
        noise = h->noise_d.as<double>();
    }
    HIP_TRY(launch_sample<Real>(ds, occ, occ + L, noise, h->seed, 0, h->B, h->state.as<Real>(), h->stream));
    if (!h->rot_up.empty()) {
        HIP_TRY(h->scratch.ensure((size_t)h->B * NE * 2 * sizeof(Real)));
        HIP_TRY(launch_rotate<Real>(L, h->rot_d.as<double>(), h->B, h->state.as<Real>(), h->scratch.as<Real>(),
                                    h->stream));
    }
    return FPSFTWA_OK;
}

struct EventTimer {
    std::vector<hipEvent_t> ev;
    ~EventTimer()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    hipError_t mark(hipStream_t s)
    {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        ev.push_back(e);
        return hipEventRecord(e, s);
    }
    // sum of (ev[2k+1] - ev[2k])
    double pairs_ms() const
    {
        double t = 0.0;
        for (size_t k = 0; k + 1 < ev.size(); k += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) t += ms;
        }
        return t;
    }
};

// Basis occupations: device copies of the bases, the output and the chunk buffers for nt_max time
// points per reduction.
int prepare_bases(fpsftwa_handle* h, int nt_max)
{
    if (h->n_bases == 0) return FPSFTWA_OK;
    const int L = h->L;
    if (!h->dense) {
        h->dense = dense_create(h->stream);
        if (!h->dense) return fail(FPSFTWA_EDEVICE, "rocBLAS handle creation failed");
    }
    HIP_TRY(h->bases_d.ensure(h->bases.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(h->bases_d.p, h->bases.data(), h->bases.size() * sizeof(double), hipMemcpyHostToDevice,
                           h->stream));
    const size_t nout = (size_t)h->n_bases * 2 * L * h->T;
    HIP_TRY(h->basis_out.ensure(nout * sizeof(double)));
    const size_t P2 = (size_t)L * (L + 1);
    HIP_TRY(h->dens_part.ensure((size_t)nt_max * reduce_chunks(h->B) * P2 * 2 * sizeof(double)));
    HIP_TRY(h->dens_S.ensure((size_t)nt_max * P2 * sizeof(double)));
    HIP_TRY(h->basis_work.ensure(basis_work_doubles(L, nt_max) * sizeof(double)));
    return FPSFTWA_OK;
}

// Density sums of records [t0, t0 + nt) and their projection on every basis.
template <typename Real>
int reduce_bases(fpsftwa_handle* h, int R, int nt, int t0)
{
    if (h->n_bases == 0) return FPSFTWA_OK;
    HIP_TRY(launch_density_reduce<Real>(h->L, h->B, R, nt, h->rec.as<Real>(), h->dens_part.as<double>(),
                                        reduce_chunks(h->B), h->dens_S.as<double>(), h->stream));
    HIP_TRY(basis_project(h->dense, h->L, nt, t0, h->T, h->n_bases, h->bases_d.as<double>(), h->dens_S.as<double>(),
                          h->basis_work.as<double>(), h->basis_out.as<double>()));
    return FPSFTWA_OK;
}

// Kernel family chosen for a batched run.
struct KernelChoice {
    bool herm = false, pipe = false, seq = false, wide = false, dense = false;
    int nt = 0;  // threads per workgroup of the full-matrix stencil kernel
    // the spin-sequential kernel (FP32 seq; FP64 herm runs it too), which can sample the initial state
    bool spin_sequential(bool fp64) const { return seq || wide || (fp64 && herm); }
};

// Whether the run's first integrator launch samples the initial states itself: fTWA with the
// generator's noise and diagonal initial states on the spin-sequential kernel.
template <typename Real>
bool samples_in_kernel(const fpsftwa_handle* h, const KernelChoice& k)
{
    return k.spin_sequential(std::is_same<Real, double>::value) && h->method == FPSFTWA_METHOD_FTWA &&
           h->noise.empty() && h->rot_up.empty();
}

template <typename Real>
int choose_kernel(fpsftwa_handle* h, const DevSystem<Real>& ds, KernelChoice& k)
{
    k = KernelChoice{};
    k.nt = h->nt_hint;
    k.dense = h->kernel == FPSFTWA_KERNEL_DENSE || h->kernel == FPSFTWA_KERNEL_DENSE_MFMA;
    if (k.dense) {
        // forced
    } else if (h->kernel == FPSFTWA_KERNEL_SEQUENTIAL) {
        k.seq = herm_seq_fits<Real>(ds) && h->n_bases == 0;
        if constexpr (std::is_same<Real, float>::value) k.wide = !k.seq && herm_wide_fits(ds) && h->n_bases == 0;
        if (!k.seq && !k.wide)
            return fail(FPSFTWA_EUNSUPPORTED, "the spin-sequential kernel needs L <= 64 (FP32: L <= 107), 4 or 8 "
                                              "neighbours, no pulse, null or on-site Coulomb and no basis occupations");
    } else if constexpr (std::is_same<Real, float>::value) {
        // automatic choice: the spin-sequential kernel where it applies (two workgroups per CU:
        // 3.12e7 trajectory-steps/s on the 8x8 workload, against 2.63e7 for the both-spin kernel), its
        // direct-commutator form for 64 < L <= 107 and where that form measured faster at L <= 64 (the
        // 8x8 workload), else the both-spin upper-triangle kernel (pulses, V interactions, basis
        // occupations)
        k.seq = h->kernel == FPSFTWA_KERNEL_AUTO && herm_seq_fits<float>(ds) && h->n_bases == 0 &&
                !herm_wide_direct_preferred(ds);
        k.wide = !k.seq && h->kernel == FPSFTWA_KERNEL_AUTO && herm_wide_fits(ds) && h->n_bases == 0;
        k.herm = !k.seq && !k.wide && h->kernel != FPSFTWA_KERNEL_FULL && herm_fits(h->L, ds.z);
        if (h->kernel == FPSFTWA_KERNEL_HERMITIAN && !k.herm)
            return fail(FPSFTWA_EUNSUPPORTED, "the upper-triangle kernel does not fit L=" + std::to_string(h->L));
        // the pipelined variant is selected only on request: on the 8x8 workload it measures 9 % slower
        // than herm_kernel (2.27e7 vs 2.48e7 trajectory-steps/s), its extra barrier and register
        // pressure outweighing the overlap of the two spins' phases
        k.pipe = k.herm && h->kernel == FPSFTWA_KERNEL_PIPELINED && herm_pipe_fits(ds);
        if (h->kernel == FPSFTWA_KERNEL_PIPELINED && !k.pipe)
            return fail(FPSFTWA_EUNSUPPORTED, "the pipelined kernel needs L <= 64, no pulse and null or on-site "
                                              "Coulomb interactions");
    } else if (h->kernel == FPSFTWA_KERNEL_PIPELINED) {
        return fail(FPSFTWA_EUNSUPPORTED, "the spin-pipelined kernel is FP32 only");
    } else {
        // FP64: the upper-triangle kernel where it applies (no pulse, null or on-site Coulomb, L <= 64)
        k.herm = h->kernel != FPSFTWA_KERNEL_FULL && herm_seq_fits<double>(ds) && h->n_bases == 0;
        if (h->kernel == FPSFTWA_KERNEL_HERMITIAN && !k.herm)
            return fail(FPSFTWA_EUNSUPPORTED, "the FP64 upper-triangle kernel needs L <= 64, 4 or 8 neighbours, "
                                              "no pulse, null or on-site Coulomb and no basis occupations");
    }
    if (!k.dense && !k.herm && !k.seq && !k.wide && !stencil_fits<Real>(h->L, ds.z, &k.nt)) {
        if (h->kernel == FPSFTWA_KERNEL_FULL)
            return fail(FPSFTWA_EUNSUPPORTED, "the full-matrix stencil kernel does not support L=" +
                                                  std::to_string(h->L) + " with " + std::to_string(ds.z) +
                                                  " neighbours per site");
        k.dense = true;  // large or long-range lattices: one GEMM per stage for the batch
    }
    if (k.dense && h->L > 1024) return fail(FPSFTWA_EUNSUPPORTED, "the batched engines support L <= 1024");
    return FPSFTWA_OK;
}

RunShape run_shape(const fpsftwa_handle* h, int R, int dens)
{
    RunShape rs;
    rs.B = h->B;
    rs.T = h->T;
    rs.nsub = h->nsub;
    rs.h = h->dt / h->nsub;
    rs.R = R;
    rs.dens = dens;
    rs.ci = h->rec_i;
    rs.cj = h->rec_j < 0 ? h->L - 1 : h->rec_j;
    rs.seed = h->seed;
    return rs;
}

// The time grid in launches of Tc time points: integrator, then the records' ensemble reduction into
// acc [T][Q][2] (and the basis occupations).  kt (may be null) brackets the integrator launches.
template <typename Real>
int integrate_chunks(fpsftwa_handle* h, const DevSystem<Real>& ds, const KernelChoice& k, const RunShape& rs, int Tc,
                     Real* state, Real* rec, double* part, double* acc, EventTimer* kt)
{
    const int L = h->L;
    const int nchunk = reduce_chunks(h->B);
    for (int t0 = 0; t0 < h->T; t0 += Tc) {
        const int t1 = std::min(h->T, t0 + Tc);
        if (kt) HIP_TRY(kt->mark(h->stream));
        if (k.dense) {
            HIP_TRY(dense_run<Real>(h->dense, ds, rs, h->dense_buf.as<Real>(), state, rec, t0, t1, t0 == 0,
                                    t1 == h->T));
        } else if (k.seq) {
            HIP_TRY(launch_herm_seq<Real>(ds, rs, t0, t1, state, rec, h->nt_hint, h->stream));
        } else if constexpr (std::is_same<Real, float>::value) {
            if (k.wide) HIP_TRY(launch_herm_wide(ds, rs, t0, t1, state, rec, h->stream));
            else if (k.pipe) HIP_TRY(launch_herm_pipe(ds, rs, t0, t1, state, rec, h->stream));
            else if (k.herm) HIP_TRY(launch_herm(ds, rs, t0, t1, state, rec, h->nt_hint, h->stream));
            else HIP_TRY(launch_stencil<Real>(ds, rs, t0, t1, state, rec, k.nt, h->stream));
        } else if (k.herm) {
            HIP_TRY(launch_herm_seq<double>(ds, rs, t0, t1, state, rec, h->nt_hint, h->stream));
        } else {
            HIP_TRY(launch_stencil<Real>(ds, rs, t0, t1, state, rec, k.nt, h->stream));
        }
        if (kt) HIP_TRY(kt->mark(h->stream));
        ++h->launches;
        HIP_TRY(launch_reduce<Real>(L, h->B, rs.R, t1 - t0, t0, rec, part, nchunk, acc, h->stream));
        if (int rc = reduce_bases<Real>(h, rs.R, t1 - t0, t0)) return rc;
    }
    return FPSFTWA_OK;
}

template <typename Real>
int run_batched(fpsftwa_handle* h, EventTimer& kt)
{
    DevSystem<Real> ds;
    int rc = upload_system<Real>(h, ds, h->T);
    if (rc) return rc;
    KernelChoice k;
    if ((rc = choose_kernel<Real>(h, ds, k))) return rc;
    h->used_herm = k.herm;
    h->used_dense = k.dense;
    if (k.dense && !h->dense) {
        h->dense = dense_create(h->stream);
        if (!h->dense) return fail(FPSFTWA_EDEVICE, "rocBLAS handle creation failed");
    }
    // FP32 long-range systems: the product fused with the update on the matrix cores by default (1.72e5
    // trajectory-steps/s on the 196-site bilayer, against 1.07e5 for the library GEMM and update kernel)
    if (h->dense)
        dense_set_fused_mfma(h->dense, h->kernel == FPSFTWA_KERNEL_DENSE_MFMA ||
                                           (h->kernel == FPSFTWA_KERNEL_AUTO &&
                                            h->engine == FPSFTWA_ENGINE_FP32_BATCHED));
    const bool fused = samples_in_kernel<Real>(h, k);
    if ((rc = sample_states<Real>(h, ds, fused))) return rc;
    const int L = h->L;
    const int dens = h->n_bases > 0 ? 1 : 0;
    const int R = 2 * L + 8 + dens * L * (L + 1);
    const int Q = 6 * L + 12;
    int Tc = h->time_chunk;
    if (Tc <= 0) {
        // records of up to 64 time points per launch; with the density part (basis occupations) the
        // records are L/2 times wider and get a larger budget, so the trajectory state is still
        // carried across several time points per launch
        const double per_tp = (double)h->B * R * sizeof(Real);
        Tc = (int)std::max(1.0, std::min(64.0, (dens ? 8e9 : 1.5e9) / per_tp));
    }
    // buffers for a whole chunk even when this run is shorter, so a longer run on the handle reuses them
    const int Tc_alloc = Tc;
    Tc = std::min(Tc, h->T);
    const int nchunk = reduce_chunks(h->B);
    HIP_TRY(h->rec.ensure((size_t)Tc_alloc * h->B * R * sizeof(Real)));
    HIP_TRY(h->part.ensure((size_t)Tc_alloc * nchunk * Q * 2 * sizeof(double)));
    if ((rc = prepare_bases(h, Tc))) return rc;
    RunShape rs = run_shape(h, R, dens);
    if (fused) rs.sample_occ = h->occ_d.as<double>();
    if (k.dense) HIP_TRY(h->dense_buf.ensure(dense_buffer_elems<Real>(L, h->B) * sizeof(Real)));
    h->graph_used = h->graph && !k.dense && h->n_bases == 0;  // the GEMM library stays out of the graph
    if (!h->graph_used)
        return integrate_chunks<Real>(h, ds, k, rs, Tc, h->state.as<Real>(), h->rec.as<Real>(), h->part.as<double>(),
                                      h->acc_d.as<double>(), &kt);
    // every integrator and reduction launch of the run captured on the handle's stream (thread-local
    // capture: other handles capture their own streams from their own threads), then launched at once
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = integrate_chunks<Real>(h, ds, k, rs, Tc, h->state.as<Real>(), h->rec.as<Real>(), h->part.as<double>(),
                                h->acc_d.as<double>(), nullptr);
    const hipError_t ec = hipStreamEndCapture(h->stream, &h->graph_g);
    if (rc) return rc;
    HIP_TRY(ec);
    HIP_TRY(hipGraphInstantiate(&h->graph_x, h->graph_g, nullptr, nullptr, 0));
    HIP_TRY(kt.mark(h->stream));
    HIP_TRY(hipGraphLaunch(h->graph_x, h->stream));
    HIP_TRY(kt.mark(h->stream));
    return FPSFTWA_OK;
}

// The last run's graph, once its stream has finished with it.
void release_graph(fpsftwa_handle* h)
{
    if (h->graph_x) (void)hipGraphExecDestroy(h->graph_x);
    if (h->graph_g) (void)hipGraphDestroy(h->graph_g);
    h->graph_x = nullptr;
    h->graph_g = nullptr;
}

// fpsftwa_run_noalloc: the batched run with the caller's state and sum buffers and the fixed scratch.
template <typename Real>
int run_noalloc(fpsftwa_handle* h, void* state, size_t state_bytes, double* sums, size_t n_sums, char* scr)
{
    const int L = h->L;
    const int R = 2 * L + 8, Q = 6 * L + 12;
    const size_t cap = FPSFTWA_NOALLOC_SCRATCH_BYTES;
    auto r256 = [](size_t b) { return (b + 255) / 256 * 256; };
    // one time point of records and partial sums must fit beside the occupations and the system tables
    const int nchunk = reduce_chunks(h->B);
    const size_t rec_tp = (size_t)h->B * R * sizeof(Real), part_tp = (size_t)nchunk * Q * 2 * sizeof(double);
    const size_t occ_bytes = r256(2 * (size_t)L * sizeof(double));
    if (occ_bytes + r256(rec_tp) + r256(part_tp) > cap)
        return fail(FPSFTWA_EINVAL, "n_traj = " + std::to_string(h->B) + " needs " + std::to_string(rec_tp + part_tp) +
                                        " B of records per time point; the fixed scratch holds " + std::to_string(cap) +
                                        " B (fpsftwa_run has no such limit)");
    if (n_sums < (size_t)h->T * Q * 2)
        return fail(FPSFTWA_EINVAL, "sums holds " + std::to_string(n_sums) + " doubles; the run needs T*(6L+12)*2 = " +
                                        std::to_string((size_t)h->T * Q * 2));
    const size_t need_state = (size_t)h->B * 2 * L * L * 2 * sizeof(Real);
    if (state_bytes < need_state)
        return fail(FPSFTWA_EINVAL, "state holds " + std::to_string(state_bytes) + " B; n_traj x 2 x L x L complex " +
                                        (sizeof(Real) == 4 ? "FP32" : "FP64") + " needs " + std::to_string(need_state));
    std::vector<double> occ(2 * (size_t)L);
    for (int i = 0; i < L; ++i) {
        occ[i] = h->n_up[i];
        occ[L + i] = h->n_dn[i];
    }
    HIP_TRY(hipMemcpyAsync(scr, occ.data(), occ.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));  // occ is a local
    size_t off = occ_bytes, used = 0;
    DevSystem<Real> ds;
    int rc = upload_system<Real>(h, ds, h->T, scr + off, cap - off, &used);
    if (rc) return rc;
    off += r256(used);
    KernelChoice k;
    if ((rc = choose_kernel<Real>(h, ds, k))) return rc;
    if (k.dense)
        return fail(FPSFTWA_EUNSUPPORTED, "the dense path (L > 64 or long-range hopping) keeps its stage input and "
                                          "RK registers (several n_traj x L x L buffers) beside the state, and its "
                                          "library GEMM form a workspace: use fpsftwa_run");
    h->used_herm = k.herm;
    h->used_dense = false;
    // time points per launch: as many as the rest of the scratch holds (at most 64, or the tuned chunk)
    int Tc = h->time_chunk > 0 ? h->time_chunk : 64;
    Tc = std::min(Tc, h->T);
    while (Tc > 1 && off + r256((size_t)Tc * rec_tp) + r256((size_t)Tc * part_tp) > cap) --Tc;
    if (off + r256((size_t)Tc * rec_tp) + r256((size_t)Tc * part_tp) > cap)
        return fail(FPSFTWA_EINVAL, "the records of one time point do not fit the fixed scratch beside the system "
                                    "tables");
    Real* rec = reinterpret_cast<Real*>(scr + off);
    double* part = reinterpret_cast<double*>(scr + off + r256((size_t)Tc * rec_tp));
    const double* occ_d = reinterpret_cast<const double*>(scr);
    RunShape rs = run_shape(h, R, 0);
    if (samples_in_kernel<Real>(h, k)) rs.sample_occ = occ_d;
    else HIP_TRY(launch_sample<Real>(ds, occ_d, occ_d + L, nullptr, h->seed, 0, h->B, static_cast<Real*>(state),
                                     h->stream));
    if ((rc = integrate_chunks<Real>(h, ds, k, rs, Tc, static_cast<Real*>(state), rec, part, sums, nullptr))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FPSFTWA_OK;
}

// GPSR on the batched engines (ftwa_gpsr.hip): records R = 2L + 8 [+ L(L+1)] + 2L (imaginary diagonals last).
template <typename Real>
int run_gpsr_batched(fpsftwa_handle* h, EventTimer& kt)
{
    DevSystem<Real> ds;
    int rc = upload_system<Real>(h, ds, h->T);
    if (rc) return rc;
    if (!gpsr_fits<Real>(ds))
        return fail(FPSFTWA_EUNSUPPORTED, "GPSR on the batched engines needs L <= 32, at most 8 neighbours per site "
                                          "and no pulse; the plain FP64 engine runs any system");
    if ((rc = sample_states<Real>(h, ds))) return rc;
    h->used_herm = false;
    h->used_dense = false;
    const int L = h->L;
    const int dens = h->n_bases > 0 ? 1 : 0;
    const int R = 2 * L + 8 + dens * L * (L + 1) + 2 * L;
    const int Q = 6 * L + 12;
    int Tc = h->time_chunk;
    if (Tc <= 0) {  // records and the chunk's diffusion noise within a few GB
        const double per_tp = (double)h->B * R * sizeof(Real) + (double)h->nsub * h->B * 4 * L * sizeof(Real);
        Tc = (int)std::max(1.0, std::min(64.0, (dens ? 8e9 : 3e9) / per_tp));
    }
    Tc = std::min(Tc, h->T);
    const int nchunk = reduce_chunks(h->B);
    HIP_TRY(h->rec.ensure((size_t)Tc * h->B * R * sizeof(Real)));
    HIP_TRY(h->part.ensure((size_t)Tc * nchunk * Q * 2 * sizeof(double)));
    HIP_TRY(h->gnoise.ensure(std::max<size_t>(gpsr_noise_elems<Real>(L, h->B, h->nsub, Tc), 1) * sizeof(Real)));
    if ((rc = prepare_bases(h, Tc))) return rc;
    RunShape rs;
    rs.B = h->B;
    rs.T = h->T;
    rs.nsub = h->nsub;
    rs.h = h->dt / h->nsub;
    rs.R = R;
    rs.dens = dens;
    rs.ci = h->rec_i;
    rs.cj = h->rec_j < 0 ? L - 1 : h->rec_j;
    rs.seed = h->seed;
    for (int t0 = 0; t0 < h->T; t0 += Tc) {
        const int t1 = std::min(h->T, t0 + Tc);
        HIP_TRY(kt.mark(h->stream));
        HIP_TRY(launch_gpsr<Real>(ds, rs, t0, t1, h->state.as<Real>(), h->rec.as<Real>(), h->gnoise.as<Real>(),
                                  h->stream));
        HIP_TRY(kt.mark(h->stream));
        ++h->launches;
        HIP_TRY(launch_reduce<Real>(L, h->B, R, t1 - t0, t0, h->rec.as<Real>(), h->part.as<double>(), nchunk,
                                    h->acc_d.as<double>(), h->stream, R - 2 * L));
        if ((rc = reduce_bases<Real>(h, R, t1 - t0, t0))) return rc;
    }
    return FPSFTWA_OK;
}

int run_plain(fpsftwa_handle* h, EventTimer& kt)
{
    DevSystem<double> ds;
    int rc = upload_system<double>(h, ds, h->T);
    if (rc) return rc;
    if ((rc = sample_states<double>(h, ds))) return rc;
    const int L = h->L;
    const int dens = h->n_bases > 0 ? 1 : 0;
    const int gp = h->method == FPSFTWA_METHOD_GPSR ? 1 : 0;
    const int R = 2 * L + 8 + dens * L * (L + 1) + gp * 2 * L;
    const int Q = 6 * L + 12;
    const int nchunk = reduce_chunks(h->B);
    HIP_TRY(h->rec.ensure((size_t)h->B * R * sizeof(double)));
    HIP_TRY(h->part.ensure((size_t)nchunk * Q * 2 * sizeof(double)));
    HIP_TRY(h->work.ensure((size_t)h->B * plain_work_doubles(L) * sizeof(double)));
    if ((rc = prepare_bases(h, 1))) return rc;
    RunShape rs;
    rs.B = h->B;
    rs.T = h->T;
    rs.nsub = h->nsub;
    rs.h = h->dt / h->nsub;
    rs.R = R;
    rs.dens = dens;
    rs.ci = h->rec_i;
    rs.cj = h->rec_j < 0 ? L - 1 : h->rec_j;
    rs.seed = h->seed;
    for (int t = 0; t < h->T; ++t) {
        HIP_TRY(kt.mark(h->stream));
        HIP_TRY(launch_plain(ds, rs, t, h->state.as<double>(), h->work.as<double>(), h->rec.as<double>(), h->stream));
        HIP_TRY(kt.mark(h->stream));
        ++h->launches;
        HIP_TRY(launch_reduce<double>(L, h->B, R, 1, t, h->rec.as<double>(), h->part.as<double>(), nchunk,
                                      h->acc_d.as<double>(), h->stream, gp ? R - 2 * L : -1));
        if ((rc = reduce_bases<double>(h, R, 1, t))) return rc;
    }
    return FPSFTWA_OK;
}

double kahan_value(const std::vector<double>& acc, int Q, int t, int q)
{
    const double* p = &acc[((size_t)t * Q + q) * 2];
    return p[0] + p[1];
}

// dJ/d(pulse parameters) from dJ/da at every RK stage time: knots by their linear-interpolation
// weights (the same rule as pulse_amplitude), the Gaussian amplitude A0 by the envelope a/A0.
void pulse_gradient(const fpsftwa_handle* h, const std::vector<double>& dJda, double* out, size_t n_out)
{
    const double hs = h->dt / h->nsub;
    const double cs[4] = {0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0};
    for (size_t k = 0; k < n_out; ++k) out[k] = 0.0;
    const int n = (int)h->knots.size();
    for (size_t q = 0; q < dJda.size(); ++q) {
        const double t = (double)(q / 4) * hs + cs[q % 4] * hs;
        if (h->pulse_type == FPSFTWA_PULSE_KNOTS) {
            if (n == 1) {
                out[0] += dJda[q];
                continue;
            }
            const double step = h->t_max / (double)(n - 1);
            const double u = t / step;
            if (!(step > 0.0) || !std::isfinite(u) || u <= 0.0) {  // same rule as pulse_amplitude
                out[0] += dJda[q];
                continue;
            }
            const int k = (int)std::floor(u);
            if (k >= n - 1) {
                out[n - 1] += dJda[q];
                continue;
            }
            const double f = u - (double)k;
            out[k] += dJda[q] * (1.0 - f);
            out[k + 1] += dJda[q] * f;
        } else if (h->pulse_type == FPSFTWA_PULSE_GAUSSIAN) {
            const double x = (t - h->t0) / h->td;
            out[0] += dJda[q] * std::exp(-(x * x) / 2.0) * std::cos(h->w0 * (t - h->t0) + h->phi);
        }
    }
}

template <typename Real>
int run_adjoint(fpsftwa_handle* h, int ck_every, double* objective, double* grad_u, double* grad_t,
                double* grad_pulse, size_t n_pulse)
{
    DevSystem<Real> ds;
    int rc = upload_system<Real>(h, ds, h->T);
    if (rc) return rc;
    if ((rc = sample_states<Real>(h, ds))) return rc;
    const int L = h->L;
    const int64_t B = h->B;
    const size_t NE = 2 * (size_t)L * L;
    const int64_t nst = (int64_t)(h->T - 1) * h->nsub;  // steps up to the last time point
    const int64_t K = std::max<int64_t>(1, std::min<int64_t>(ck_every, std::max<int64_t>(nst, 1)));
    const int64_t nseg = (nst + K - 1) / K;
    const int64_t n_stage = ds.amp ? nst * 4 : 0;
    const int64_t G = 2 * (int64_t)L + (int64_t)ds.z * L + n_stage;
    RunShape rs;
    rs.B = B;
    rs.T = h->T;
    rs.nsub = h->nsub;
    rs.h = h->dt / h->nsub;
    DevBuf ck, wk, gacc, jb, out, tt;
    HIP_TRY(gacc.ensure((size_t)B * G * sizeof(double)));
    HIP_TRY(jb.ensure((size_t)B * sizeof(double)));
    HIP_TRY(out.ensure((size_t)(G + 1) * sizeof(double)));
    HIP_TRY(hipMemsetAsync(gacc.p, 0, (size_t)B * G * sizeof(double), h->stream));
    Real* ttab = nullptr;
    if (n_stage) {
        HIP_TRY(tt.ensure((size_t)n_stage * ds.z * L * 2 * sizeof(Real)));
        ttab = tt.as<Real>();
        HIP_TRY(launch_adjoint_table<Real>(ds, n_stage, ttab, h->stream));
    }
    Real* state = h->state.as<Real>();
    EventTimer kt;
    const int want_nt = h->nt_hint == 512 || h->nt_hint == 1024 ? h->nt_hint : 0;
    int lds_slots = 0;
    const int lds_nt = adjoint_lds_threads<Real>(L, ds.z, &lds_slots, want_nt);
    if (lds_nt) {
        // LDS-resident form: one forward pass writes the checkpoints, then each segment is recomputed and
        // reversed in one launch; states and cotangents in the kernel's slot order between launches
        const size_t SL = (size_t)lds_nt * lds_slots * 2;  // reals per trajectory state
        DevBuf seg, lam;
        HIP_TRY(ck.ensure(std::max<int64_t>(nseg, 1) * B * SL * sizeof(Real)));
        HIP_TRY(seg.ensure((size_t)B * K * SL * sizeof(Real)));
        HIP_TRY(lam.ensure((size_t)B * SL * sizeof(Real)));
        HIP_TRY(kt.mark(h->stream));
        HIP_TRY(launch_adjoint_lds<Real>(ds, ttab, rs, want_nt, 0, 0, nst, nst, K, n_stage, state, ck.as<Real>(),
                                         seg.as<Real>(), lam.as<Real>(), gacc.as<double>(), jb.as<double>(), h->stream));
        for (int64_t s = nseg - 1; s >= 0; --s)
            HIP_TRY(launch_adjoint_lds<Real>(ds, ttab, rs, want_nt, 1, s * K, std::min(nst, (s + 1) * K), nst, K,
                                             n_stage, state, ck.as<Real>(), seg.as<Real>(), lam.as<Real>(),
                                             gacc.as<double>(), jb.as<double>(), h->stream));
    } else {
        HIP_TRY(wk.ensure(adjoint_work_bytes<Real>(L, B, K)));
        HIP_TRY(ck.ensure(std::max<int64_t>(nseg, 1) * B * NE * 2 * sizeof(Real)));
        Real* work = wk.as<Real>();
        Real* segbuf = work + (size_t)B * NE * 2 * 10;
        const size_t ck_bytes = (size_t)B * NE * 2 * sizeof(Real);
        HIP_TRY(kt.mark(h->stream));
        for (int64_t s = 0; s < nseg; ++s) {
            HIP_TRY(hipMemcpyAsync(ck.as<char>() + s * ck_bytes, state, ck_bytes, hipMemcpyDeviceToDevice, h->stream));
            HIP_TRY(launch_adjoint_forward<Real>(ds, ttab, rs, s * K, std::min(nst, (s + 1) * K), state, work,
                                                 h->stream));
        }
        HIP_TRY(launch_adjoint_terminal<Real>(ds, B, state, work, jb.as<double>(), h->stream));
        for (int64_t s = nseg - 1; s >= 0; --s)
            HIP_TRY(launch_adjoint_backward<Real>(ds, ttab, rs, s * K, std::min(nst, (s + 1) * K),
                                                  reinterpret_cast<const Real*>(ck.as<char>() + s * ck_bytes), segbuf,
                                                  work, gacc.as<double>(), n_stage, h->stream));
    }
    HIP_TRY(launch_adjoint_reduce(gacc.as<double>(), B, G, out.as<double>(), h->stream));
    HIP_TRY(launch_adjoint_reduce(jb.as<double>(), B, 1, out.as<double>() + G, h->stream));
    HIP_TRY(kt.mark(h->stream));
    std::vector<double> g(G + 1);
    HIP_TRY(hipMemcpyAsync(g.data(), out.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->kernel_ms = kt.pairs_ms();
    h->launches = lds_nt ? nseg + 1 : 2 * nseg + 1;
    const double inv = 1.0 / (double)B;
    if (objective) *objective = g[G] * inv;
    if (grad_u)
        for (int i = 0; i < L; ++i) grad_u[i] = g[i] * inv;
    if (grad_t) {
        for (size_t e = 0; e < (size_t)L * L; ++e) grad_t[e] = 0.0;
        for (int k = 0; k < L; ++k) grad_t[(size_t)k * L + k] = g[L + k] * inv;
        // the entries of upload_system's neighbour tables
        int zz = 0;
        const std::vector<std::vector<int>> ent = neighbour_entries(h, zz, false);
        for (int k = 0; k < L; ++k)
            for (int n = 0; n < ds.z && n < zz; ++n)
                if (ent[k][n] >= 0) grad_t[(size_t)k * L + ent[k][n]] = g[2 * L + (size_t)n * L + k] * inv;
    }
    if (grad_pulse && n_pulse && n_stage) {
        std::vector<double> dJda(g.begin() + 2 * L + (size_t)ds.z * L, g.begin() + G);
        for (double& v : dJda) v *= inv;
        pulse_gradient(h, dJda, grad_pulse, n_pulse);
    } else if (grad_pulse) {
        for (size_t k = 0; k < n_pulse; ++k) grad_pulse[k] = 0.0;
    }
    return FPSFTWA_OK;
}

}  // namespace

extern "C" {

const char* fpsftwa_last_error(void) { return g_last_error.c_str(); }

int fpsftwa_version(int* major, int* minor, int* patch)
{
    if (major) *major = FPSFTWA_VERSION_MAJOR;
    if (minor) *minor = FPSFTWA_VERSION_MINOR;
    if (patch) *patch = FPSFTWA_VERSION_PATCH;
    return FPSFTWA_OK;
}

int fpsftwa_device_info(char* name, size_t name_len, int* cu_count, double* hbm_gib)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(FPSFTWA_EDEVICE, "no HIP device visible");
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_gib) *hbm_gib = prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    return FPSFTWA_OK;
}

int fpsftwa_create(const fpsftwa_system* sys, fpsftwa_handle** out)
{
    if (!sys || !out) return fail(FPSFTWA_EINVAL, "null argument");
    *out = nullptr;
    const int L = sys->n_sites;
    if (L < 1) return fail(FPSFTWA_EINVAL, "n_sites must be >= 1");
    if (!sys->hopping) return fail(FPSFTWA_EINVAL, "hopping matrix is required");
    if (sys->coulomb_type < 0 || sys->coulomb_type > 2) return fail(FPSFTWA_EINVAL, "bad coulomb_type");
    if (sys->coulomb_type != FPSFTWA_COULOMB_NULL && !sys->coulomb_u) return fail(FPSFTWA_EINVAL, "coulomb_u is required");
    if (sys->coulomb_type == FPSFTWA_COULOMB_GENERAL && !sys->coulomb_v)
        return fail(FPSFTWA_EINVAL, "coulomb_v is required for COULOMB_GENERAL");
    if (sys->pulse_type < 0 || sys->pulse_type > 2) return fail(FPSFTWA_EINVAL, "bad pulse_type");
    if (sys->pulse_type != FPSFTWA_PULSE_NONE && !sys->displacement)
        return fail(FPSFTWA_EINVAL, "the pulse needs the displacement array");
    if (sys->pulse_type == FPSFTWA_PULSE_KNOTS && (sys->pulse_n_knots < 1 || !sys->pulse_knots))
        return fail(FPSFTWA_EINVAL, "PULSE_KNOTS needs at least one knot");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(FPSFTWA_EDEVICE, "no HIP device visible: this engine runs only on a gfx950 GPU");
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::string arch;
    if (!is_gfx950(dev, &arch))
        return fail(FPSFTWA_EDEVICE, "device " + std::to_string(dev) + " is " + arch + ", not gfx950");
    void* scr = nullptr;  // the no-allocation entry point's scratch on this device (looked up once)
    if (int rc = scratch_address(dev, &scr)) return rc;
    auto* h = new (std::nothrow) fpsftwa_handle();
    if (!h) return fail(FPSFTWA_ENOMEM, "host allocation failed");
    h->device = dev;
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(FPSFTWA_EDEVICE, "hipStreamCreate failed");
    }
    h->stream = h->own_stream;
    const size_t LL = (size_t)L * L;
    h->L = L;
    h->coulomb_type = sys->coulomb_type;
    h->tM.assign(sys->hopping, sys->hopping + LL);
    if (sys->coulomb_u) h->U.assign(sys->coulomb_u, sys->coulomb_u + LL);
    else h->U.assign(LL, 0.0);
    if (sys->coulomb_v && sys->coulomb_type == FPSFTWA_COULOMB_GENERAL) h->V.assign(sys->coulomb_v, sys->coulomb_v + LL);
    else h->V.assign(LL, 0.0);
    if (sys->potential) h->pot.assign(sys->potential, sys->potential + L);
    h->pulse_type = sys->pulse_type;
    h->A0 = sys->pulse_A0;
    h->t0 = sys->pulse_t0;
    h->td = sys->pulse_td;
    h->w0 = sys->pulse_w0;
    h->phi = sys->pulse_phi;
    for (int k = 0; k < 3; ++k) h->pol[k] = sys->pulse_pol[k];
    if (sys->pulse_type == FPSFTWA_PULSE_KNOTS) h->knots.assign(sys->pulse_knots, sys->pulse_knots + sys->pulse_n_knots);
    if (sys->displacement) h->disp.assign(sys->displacement, sys->displacement + LL * 3);
    h->n_up.assign(L, 0.0);
    h->n_dn.assign(L, 0.0);
    *out = h;
    return FPSFTWA_OK;
}

int fpsftwa_destroy(fpsftwa_handle* h)
{
    if (!h) return FPSFTWA_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->sys_buf.release();
    h->state.release();
    h->rec.release();
    h->part.release();
    h->acc_d.release();
    h->work.release();
    h->noise_d.release();
    h->rot_d.release();
    h->dense_buf.release();
    dense_destroy(h->dense);
    h->dense = nullptr;
    h->scratch.release();
    h->occ_d.release();
    h->bases_d.release();
    h->dens_part.release();
    h->dens_S.release();
    h->basis_work.release();
    h->basis_out.release();
    h->gnoise.release();
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return FPSFTWA_OK;
}

int fpsftwa_set_params(fpsftwa_handle* h, double dt, double t_max, int32_t n_sub, int64_t n_traj, int32_t int_order,
                       uint64_t seed, int32_t engine)
{
    if (int rc = check_handle(h)) return rc;
    (void)int_order;  // the reference forces RK4 "3/8" (src/FHfTWASolver.cpp:90)
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(FPSFTWA_EINVAL, "dt must be > 0");
    if (!(t_max >= 0.0) || !std::isfinite(t_max)) return fail(FPSFTWA_EINVAL, "t_max must be >= 0");
    if (n_sub < 1) return fail(FPSFTWA_EINVAL, "n_sub must be >= 1");
    if (n_traj < 1) return fail(FPSFTWA_EINVAL, "n_traj must be >= 1");
    if (engine < 0 || engine > 2) return fail(FPSFTWA_EINVAL, "bad engine");
    const double nt = std::floor(t_max / dt + 1.0);
    if (nt > 2.0e9) return fail(FPSFTWA_EINVAL, "too many time points");
    h->dt = dt;
    h->t_max = t_max;
    h->nsub = n_sub;
    h->B = n_traj;
    h->seed = seed;
    h->engine = engine;
    h->has_result = false;
    return FPSFTWA_OK;
}

int fpsftwa_set_results(fpsftwa_handle* h, uint32_t flags)
{
    if (int rc = check_handle(h)) return rc;
    h->flags = flags;
    return FPSFTWA_OK;
}

int fpsftwa_set_initial_state(fpsftwa_handle* h, const double* n_up, const double* n_dn, const double* rot_up,
                              const double* rot_dn)
{
    if (int rc = check_handle(h)) return rc;
    if (!n_up || !n_dn) return fail(FPSFTWA_EINVAL, "occupations are required");
    const int L = h->L;
    for (int i = 0; i < L; ++i)
        if (!(n_up[i] >= 0.0 && n_up[i] <= 1.0 && n_dn[i] >= 0.0 && n_dn[i] <= 1.0))
            return fail(FPSFTWA_EINVAL, "occupations must lie in [0, 1]");
    if ((rot_up == nullptr) != (rot_dn == nullptr))
        return fail(FPSFTWA_EINVAL, "give both rotations or neither");
    h->n_up.assign(n_up, n_up + L);
    h->n_dn.assign(n_dn, n_dn + L);
    if (rot_up) {
        h->rot_up.assign(rot_up, rot_up + (size_t)L * L * 2);
        h->rot_dn.assign(rot_dn, rot_dn + (size_t)L * L * 2);
    } else {
        h->rot_up.clear();
        h->rot_dn.clear();
    }
    h->has_result = false;
    return FPSFTWA_OK;
}

int fpsftwa_set_bases(fpsftwa_handle* h, int32_t n_bases, const double* bases)
{
    if (int rc = check_handle(h)) return rc;
    if (n_bases < 0 || n_bases > 2) return fail(FPSFTWA_EINVAL, "n_bases must lie in [0, 2]");
    if (n_bases > 0 && !bases) return fail(FPSFTWA_EINVAL, "bases are required when n_bases > 0");
    if (n_bases > 0 && h->L > 1024) return fail(FPSFTWA_EUNSUPPORTED, "basis occupations support L <= 1024");
    const size_t n = (size_t)n_bases * 2 * h->L * h->L;
    for (size_t e = 0; e < n; ++e)
        if (!std::isfinite(bases[e])) return fail(FPSFTWA_EINVAL, "bases must be finite");
    h->n_bases = n_bases;
    h->bases.assign(bases, bases + n);
    h->has_result = false;
    return FPSFTWA_OK;
}

int fpsftwa_set_method(fpsftwa_handle* h, int32_t method)
{
    if (int rc = check_handle(h)) return rc;
    if (method != FPSFTWA_METHOD_FTWA && method != FPSFTWA_METHOD_HF && method != FPSFTWA_METHOD_GPSR)
        return fail(FPSFTWA_EINVAL, "unknown method");
    h->method = method;
    h->has_result = false;
    return FPSFTWA_OK;
}

int fpsftwa_set_graph(fpsftwa_handle* h, int32_t on)
{
    if (int rc = check_handle(h)) return rc;
    h->graph = on != 0;
    return FPSFTWA_OK;
}

int fpsftwa_set_stream(fpsftwa_handle* h, void* stream)
{
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));  // nothing of this handle is in flight on the old stream
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
    dense_set_stream(h->dense, h->stream);
    return FPSFTWA_OK;
}

int fpsftwa_set_recorded_element(fpsftwa_handle* h, int32_t i, int32_t j)
{
    if (int rc = check_handle(h)) return rc;
    if (i < 0 || j < 0 || i >= h->L || j >= h->L) return fail(FPSFTWA_EINVAL, "element index outside [0, L)");
    h->rec_i = i;
    h->rec_j = j;
    h->has_result = false;
    return FPSFTWA_OK;
}

int fpsftwa_set_noise(fpsftwa_handle* h, const double* noise, size_t n_values)
{
    if (int rc = check_handle(h)) return rc;
    if (!noise) {
        h->noise.clear();
        return FPSFTWA_OK;
    }
    const size_t need = (size_t)h->B * 2 * ((size_t)h->L * (h->L - 1) / 2) * 2;
    if (n_values != need)
        return fail(FPSFTWA_EINVAL, "noise must hold n_traj*2*L(L-1)/2*2 = " + std::to_string(need) + " values");
    h->noise.assign(noise, noise + n_values);
    return FPSFTWA_OK;
}

int fpsftwa_sample_noise(fpsftwa_handle* h, int64_t first, int64_t count, double* out, size_t n_values)
{
    if (int rc = check_handle(h)) return rc;
    if (first < 0 || count < 0 || !out) return fail(FPSFTWA_EINVAL, "bad range");
    const size_t need = (size_t)count * 2 * ((size_t)h->L * (h->L - 1) / 2) * 2;
    if (n_values < need) return fail(FPSFTWA_EINVAL, "output too small");
    if (need == 0) return FPSFTWA_OK;
    HIP_TRY(hipSetDevice(h->device));
    DevBuf buf;
    HIP_TRY(buf.ensure(need * sizeof(double)));
    HIP_TRY(launch_noise(h->L, h->seed, first, count, buf.as<double>(), h->stream));
    HIP_TRY(hipMemcpyAsync(out, buf.p, need * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FPSFTWA_OK;
}

int fpsftwa_run(fpsftwa_handle* h)
{
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    if (!h->noise.empty()) {
        const size_t need = (size_t)h->B * 2 * ((size_t)h->L * (h->L - 1) / 2) * 2;
        if (h->noise.size() != need) return fail(FPSFTWA_EINVAL, "noise size does not match n_traj");
    }
    if (h->method == FPSFTWA_METHOD_HF && h->coulomb_type == FPSFTWA_COULOMB_GENERAL &&
        h->engine != FPSFTWA_ENGINE_FP64_PLAIN)
        return fail(FPSFTWA_EUNSUPPORTED, "Hartree-Fock with a V interaction (the Fock term) runs on the plain FP64 "
                                          "engine only");
    h->has_result = false;
    h->T = (int)std::floor(h->t_max / h->dt + 1.0);  // src/FHSolver.cpp:100
    if (h->method == FPSFTWA_METHOD_GPSR && (double)(h->T - 1) * h->nsub >= 4294967296.0)
        return fail(FPSFTWA_EINVAL, "GPSR runs at most 2^32 steps (the diffusion noise counter)");
    const int Q = 6 * h->L + 12;
    h->launches = 0;
    h->kernel_ms = 0.0;
    // sized in steps of 1,024 time points, so a somewhat longer run on the handle reuses the buffer
    HIP_TRY(h->acc_d.ensure((size_t)((h->T + 1023) / 1024 * 1024) * Q * 2 * sizeof(double)));
    HIP_TRY(hipMemsetAsync(h->acc_d.p, 0, (size_t)h->T * Q * 2 * sizeof(double), h->stream));
    int rc = prepare_initial(h);
    if (rc) return rc;
    EventTimer whole, kt;
    HIP_TRY(whole.mark(h->stream));
    if (h->engine == FPSFTWA_ENGINE_FP64_PLAIN) {
        h->state_double = true;
        rc = run_plain(h, kt);
    } else if (h->engine == FPSFTWA_ENGINE_FP64_BATCHED) {
        h->state_double = true;
        rc = h->method == FPSFTWA_METHOD_GPSR ? run_gpsr_batched<double>(h, kt) : run_batched<double>(h, kt);
    } else {
        h->state_double = false;
        rc = h->method == FPSFTWA_METHOD_GPSR ? run_gpsr_batched<float>(h, kt) : run_batched<float>(h, kt);
    }
    if (rc) {
        (void)hipStreamSynchronize(h->stream);
        release_graph(h);
        return rc;
    }
    HIP_TRY(whole.mark(h->stream));
    h->acc.resize((size_t)h->T * Q * 2);
    HIP_TRY(hipMemcpyAsync(h->acc.data(), h->acc_d.p, h->acc.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
    h->n_bases_run = h->n_bases;
    h->basis_acc.assign((size_t)h->n_bases * 2 * h->L * h->T, 0.0);
    if (h->n_bases)
        HIP_TRY(hipMemcpyAsync(h->basis_acc.data(), h->basis_out.p, h->basis_acc.size() * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream));
    const hipError_t es = hipStreamSynchronize(h->stream);
    release_graph(h);
    HIP_TRY(es);
    h->kernel_ms = kt.pairs_ms();
    h->run_ms = whole.pairs_ms();
    h->B_run = h->B;
    h->method_run = h->method;
    h->has_result = true;
    return FPSFTWA_OK;
}

int fpsftwa_run_noalloc(fpsftwa_handle* h, void* state, size_t state_bytes, double* sums, size_t n_sums)
{
    if (int rc = check_handle(h)) return rc;
    if (!state || !sums) return fail(FPSFTWA_EINVAL, "state and sums must be device buffers of the caller");
    HIP_TRY(hipSetDevice(h->device));
    void* scr = nullptr;
    if (int rc = scratch_address(h->device, &scr)) return rc;
    if (h->engine == FPSFTWA_ENGINE_FP64_PLAIN)
        return fail(FPSFTWA_EUNSUPPORTED, "fpsftwa_run_noalloc runs the batched engines (FP32, FP64) only");
    if (h->method != FPSFTWA_METHOD_FTWA)
        return fail(FPSFTWA_EUNSUPPORTED, "fpsftwa_run_noalloc runs the fTWA method only");
    if (!h->noise.empty() || !h->rot_up.empty() || h->n_bases > 0)
        return fail(FPSFTWA_EUNSUPPORTED, "fpsftwa_run_noalloc takes the generator's noise, diagonal initial states "
                                          "(no rotations) and no basis occupations");
    h->has_result = false;
    h->T = (int)std::floor(h->t_max / h->dt + 1.0);  // src/FHSolver.cpp:100
    h->launches = 0;
    h->kernel_ms = h->run_ms = 0.0;
    int rc = h->engine == FPSFTWA_ENGINE_FP32_BATCHED
                 ? run_noalloc<float>(h, state, state_bytes, sums, n_sums, static_cast<char*>(scr))
                 : run_noalloc<double>(h, state, state_bytes, sums, n_sums, static_cast<char*>(scr));
    if (rc) (void)hipStreamSynchronize(h->stream);
    return rc;
}

int fpsftwa_run_adjoint(fpsftwa_handle* h, const fpsftwa_adjoint_spec* spec, double* objective, double* grad_u,
                        double* grad_t, double* grad_pulse, size_t n_pulse)
{
    if (int rc = check_handle(h)) return rc;
    if (spec && spec->objective != FPSFTWA_OBJ_DOUBLE_OCCUPANCY) return fail(FPSFTWA_EINVAL, "unknown objective");
    if (h->coulomb_type == FPSFTWA_COULOMB_GENERAL)
        return fail(FPSFTWA_EUNSUPPORTED, "the adjoint supports null and on-site Coulomb interactions");
    if (h->method != FPSFTWA_METHOD_FTWA) return fail(FPSFTWA_EUNSUPPORTED, "the adjoint is defined for fTWA runs");
    if (h->L > 256) return fail(FPSFTWA_EUNSUPPORTED, "the adjoint supports L <= 256");
    if (grad_pulse && n_pulse) {
        const size_t need = h->pulse_type == FPSFTWA_PULSE_KNOTS ? h->knots.size()
                                                                 : (h->pulse_type == FPSFTWA_PULSE_GAUSSIAN ? 1 : 0);
        if (n_pulse != need) return fail(FPSFTWA_EINVAL, "grad_pulse length does not match the pulse parameters");
    }
    if (!h->noise.empty()) {
        const size_t need = (size_t)h->B * 2 * ((size_t)h->L * (h->L - 1) / 2) * 2;
        if (h->noise.size() != need) return fail(FPSFTWA_EINVAL, "noise size does not match n_traj");
    }
    HIP_TRY(hipSetDevice(h->device));
    h->has_result = false;
    h->T = (int)std::floor(h->t_max / h->dt + 1.0);  // src/FHSolver.cpp:100
    int rc = prepare_initial(h);
    if (rc) return rc;
    const int ck = spec && spec->checkpoint_every > 0 ? spec->checkpoint_every : 50;
    if (h->engine == FPSFTWA_ENGINE_FP32_BATCHED) {
        h->state_double = false;
        rc = run_adjoint<float>(h, ck, objective, grad_u, grad_t, grad_pulse, n_pulse);
    } else {
        h->state_double = true;
        rc = run_adjoint<double>(h, ck, objective, grad_u, grad_t, grad_pulse, n_pulse);
    }
    (void)hipStreamSynchronize(h->stream);
    return rc;
}

int fpsftwa_get_timing(fpsftwa_handle* h, int64_t* launches, double* kernel_ms, double* run_ms)
{
    if (int rc = check_handle(h)) return rc;
    if (!h->has_result) return fail(FPSFTWA_ESTATE, "no completed run");
    if (launches) *launches = h->launches;
    if (kernel_ms) *kernel_ms = h->kernel_ms;
    if (run_ms) *run_ms = h->run_ms;
    return FPSFTWA_OK;
}

int fpsftwa_get_result(fpsftwa_handle* h, int32_t which, double* out, size_t n_values, size_t* needed)
{
    if (int rc = check_handle(h)) return rc;
    if (!h->has_result) return fail(FPSFTWA_ESTATE, "no completed run");
    const int L = h->L, T = h->T, Q = 6 * L + 12;
    const double B = (double)h->B_run;
    size_t n = 0;
    switch (which) {
    case FPSFTWA_OUT_TIME: n = T; break;
    case FPSFTWA_OUT_ENERGY: n = 5 * (size_t)T; break;
    case FPSFTWA_OUT_ENERGY_ERR:
    case FPSFTWA_OUT_SUM_ENERGY: n = 4 * (size_t)T; break;
    case FPSFTWA_OUT_N1HU:
    case FPSFTWA_OUT_N1HD: n = 2 * (size_t)T; break;
    case FPSFTWA_OUT_BASIS:
    case FPSFTWA_OUT_SUM_BASIS:
        if (h->n_bases_run == 0) return fail(FPSFTWA_ESTATE, "the run recorded no basis occupations (fpsftwa_set_bases)");
        n = h->basis_acc.size();
        break;
    default:
        if (which < 0 || which >= FPSFTWA_OUT_COUNT) return fail(FPSFTWA_EINVAL, "unknown result id");
        n = (size_t)L * T;
    }
    if (needed) *needed = n;
    if (!out) return FPSFTWA_OK;
    if (n_values < n) return fail(FPSFTWA_EINVAL, "output buffer too small");
    // GPSR's rho is normally ordered: plain means, no 1/2 (src/FHGPSRSolver.cpp:895-929)
    const bool gpsr = h->method_run == FPSFTWA_METHOD_GPSR;
    const double hf = gpsr ? 0.0 : 0.5;
    if (which == FPSFTWA_OUT_BASIS || which == FPSFTWA_OUT_SUM_BASIS) {
        for (size_t e = 0; e < n; ++e)  // Mkk = mkk / nTraj + 1/2 (src/FHfTWASolver.cpp:531-540)
            out[e] = which == FPSFTWA_OUT_BASIS ? h->basis_acc[e] / B + hf : h->basis_acc[e];
        return FPSFTWA_OK;
    }
    auto S = [&](int t, int q) { return kahan_value(h->acc, Q, t, q); };
    const bool on_diag = h->rec_i == (h->rec_j < 0 ? L - 1 : h->rec_j);
    auto stderr_of = [&](double s1, double s2) {
        if (B < 2) return 0.0;
        double var = (s2 - s1 * s1 / B) / (B - 1.0);
        return var > 0 ? std::sqrt(var / B) : 0.0;
    };
    for (int t = 0; t < T; ++t) {
        switch (which) {
        case FPSFTWA_OUT_TIME: out[t] = t * h->dt; break;
        case FPSFTWA_OUT_ENERGY: {
            double tot = 0.0;
            for (int k = 0; k < 4; ++k) {
                const double e = S(t, 6 * L + k) / B;
                out[(size_t)k * T + t] = e;
                tot += e;
            }
            out[(size_t)4 * T + t] = tot;
            break;
        }
        case FPSFTWA_OUT_ENERGY_ERR:
            for (int k = 0; k < 4; ++k) out[(size_t)k * T + t] = stderr_of(S(t, 6 * L + k), S(t, 6 * L + 4 + k));
            break;
        case FPSFTWA_OUT_SUM_ENERGY:
            for (int k = 0; k < 4; ++k) out[(size_t)k * T + t] = S(t, 6 * L + k);
            break;
        case FPSFTWA_OUT_N1HU:
            out[t] = S(t, 6 * L + 8) / B + (on_diag ? hf : 0.0);
            out[T + t] = S(t, 6 * L + 9) / B;
            break;
        case FPSFTWA_OUT_N1HD:
            out[t] = S(t, 6 * L + 10) / B + (on_diag ? hf : 0.0);
            out[T + t] = S(t, 6 * L + 11) / B;
            break;
        default:
            for (int i = 0; i < L; ++i) {
                const double su = S(t, 6 * i), sd = S(t, 6 * i + 1), suu = S(t, 6 * i + 2), sdd = S(t, 6 * i + 3),
                             sud = S(t, 6 * i + 4), snn = S(t, 6 * i + 5);
                double v = 0.0;
                if (gpsr) {
                    // Niiu = <rho_ii>, N2iiu = <|rho_ii|^2>, NuNd = <rho_up_ii rho_dn_ii>, real parts
                    // (src/FHGPSRSolver.cpp:895-914); the error of Niiu uses |rho_ii|^2, an upper bound of
                    // (Re rho_ii)^2
                    switch (which) {
                    case FPSFTWA_OUT_NIIU: v = su / B; break;
                    case FPSFTWA_OUT_NIID: v = sd / B; break;
                    case FPSFTWA_OUT_N2IIU: v = suu / B; break;
                    case FPSFTWA_OUT_N2IID: v = sdd / B; break;
                    case FPSFTWA_OUT_NUND: v = sud / B; break;
                    case FPSFTWA_OUT_NIIU_ERR: v = stderr_of(su, suu); break;
                    case FPSFTWA_OUT_NIID_ERR: v = stderr_of(sd, sdd); break;
                    case FPSFTWA_OUT_NUND_ERR: v = stderr_of(sud, snn); break;
                    case FPSFTWA_OUT_SUM_RHO_U: v = su; break;
                    case FPSFTWA_OUT_SUM_RHO_D: v = sd; break;
                    case FPSFTWA_OUT_SUM_RHO2_U: v = suu; break;
                    case FPSFTWA_OUT_SUM_RHO2_D: v = sdd; break;
                    case FPSFTWA_OUT_SUM_RHOUD: v = sud; break;
                    default: break;
                    }
                    out[(size_t)i * T + t] = v;
                    continue;
                }
                switch (which) {
                case FPSFTWA_OUT_NIIU: v = su / B + 0.5; break;              // src/FHfTWASolver.cpp:509
                case FPSFTWA_OUT_NIID: v = sd / B + 0.5; break;              // :510
                case FPSFTWA_OUT_N2IIU: v = suu / B + (2.0 * su) / (2.0 * B) + 0.25; break;  // :520
                case FPSFTWA_OUT_N2IID: v = sdd / B + (2.0 * sd) / (2.0 * B) + 0.25; break;  // :521
                case FPSFTWA_OUT_NUND: v = sud / B + (su + sd) / (2.0 * B) + 0.25; break;    // :530
                case FPSFTWA_OUT_NIIU_ERR: v = stderr_of(su, suu); break;
                case FPSFTWA_OUT_NIID_ERR: v = stderr_of(sd, sdd); break;
                case FPSFTWA_OUT_NUND_ERR: {
                    const double sx = sud + (su + sd) / 2.0 + B / 4.0;  // sum of (u+1/2)(d+1/2)
                    v = stderr_of(sx, snn);
                    break;
                }
                case FPSFTWA_OUT_SUM_RHO_U: v = su; break;
                case FPSFTWA_OUT_SUM_RHO_D: v = sd; break;
                case FPSFTWA_OUT_SUM_RHO2_U: v = suu; break;
                case FPSFTWA_OUT_SUM_RHO2_D: v = sdd; break;
                case FPSFTWA_OUT_SUM_RHOUD: v = sud; break;
                default: break;
                }
                out[(size_t)i * T + t] = v;
            }
        }
    }
    return FPSFTWA_OK;
}

int fpsftwa_get_states(fpsftwa_handle* h, int64_t first, int64_t count, double* out, size_t n_values)
{
    if (int rc = check_handle(h)) return rc;
    if (!h->has_result) return fail(FPSFTWA_ESTATE, "no completed run");
    if (first < 0 || count < 0 || first + count > h->B_run || !out) return fail(FPSFTWA_EINVAL, "bad range");
    const size_t per = 2 * (size_t)h->L * h->L * 2;
    if (n_values < per * count) return fail(FPSFTWA_EINVAL, "output too small");
    HIP_TRY(hipSetDevice(h->device));
    if (h->state_double) {
        HIP_TRY(hipMemcpy(out, h->state.as<double>() + (size_t)first * per, per * count * sizeof(double),
                          hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(per * count);
        HIP_TRY(hipMemcpy(tmp.data(), h->state.as<float>() + (size_t)first * per, per * count * sizeof(float),
                          hipMemcpyDeviceToHost));
        for (size_t k = 0; k < tmp.size(); ++k) out[k] = tmp[k];
    }
    return FPSFTWA_OK;
}

}  // extern "C"

namespace fpsamd {
void set_last_error(const std::string& msg) { g_last_error = msg; }

double pulse_amplitude(const PulseSpec& p, double t)
{
    if (p.type == FPSFTWA_PULSE_GAUSSIAN) {  // src/FHExternalPotentialPulse.cpp:88
        const double x = (t - p.t0) / p.td;
        return p.A0 * std::exp(-(x * x) / 2.0) * std::cos(p.w0 * (t - p.t0) + p.phi);
    }
    if (p.type == FPSFTWA_PULSE_KNOTS) {  // piecewise linear through knots at k*t_max/(n-1)
        const int n = p.n_knots;
        if (n <= 0) return 0.0;
        if (n == 1) return p.knots[0];
        const double step = p.t_max / (double)(n - 1);
        const double u = t / step;
        // t_max == 0 puts every knot at t = 0 (u is 0/0 or infinite): the first knot holds
        if (!(step > 0.0) || !std::isfinite(u) || u <= 0.0) return p.knots[0];
        const int k = (int)std::floor(u);
        if (k >= n - 1) return p.knots[n - 1];
        const double f = u - (double)k;
        return p.knots[k] * (1.0 - f) + p.knots[k + 1] * f;
    }
    return 0.0;
}
}  // namespace fpsamd

extern "C" int fpsftwa_set_tuning(fpsftwa_handle* h, int32_t block_threads, int32_t time_chunk, int32_t kernel)
{
    if (int rc = check_handle(h)) return rc;
    if (kernel < FPSFTWA_KERNEL_AUTO || kernel > FPSFTWA_KERNEL_DENSE_MFMA) return fail(FPSFTWA_EINVAL, "bad kernel");
    h->kernel = kernel;
    if ((block_threads != 0 && block_threads != 512 && block_threads != 832 && block_threads != 1024) || time_chunk < 0)
        return fail(FPSFTWA_EINVAL, "block_threads must be 0 (automatic), 512, 832 or 1024, and time_chunk >= 0");
    h->nt_hint = block_threads;
    h->time_chunk = time_chunk;
    return FPSFTWA_OK;
}


// After a trial run on the dense path (which leaves its buffers filled), pick the fastest library
// solution for the stage GEMM and re-time the trial; returns its rate, or -1 when nothing was tuned.
template <typename Real>
double tune_dense_gemm_t(fpsftwa_handle* h, int trial_steps)
{
    DevSystem<Real> ds;
    if (upload_system<Real>(h, ds, h->T) != FPSFTWA_OK || !h->dense || !dense_uses_gemm<Real>(ds)) return -1.0;
    int sol = 0, tried = 0;
    double ms = 0.0;
    if (dense_tune_gemm<Real>(h->dense, ds, h->B, h->dense_buf.as<Real>(), &sol, &ms, &tried) != hipSuccess)
        return -1.0;
    if (fpsftwa_run(h) != FPSFTWA_OK) return -1.0;
    return (double)h->B * trial_steps / std::max(h->kernel_ms * 1e-3, 1e-12);
}

double tune_dense_gemm(fpsftwa_handle* h, int trial_steps)
{
    if (fpsftwa_run(h) != FPSFTWA_OK) return -1.0;  // fills the dense buffers for this trial batch
    return h->engine == FPSFTWA_ENGINE_FP32_BATCHED ? tune_dense_gemm_t<float>(h, trial_steps)
                                                    : tune_dense_gemm_t<double>(h, trial_steps);
}

extern "C" int fpsftwa_autotune(fpsftwa_handle* h, int32_t trial_steps, int32_t* kernel, int32_t* block_threads,
                                int32_t* time_chunk, double* traj_steps_per_s)
{
    if (int rc = check_handle(h)) return rc;
    if (trial_steps < 1) return fail(FPSFTWA_EINVAL, "trial_steps must be >= 1");
    if (h->engine == FPSFTWA_ENGINE_FP64_PLAIN) return fail(FPSFTWA_EINVAL, "the plain engine has no tuning");
    if (h->method == FPSFTWA_METHOD_GPSR) return fail(FPSFTWA_EINVAL, "GPSR has a single batched kernel");
    struct Cand {
        int kernel, block, chunk;
    };
    std::vector<Cand> cands;
    const bool fp32 = h->engine == FPSFTWA_ENGINE_FP32_BATCHED;
    cands.push_back({FPSFTWA_KERNEL_HERMITIAN, 0, 0});
    if (fp32) {
        cands.push_back({FPSFTWA_KERNEL_PIPELINED, 0, 0});
        cands.push_back({FPSFTWA_KERNEL_SEQUENTIAL, 0, 0});
    }
    else cands.push_back({FPSFTWA_KERNEL_HERMITIAN, 1024, 0});
    for (int blk : {0, 512, 1024}) cands.push_back({FPSFTWA_KERNEL_FULL, blk, 0});
    cands.push_back({FPSFTWA_KERNEL_DENSE, 0, 0});
    if (fp32) {
        cands.push_back({FPSFTWA_KERNEL_HERMITIAN, 0, 8});
        cands.push_back({FPSFTWA_KERNEL_DENSE_MFMA, 0, 0});  // the fused matrix-core form of the dense path
    }
    // the automatic choice last, kept only when strictly faster than every named kernel: it holds a form no
    // named kernel selects (the direct-commutator kernel on the 8x8 lattice)
    cands.push_back({FPSFTWA_KERNEL_AUTO, 0, 0});
    // a shortened copy of the configured run: trial_steps steps of at most 16,384 trajectories
    const double t_max = h->t_max;
    const int64_t B = h->B;
    const int nsub = h->nsub;
    const std::vector<double> noise = h->noise;
    const int saved_kernel = h->kernel, saved_block = h->nt_hint, saved_chunk = h->time_chunk;
    h->t_max = h->dt * trial_steps;
    h->B = std::min<int64_t>(B, 16384);
    h->nsub = 1;
    if (!noise.empty()) h->noise.clear();
    double best = -1.0;
    Cand pick{saved_kernel, saved_block, saved_chunk};
    const std::string keep_err = g_last_error;
    for (const Cand& c : cands) {
        h->kernel = c.kernel;
        h->nt_hint = c.block;
        h->time_chunk = c.chunk;
        double rate = -1.0;
        for (int rep = 0; rep < 2; ++rep) {  // the first run pays one-time costs (library set-up)
            if (fpsftwa_run(h) != FPSFTWA_OK) {
                rate = -1.0;
                break;
            }
            rate = (double)h->B * trial_steps / std::max(h->kernel_ms * 1e-3, 1e-12);
        }
        if (rate > best) {
            best = rate;
            pick = c;
        }
    }
    g_last_error = keep_err;  // unsupported candidates are expected failures
    if (best > 0.0 && pick.kernel == FPSFTWA_KERNEL_DENSE) {
        // the dense path's stage GEMM: time the library's solutions on the trial batch
        h->kernel = pick.kernel;
        h->nt_hint = pick.block;
        h->time_chunk = pick.chunk;
        const double rate = tune_dense_gemm(h, trial_steps);
        if (rate > best) best = rate;
        g_last_error = keep_err;
    }
    h->t_max = t_max;
    h->B = B;
    h->nsub = nsub;
    h->noise = noise;
    h->has_result = false;
    if (best <= 0.0) {
        h->kernel = saved_kernel;
        h->nt_hint = saved_block;
        h->time_chunk = saved_chunk;
        return fail(FPSFTWA_EUNSUPPORTED, "no batched kernel runs this system");
    }
    h->kernel = pick.kernel;
    h->nt_hint = pick.block;
    h->time_chunk = pick.chunk;
    if (kernel) *kernel = pick.kernel;
    if (block_threads) *block_threads = pick.block;
    if (time_chunk) *time_chunk = pick.chunk;
    if (traj_steps_per_s) *traj_steps_per_s = best;
    return FPSFTWA_OK;
}

extern "C" int fpsftwa_run_many(fpsftwa_handle** handles, int32_t n)
{
    if (!handles || n < 0) return fail(FPSFTWA_EINVAL, "bad handle list");
    for (int k = 0; k < n; ++k)
        for (int m = 0; m < k; ++m)
            if (handles[k] == handles[m]) return fail(FPSFTWA_EINVAL, "a handle appears twice");
    std::vector<int> rc(n, FPSFTWA_OK);
    std::vector<std::string> msg(n);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int k = 0; k < n; ++k)
        th.emplace_back([&, k]() {
            rc[k] = fpsftwa_run(handles[k]);
            if (rc[k] != FPSFTWA_OK) msg[k] = g_last_error;  // thread-local message of that thread
        });
    for (auto& t : th) t.join();
    for (int k = 0; k < n; ++k)
        if (rc[k] != FPSFTWA_OK) return fail(rc[k], "system " + std::to_string(k) + ": " + msg[k]);
    return FPSFTWA_OK;
}
