"""ctypes binding of the CPU FP64 fTWA / GPSR oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker or the timed CPU baseline, never by the product package.

The C restatement follows the reference FHfTWASolver (src/FHfTWASolver.cpp:74-213 initiateSolver,
:222-556 runSimulation, :568-600 derivative, :629-661 step_RK4_38); `normalise` restates the
reference's ensemble normalisation (src/FHfTWASolver.cpp:489-551).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.POINTER
    lib.oracle_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
    lib.oracle_n_time_points.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_n_time_points.restype = ctypes.c_int64
    lib.oracle_noise.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                 P(ctypes.c_double)]
    lib.oracle_pulse_amplitude.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    lib.oracle_pulse_amplitude.restype = ctypes.c_double
    lib.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int,
                               ctypes.c_void_p]
    lib.oracle_run.restype = ctypes.c_int
    lib.oracle_run_seeded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_void_p]
    lib.oracle_run_seeded.restype = ctypes.c_int
    lib.oracle_derivative.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                      P(ctypes.c_double), P(ctypes.c_double),
                                      P(ctypes.c_double), P(ctypes.c_double)]
    lib.oracle_gpsr_noise.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                      P(ctypes.c_double)]
    lib.oracle_gpsr_diffusion.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double),
                                          P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double)]
    _lib = lib
    return lib


class _OrSystem(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int),
        ("tM", ctypes.POINTER(ctypes.c_double)),
        ("coulomb_type", ctypes.c_int),
        ("U", ctypes.POINTER(ctypes.c_double)),
        ("V", ctypes.POINTER(ctypes.c_double)),
        ("pot", ctypes.POINTER(ctypes.c_double)),
        ("pulse_type", ctypes.c_int),
        ("A0", ctypes.c_double), ("t0", ctypes.c_double), ("td", ctypes.c_double),
        ("w0", ctypes.c_double), ("phi", ctypes.c_double),
        ("pol", ctypes.c_double * 3),
        ("n_knots", ctypes.c_int),
        ("knots", ctypes.POINTER(ctypes.c_double)),
        ("disp", ctypes.POINTER(ctypes.c_double)),
        ("fock", ctypes.c_int),
        ("gpsr", ctypes.c_int),
        ("gpsr_seed", ctypes.c_uint64),
    ]


class _OrParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("t_max", ctypes.c_double),
        ("n_sub", ctypes.c_int), ("n_traj", ctypes.c_int64),
        ("n_up", ctypes.POINTER(ctypes.c_double)), ("n_dn", ctypes.POINTER(ctypes.c_double)),
        ("rot_up", ctypes.POINTER(ctypes.c_double)), ("rot_dn", ctypes.POINTER(ctypes.c_double)),
        ("n_bases", ctypes.c_int), ("bases", ctypes.POINTER(ctypes.c_double)),
        ("corner_i", ctypes.c_int), ("corner_j", ctypes.c_int),
    ]


class _OrOut(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in
                ("sum_u", "sum_d", "sum2_u", "sum2_d", "sum_ud", "sum_e", "sum_corner", "final_states",
                 "sum_basis")]


def _ptr(a):
    if a is None:
        return ctypes.POINTER(ctypes.c_double)()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class System:
    """Plain-array fTWA system: hopping tM, U, V, constant potential, optional pulse."""
    tM: np.ndarray
    coulomb_type: int = 1
    U: np.ndarray | None = None
    V: np.ndarray | None = None
    pot: np.ndarray | None = None
    pulse_type: int = 0
    A0: float = 0.0
    t0: float = 0.0
    td: float = 1.0
    w0: float = 0.0
    phi: float = 0.0
    pol: tuple = (0.0, 0.0, 0.0)
    knots: np.ndarray | None = None
    disp: np.ndarray | None = None
    fock: bool = False  # Hartree-Fock dynamics (exchange with V), with zero noise: the HF solver
    gpsr: bool = False  # Gaussian phase-space representation (the GPSR solver), diffusion keyed by gpsr_seed
    gpsr_seed: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def L(self) -> int:
        return int(self.tM.shape[0])

    def _c(self) -> _OrSystem:
        L = self.L
        arrs = {}
        for name in ("tM", "U", "V", "pot", "knots", "disp"):
            v = getattr(self, name)
            arrs[name] = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        if arrs["U"] is None:
            arrs["U"] = np.zeros((L, L))
        if arrs["V"] is None:
            arrs["V"] = np.zeros((L, L))
        if arrs["disp"] is None:
            arrs["disp"] = np.zeros((L, L, 3))
        self._keep = list(arrs.values())
        s = _OrSystem()
        s.L = L
        s.tM = _ptr(arrs["tM"])
        s.coulomb_type = int(self.coulomb_type)
        s.fock = 1 if self.fock else 0
        s.gpsr = 1 if self.gpsr else 0
        s.gpsr_seed = int(self.gpsr_seed) & 0xFFFFFFFFFFFFFFFF
        s.U = _ptr(arrs["U"])
        s.V = _ptr(arrs["V"])
        s.pot = _ptr(arrs["pot"])
        s.pulse_type = int(self.pulse_type)
        s.A0, s.t0, s.td, s.w0, s.phi = self.A0, self.t0, self.td, self.w0, self.phi
        s.pol = (ctypes.c_double * 3)(*self.pol)
        s.n_knots = 0 if arrs["knots"] is None else int(arrs["knots"].size)
        s.knots = _ptr(arrs["knots"])
        s.disp = _ptr(arrs["disp"])
        return s


def philox4x32_10(ctr, key) -> np.ndarray:
    lib = _load()
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib.oracle_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def n_time_points(dt: float, t_max: float) -> int:
    return int(_load().oracle_n_time_points(dt, t_max))


def noise(L: int, first: int, count: int, seed: int) -> np.ndarray:
    """Counter-based noise, [count][2][L(L-1)/2][2] (re, im)."""
    npairs = L * (L - 1) // 2
    out = np.zeros((count, 2, max(npairs, 0), 2))
    _load().oracle_noise(L, first, count, seed, _ptr(out))
    return out


def pulse_amplitude(system: System, t: float, t_max: float) -> float:
    s = system._c()
    return float(_load().oracle_pulse_amplitude(ctypes.byref(s), t, t_max))


def derivative(system: System, t: float, t_max: float, rho_u: np.ndarray, rho_d: np.ndarray):
    """f = -i(rho Delta - Delta rho) for complex [L][L] rho_up, rho_dn."""
    L = system.L
    s = system._c()
    ru = np.ascontiguousarray(np.stack([rho_u.real, rho_u.imag], -1), dtype=np.float64)
    rd = np.ascontiguousarray(np.stack([rho_d.real, rho_d.imag], -1), dtype=np.float64)
    fu = np.zeros((L, L, 2))
    fd = np.zeros((L, L, 2))
    _load().oracle_derivative(ctypes.byref(s), t, t_max, _ptr(ru), _ptr(rd), _ptr(fu), _ptr(fd))
    return fu[..., 0] + 1j * fu[..., 1], fd[..., 0] + 1j * fd[..., 1]


def gpsr_noise(L: int, traj: int, step: int, seed: int) -> np.ndarray:
    """Unscaled GPSR diffusion noise of one step, complex [2][L] (xi_1, xi_2)."""
    out = np.zeros((2, L, 2))
    _load().oracle_gpsr_noise(L, traj, step, seed, _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def gpsr_diffusion(system: System, rho_u: np.ndarray, rho_d: np.ndarray, xi: np.ndarray):
    """GPSR diffusion increments (d_up, d_dn) at complex [L][L] rho for the scaled noise xi [2][L]."""
    L = system.L
    s = system._c()

    def c2(a):
        a = np.asarray(a, dtype=np.complex128)
        return np.ascontiguousarray(np.stack([a.real, a.imag], -1), dtype=np.float64)

    ru, rd, x = c2(rho_u), c2(rho_d), c2(xi)
    du = np.zeros((L, L, 2))
    dd = np.zeros((L, L, 2))
    _load().oracle_gpsr_diffusion(ctypes.byref(s), _ptr(ru), _ptr(rd), _ptr(x), _ptr(du), _ptr(dd))
    return du[..., 0] + 1j * du[..., 1], dd[..., 0] + 1j * dd[..., 1]


@dataclass
class Sums:
    T: int
    sum_u: np.ndarray
    sum_d: np.ndarray
    sum2_u: np.ndarray
    sum2_d: np.ndarray
    sum_ud: np.ndarray
    sum_e: np.ndarray
    sum_corner: np.ndarray
    final_states: np.ndarray | None
    sum_basis: np.ndarray | None = None  # [n_bases][2][L][T]


def run(system: System, dt: float, t_max: float, n_sub: int, n_up, n_dn, *, noise_arr=None,
        n_traj: int | None = None, seed: int | None = None, rot_up=None, rot_dn=None,
        nthreads: int = 0, keep_states: bool = False, bases=None, corner=None) -> Sums:
    """Run the oracle on explicit noise ([B][2][npairs][2]) or on the seeded generator.

    bases: optional [n_bases][2][L][L] (real or complex) basis matrices whose occupations
    diag(V^dagger rho V) are summed into Sums.sum_basis.  corner: the (i, j) element summed into
    Sums.sum_corner (default (0, L-1))."""
    lib = _load()
    L = system.L
    T = n_time_points(dt, t_max)
    if noise_arr is not None:
        noise_arr = np.ascontiguousarray(noise_arr, dtype=np.float64)
        B = noise_arr.shape[0]
    else:
        assert seed is not None and n_traj is not None
        B = int(n_traj)
    nu = np.ascontiguousarray(n_up, dtype=np.float64)
    nd = np.ascontiguousarray(n_dn, dtype=np.float64)

    def cplx(r):
        if r is None:
            return None
        r = np.asarray(r, dtype=np.complex128)
        return np.ascontiguousarray(np.stack([r.real, r.imag], -1))

    ru, rd = cplx(rot_up), cplx(rot_dn)
    bs = cplx(bases)
    nb = 0 if bases is None else int(np.asarray(bases).shape[0])
    p = _OrParams()
    p.dt, p.t_max, p.n_sub, p.n_traj = dt, t_max, n_sub, B
    p.n_up, p.n_dn, p.rot_up, p.rot_dn = _ptr(nu), _ptr(nd), _ptr(ru), _ptr(rd)
    p.n_bases, p.bases = nb, _ptr(bs)
    p.corner_i, p.corner_j = (0, -1) if corner is None else (int(corner[0]), int(corner[1]))
    arrays = {n: np.zeros((L, T)) for n in ("sum_u", "sum_d", "sum2_u", "sum2_d", "sum_ud")}
    arrays["sum_e"] = np.zeros((4, T))
    arrays["sum_corner"] = np.zeros((4, T))
    fs = np.zeros((B, 2, L, L, 2)) if keep_states else None
    o = _OrOut()
    for k, v in arrays.items():
        setattr(o, k, _ptr(v))
    o.final_states = _ptr(fs)
    sb = np.zeros((nb, 2, L, T)) if nb else None
    o.sum_basis = _ptr(sb)
    s = system._c()
    if noise_arr is not None:
        rc = lib.oracle_run(ctypes.byref(s), ctypes.byref(p), _ptr(noise_arr), nthreads, ctypes.byref(o))
    else:
        rc = lib.oracle_run_seeded(ctypes.byref(s), ctypes.byref(p), int(seed), nthreads, ctypes.byref(o))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed ({rc})")
    final = None if fs is None else fs[..., 0] + 1j * fs[..., 1]
    return Sums(T=T, final_states=final, sum_basis=sb, **arrays)


def normalise(s: Sums, B: int, gpsr: bool = False) -> dict:
    """The reference's ensemble normalisation (src/FHfTWASolver.cpp:506-547); GPSR: plain averages of the
    normally ordered quantities (src/FHGPSRSolver.cpp:895-929)."""
    if gpsr:
        out = {"Niiu": s.sum_u / B, "Niid": s.sum_d / B, "N2iiu": s.sum2_u / B, "N2iid": s.sum2_d / B,
               "NuNd": s.sum_ud / B}
        E = s.sum_e / B
        out["E"] = np.vstack([E, E.sum(axis=0, keepdims=True)])
        if s.sum_basis is not None:
            out["M"] = s.sum_basis / B
        return out
    out = {
        "Niiu": s.sum_u / B + 0.5,
        "Niid": s.sum_d / B + 0.5,
        "N2iiu": s.sum2_u / B + (2.0 * s.sum_u) / (2 * B) + 0.25,
        "N2iid": s.sum2_d / B + (2.0 * s.sum_d) / (2 * B) + 0.25,
        "NuNd": s.sum_ud / B + (s.sum_u + s.sum_d) / (2 * B) + 0.25,
    }
    E = s.sum_e / B
    out["E"] = np.vstack([E, E.sum(axis=0, keepdims=True)])
    if s.sum_basis is not None:  # Mkk = mkk / nTraj + 1/2 (src/FHfTWASolver.cpp:531-540)
        out["M"] = s.sum_basis / B + 0.5
    return out


def gauge_columns(V: np.ndarray) -> np.ndarray:
    """Fix each eigenvector's sign: its largest-magnitude entry (the first one within 1e-8 on ties) positive."""
    V = np.array(V, dtype=np.float64, copy=True)
    for c in range(V.shape[1]):
        big = 0
        for k in range(1, V.shape[0]):
            if abs(V[k, c]) > abs(V[big, c]) * (1.0 + 1e-8):
                big = k
        if V[big, c] < 0:
            V[:, c] = -V[:, c]
    return V


def ground_state(tM, U, n_up: int, n_dn: int, mean_field: bool, energies: bool = False):
    """Eigenbasis of the tbGroundState / mfGroundState initial states with LAPACK (numpy.linalg.eigh).

    tight binding: eig_sym(tM) (src/FHSystem.cpp:909-937); mean field: per spin eig_sym(tM + diag(U n_other))
    from zero densities, the lowest n_up / n_dn states occupied, plain substitution until
    |n_up' - n_up| + |n_dn' - n_dn| < 1e-4 L, at most 100 iterations (src/FHSystem.cpp:940-1019).
    Returns (rot_up, rot_dn, iterations, residue) with columns in the gauge of gauge_columns; with
    energies=True also (energy_up, energy_dn), the ascending eigenvalues the results file echoes
    (tbgs_E for both spins, or the last iteration's mfgs_Eu / mfgs_Ed; src/FHSystem.cpp:1143-1230).
    """
    tM = np.asarray(tM, dtype=np.float64)
    L = tM.shape[0]
    if not mean_field:
        E, V = np.linalg.eigh(tM)
        V = gauge_columns(V)
        return (V, V.copy(), 0, 0.0, E, E.copy()) if energies else (V, V.copy(), 0, 0.0)
    U = np.zeros((L, L)) if U is None else np.asarray(U, dtype=np.float64)
    nu, nd = np.zeros(L), np.zeros(L)
    it, res = 0, 0.0
    while it < 100:
        Eu, Vu = np.linalg.eigh(tM + np.diag(U @ nd))
        Ed, Vd = np.linalg.eigh(tM + np.diag(U @ nu))
        Vu, Vd = gauge_columns(Vu), gauge_columns(Vd)
        nu_new = (Vu[:, :n_up] ** 2).sum(axis=1)
        nd_new = (Vd[:, :n_dn] ** 2).sum(axis=1)
        it += 1
        res = float(np.linalg.norm(nu_new - nu) + np.linalg.norm(nd_new - nd))
        if res < 1e-4 * L:
            break
        nu, nd = nu_new, nd_new
    return (Vu, Vd, it, res, Eu, Ed) if energies else (Vu, Vd, it, res)
