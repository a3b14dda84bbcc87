"""ctypes binding of the C ABI in include/fps_ftwa.h (libfps_ftwa.so, built in-tree).

There is no fallback: if the library is missing the import of an engine raises, and if no gfx950
device is visible `Engine(...)` raises with the library's own message.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfps_ftwa.so")

OK = 0
EINVAL, EDEVICE, ENOMEM, ESTATE, EUNSUPPORTED, EIO = -1, -2, -3, -4, -5, -6

COULOMB_NULL, COULOMB_ONSITE, COULOMB_GENERAL = 0, 1, 2
PULSE_NONE, PULSE_GAUSSIAN, PULSE_KNOTS = 0, 1, 2
METHOD_FTWA, METHOD_HF, METHOD_GPSR = 0, 1, 2
ENGINE_FP64_PLAIN, ENGINE_FP64_BATCHED, ENGINE_FP32_BATCHED = 0, 1, 2
ENGINES = {"fp64_plain": ENGINE_FP64_PLAIN, "fp64": ENGINE_FP64_BATCHED, "fp32": ENGINE_FP32_BATCHED}
KERNEL_AUTO, KERNEL_FULL, KERNEL_HERMITIAN, KERNEL_DENSE, KERNEL_PIPELINED, KERNEL_SEQUENTIAL = 0, 1, 2, 3, 4, 5
KERNEL_DENSE_MFMA = 6
KERNELS = {"auto": KERNEL_AUTO, "full": KERNEL_FULL, "hermitian": KERNEL_HERMITIAN, "dense": KERNEL_DENSE,
           "pipelined": KERNEL_PIPELINED, "sequential": KERNEL_SEQUENTIAL, "dense_mfma": KERNEL_DENSE_MFMA}

RES_DIAG, RES_DIAG2, RES_DOUBLONS, RES_ENERGIES, RES_WHOLE = 0x01, 0x02, 0x04, 0x08, 0x10

OUT = {
    "time": 0, "Niiu": 1, "Niid": 2, "N2iiu": 3, "N2iid": 4, "NuNd": 5, "E": 6, "N1hu": 7, "N1hd": 8,
    "Niiu_err": 9, "Niid_err": 10, "NuNd_err": 11, "E_err": 12,
    "sum_u": 13, "sum_d": 14, "sum2_u": 15, "sum2_d": 16, "sum_ud": 17, "sum_e": 18,
    "basis": 19, "sum_basis": 20,
}

# every symbol include/fps_ftwa.h declares
EXPORTED = [
    "fpsftwa_last_error", "fpsftwa_version", "fpsftwa_device_info", "fpsftwa_create", "fpsftwa_destroy",
    "fpsftwa_set_params", "fpsftwa_set_results", "fpsftwa_set_initial_state", "fpsftwa_set_noise",
    "fpsftwa_sample_noise", "fpsftwa_run", "fpsftwa_get_result", "fpsftwa_get_states", "fpsftwa_get_timing",
    "fpsftwa_set_tuning", "fpsftwa_run_json_file", "fpsftwa_load_system_json", "fpsftwa_run_adjoint",
    "fpsftwa_autotune", "fpsftwa_run_many", "fpsftwa_load_initial_state_json", "fpsftwa_ground_state",
    "fpsftwa_set_bases", "fpsftwa_set_recorded_element", "fpsftwa_set_method", "fpsftwa_set_stream",
    "fpsftwa_measure_peaks", "fpsftwa_run_noalloc", "fpsftwa_set_graph",
]


class FtwaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _AdjointSpec(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_int32), ("checkpoint_every", ctypes.c_int32)]


class _Peaks(ctypes.Structure):
    _fields_ = [("hbm_copy_gbps", ctypes.c_double), ("copy_bytes", ctypes.c_double),
                ("fp32_fma_tflops", ctypes.c_double), ("fp64_fma_tflops", ctypes.c_double),
                ("fp32_mfma_tflops", ctypes.c_double)]


class _System(ctypes.Structure):
    _fields_ = [
        ("n_sites", ctypes.c_int32),
        ("hopping", ctypes.POINTER(ctypes.c_double)),
        ("coulomb_type", ctypes.c_int32),
        ("coulomb_u", ctypes.POINTER(ctypes.c_double)),
        ("coulomb_v", ctypes.POINTER(ctypes.c_double)),
        ("potential", ctypes.POINTER(ctypes.c_double)),
        ("pulse_type", ctypes.c_int32),
        ("pulse_A0", ctypes.c_double), ("pulse_t0", ctypes.c_double), ("pulse_td", ctypes.c_double),
        ("pulse_w0", ctypes.c_double), ("pulse_phi", ctypes.c_double),
        ("pulse_pol", ctypes.c_double * 3),
        ("pulse_n_knots", ctypes.c_int32),
        ("pulse_knots", ctypes.POINTER(ctypes.c_double)),
        ("displacement", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def load(path: str = LIB_PATH):
    """Load libfps_ftwa.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FtwaError(EDEVICE, f"{path} is missing: build it with `python -m fermiphasespace_amd.build`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 under the same soname.  Loaded first,
    # it also serves this library; loaded after the system runtime this library links, it would be a
    # second runtime in the process and fail to enumerate the device.  So torch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    d, i32, i64, u32, u64, sz = (ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32,
                                 ctypes.c_uint64, ctypes.c_size_t)
    H = ctypes.c_void_p
    sig = {
        "fpsftwa_last_error": ([], ctypes.c_char_p),
        "fpsftwa_version": ([P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)], ctypes.c_int),
        "fpsftwa_device_info": ([ctypes.c_char_p, sz, P(ctypes.c_int), P(d)], ctypes.c_int),
        "fpsftwa_measure_peaks": ([P(_Peaks)], ctypes.c_int),
        "fpsftwa_create": ([P(_System), P(H)], ctypes.c_int),
        "fpsftwa_destroy": ([H], ctypes.c_int),
        "fpsftwa_set_params": ([H, d, d, i32, i64, i32, u64, i32], ctypes.c_int),
        "fpsftwa_set_results": ([H, u32], ctypes.c_int),
        "fpsftwa_set_initial_state": ([H, P(d), P(d), P(d), P(d)], ctypes.c_int),
        "fpsftwa_set_noise": ([H, P(d), sz], ctypes.c_int),
        "fpsftwa_sample_noise": ([H, i64, i64, P(d), sz], ctypes.c_int),
        "fpsftwa_run": ([H], ctypes.c_int),
        "fpsftwa_get_result": ([H, i32, P(d), sz, P(sz)], ctypes.c_int),
        "fpsftwa_get_states": ([H, i64, i64, P(d), sz], ctypes.c_int),
        "fpsftwa_get_timing": ([H, P(i64), P(d), P(d)], ctypes.c_int),
        "fpsftwa_set_tuning": ([H, i32, i32, i32], ctypes.c_int),
        "fpsftwa_run_json_file": ([ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
        "fpsftwa_run_adjoint": ([H, P(_AdjointSpec), P(d), P(d), P(d), P(d), sz], ctypes.c_int),
        "fpsftwa_autotune": ([H, i32, P(i32), P(i32), P(i32), P(d)], ctypes.c_int),
        "fpsftwa_run_many": ([P(H), i32], ctypes.c_int),
        "fpsftwa_load_system_json": ([ctypes.c_char_p, P(i32), P(i32), P(d), P(d), P(d), P(d), P(d), P(d), P(d)],
                                     ctypes.c_int),
        "fpsftwa_load_initial_state_json": ([ctypes.c_char_p, P(i32), P(d), P(d), P(d), P(d)], ctypes.c_int),
        "fpsftwa_ground_state": ([i32, P(d), P(d), i32, i32, i32, P(d), P(d), P(i32), P(d), P(d), P(d)],
                                 ctypes.c_int),
        "fpsftwa_set_bases": ([H, i32, P(d)], ctypes.c_int),
        "fpsftwa_set_recorded_element": ([H, i32, i32], ctypes.c_int),
        "fpsftwa_set_stream": ([H, ctypes.c_void_p], ctypes.c_int),
        "fpsftwa_set_method": ([H, i32], ctypes.c_int),
        "fpsftwa_run_noalloc": ([H, ctypes.c_void_p, sz, ctypes.c_void_p, sz], ctypes.c_int),
        "fpsftwa_set_graph": ([H, i32], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def last_error() -> str:
    return load().fpsftwa_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        raise FtwaError(rc, f"{what}: {last_error()}" if what else last_error())


def _dptr(a):
    if a is None:
        return ctypes.POINTER(ctypes.c_double)()
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def device_info() -> dict:
    lib = load()
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int(0)
    gib = ctypes.c_double(0)
    check(lib.fpsftwa_device_info(name, 256, ctypes.byref(cu), ctypes.byref(gib)), "device_info")
    return {"name": name.value.decode(), "cu": cu.value, "hbm_gib": gib.value}


def measure_peaks() -> dict:
    """HBM copy bandwidth and FMA / MFMA rates measured on the current device (fpsftwa_measure_peaks)."""
    lib = load()
    pk = _Peaks()
    check(lib.fpsftwa_measure_peaks(ctypes.byref(pk)), "measure_peaks")
    return {name: getattr(pk, name) for name, _ in _Peaks._fields_}


def load_system_json(path: str) -> dict:
    """Matrices and occupations the JSON front end builds (host only, no device needed)."""
    lib = load()
    L = ctypes.c_int32(0)
    ct = ctypes.c_int32(0)
    P = ctypes.POINTER(ctypes.c_double)
    check(lib.fpsftwa_load_system_json(path.encode(), ctypes.byref(L), ctypes.byref(ct), P(), P(), P(), P(), P(),
                                       P(), P()), "load_system_json")
    n = L.value
    out = {k: np.zeros((n, n)) for k in ("tM", "U", "V")}
    out.update({k: np.zeros(n) for k in ("pot", "n_up", "n_dn")})
    out["positions"] = np.zeros((n, 3))
    check(lib.fpsftwa_load_system_json(path.encode(), ctypes.byref(L), ctypes.byref(ct), _dptr(out["tM"]),
                                       _dptr(out["U"]), _dptr(out["V"]), _dptr(out["pot"]), _dptr(out["n_up"]),
                                       _dptr(out["n_dn"]), _dptr(out["positions"])), "load_system_json")
    out["L"] = n
    out["coulomb_type"] = ct.value
    return out


def load_initial_state_json(path: str) -> dict:
    """Occupations and rotations of the initial state the JSON front end builds (host only)."""
    lib = load()
    L = ctypes.c_int32(0)
    P = ctypes.POINTER(ctypes.c_double)
    check(lib.fpsftwa_load_initial_state_json(path.encode(), ctypes.byref(L), P(), P(), P(), P()),
          "load_initial_state_json")
    n = L.value
    out = {"n_up": np.zeros(n), "n_dn": np.zeros(n), "rot_up": np.zeros((n, n, 2)), "rot_dn": np.zeros((n, n, 2))}
    check(lib.fpsftwa_load_initial_state_json(path.encode(), ctypes.byref(L), _dptr(out["n_up"]), _dptr(out["n_dn"]),
                                              _dptr(out["rot_up"]), _dptr(out["rot_dn"])), "load_initial_state_json")
    for k in ("rot_up", "rot_dn"):
        out[k] = out[k][..., 0] + 1j * out[k][..., 1]
    return out


def ground_state(tM, U=None, n_up=None, n_dn=None, mean_field=False) -> dict:
    """Tight-binding or mean-field eigenbasis (fpsftwa_ground_state): real [L][L] eigenvector columns,
    ascending eigenvalues (energy_up / energy_dn), each column's largest-magnitude entry positive.  Host only."""
    lib = load()
    tM = _f64(tM)
    L = tM.shape[0]
    U = _f64(U) if U is not None else (np.zeros((L, L)) if mean_field else None)
    n_up = int(np.ceil(L / 2)) if n_up is None else int(n_up)
    n_dn = int(np.floor(L / 2)) if n_dn is None else int(n_dn)
    ru, rd = np.zeros((L, L)), np.zeros((L, L))
    eu, ed = np.zeros(L), np.zeros(L)
    it, res = ctypes.c_int32(0), ctypes.c_double(0)
    check(lib.fpsftwa_ground_state(L, _dptr(tM), _dptr(U), n_up, n_dn, 1 if mean_field else 0, _dptr(ru), _dptr(rd),
                                   ctypes.byref(it), ctypes.byref(res), _dptr(eu), _dptr(ed)), "ground_state")
    return {"rot_up": ru, "rot_dn": rd, "energy_up": eu, "energy_dn": ed, "iterations": it.value,
            "residue": res.value}


def run_many(engines) -> None:
    """Run several engines at once, each on its own stream (fpsftwa_run_many)."""
    lib = load()
    arr = (ctypes.c_void_p * len(engines))(*[e._h.value for e in engines])
    check(lib.fpsftwa_run_many(arr, len(engines)), "run_many")


def run_json_file(path: str, out_path: str | None = None) -> None:
    check(load().fpsftwa_run_json_file(path.encode(), None if out_path is None else out_path.encode()),
          "run_json_file")


class Engine:
    """One handle of the C ABI: a system on the gfx950 device."""

    def __init__(self, tM, *, coulomb_type=COULOMB_ONSITE, U=None, V=None, pot=None, pulse_type=PULSE_NONE,
                 A0=0.0, t0=0.0, td=1.0, w0=0.0, phi=0.0, pol=(0.0, 0.0, 0.0), knots=None, disp=None):
        self._lib = load()
        tM = _f64(tM)
        L = tM.shape[0]
        self.L = L
        keep = [tM, _f64(U), _f64(V), _f64(pot), _f64(knots), _f64(disp)]
        s = _System()
        s.n_sites = L
        s.hopping = _dptr(keep[0])
        s.coulomb_type = coulomb_type
        s.coulomb_u = _dptr(keep[1])
        s.coulomb_v = _dptr(keep[2])
        s.potential = _dptr(keep[3])
        s.pulse_type = pulse_type
        s.pulse_A0, s.pulse_t0, s.pulse_td, s.pulse_w0, s.pulse_phi = A0, t0, td, w0, phi
        s.pulse_pol = (ctypes.c_double * 3)(*pol)
        s.pulse_n_knots = 0 if keep[4] is None else int(keep[4].size)
        s.pulse_knots = _dptr(keep[4])
        s.displacement = _dptr(keep[5])
        h = ctypes.c_void_p()
        check(self._lib.fpsftwa_create(ctypes.byref(s), ctypes.byref(h)), "create")
        self._n_pulse = (s.pulse_n_knots if pulse_type == PULSE_KNOTS else (1 if pulse_type == PULSE_GAUSSIAN else 0))
        self._h = h
        self.n_traj = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fpsftwa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, dt, t_max, n_sub, n_traj, seed=0, engine="fp64", int_order=4):
        e = ENGINES[engine] if isinstance(engine, str) else int(engine)
        check(self._lib.fpsftwa_set_params(self._h, dt, t_max, n_sub, n_traj, int_order, seed, e), "set_params")
        self.n_traj = n_traj

    def set_results(self, flags: int):
        check(self._lib.fpsftwa_set_results(self._h, flags), "set_results")

    def set_initial_state(self, n_up, n_dn, rot_up=None, rot_dn=None):
        nu, nd = _f64(n_up), _f64(n_dn)

        def cplx(r):
            if r is None:
                return None
            r = np.asarray(r, dtype=np.complex128)
            return np.ascontiguousarray(np.stack([r.real, r.imag], -1))

        ru, rd = cplx(rot_up), cplx(rot_dn)
        check(self._lib.fpsftwa_set_initial_state(self._h, _dptr(nu), _dptr(nd), _dptr(ru), _dptr(rd)),
              "set_initial_state")

    def set_noise(self, noise):
        if noise is None:
            check(self._lib.fpsftwa_set_noise(self._h, None, 0), "set_noise")
            return
        w = _f64(noise)
        check(self._lib.fpsftwa_set_noise(self._h, _dptr(w), w.size), "set_noise")

    def sample_noise(self, first, count):
        npairs = self.L * (self.L - 1) // 2
        out = np.zeros((count, 2, npairs, 2))
        check(self._lib.fpsftwa_sample_noise(self._h, first, count, _dptr(out), out.size), "sample_noise")
        return out

    def set_tuning(self, block_threads=0, time_chunk=0, kernel="auto"):
        k = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        check(self._lib.fpsftwa_set_tuning(self._h, block_threads, time_chunk, k), "set_tuning")

    def run(self):
        check(self._lib.fpsftwa_run(self._h), "run")

    def result(self, name: str) -> np.ndarray:
        which = OUT[name]
        n = ctypes.c_size_t(0)
        check(self._lib.fpsftwa_get_result(self._h, which, None, 0, ctypes.byref(n)), "get_result")
        out = np.zeros(n.value)
        check(self._lib.fpsftwa_get_result(self._h, which, _dptr(out), out.size, None), "get_result")
        T = self.T
        if name == "time":
            return out
        if name in ("basis", "sum_basis"):  # [n_bases][2][L][T]
            return out.reshape(-1, 2, self.L, T)
        return out.reshape(-1, T)

    def set_method(self, method: str):
        """"ftwa" (default), "hf" (time-dependent Hartree-Fock: one deterministic state, Fock exchange) or
        "gpsr" (Gaussian phase-space representation: stochastic trajectories of the normally ordered density
        matrix, diffusion noise from the run's seed)."""
        m = {"ftwa": METHOD_FTWA, "hf": METHOD_HF, "gpsr": METHOD_GPSR}[method.lower()]
        check(self._lib.fpsftwa_set_method(self._h, m), "set_method")

    def set_stream(self, stream: int | None):
        """Enqueue this engine's work on a HIP stream given by its handle (an int, e.g. a torch stream's
        cuda_stream); None returns to the engine's own stream."""
        check(self._lib.fpsftwa_set_stream(self._h, ctypes.c_void_p(stream) if stream else None), "set_stream")

    def set_recorded_element(self, i: int, j: int):
        """The element rho(i, j) behind result("N1hu") / result("N1hd") (default (0, L-1))."""
        check(self._lib.fpsftwa_set_recorded_element(self._h, int(i), int(j)), "set_recorded_element")

    def set_bases(self, bases):
        """Bases whose occupations diag(V^T rho V) are recorded: [n_bases][2][L][L] real (columns the
        basis states), or None to stop recording them."""
        if bases is None:
            check(self._lib.fpsftwa_set_bases(self._h, 0, None), "set_bases")
            return
        b = np.ascontiguousarray(bases, dtype=np.float64)
        if b.ndim != 4 or b.shape[1:] != (2, self.L, self.L):
            raise ValueError(f"bases must have shape (n, 2, {self.L}, {self.L})")
        check(self._lib.fpsftwa_set_bases(self._h, b.shape[0], _dptr(b)), "set_bases")

    @property
    def T(self) -> int:
        n = ctypes.c_size_t(0)
        check(self._lib.fpsftwa_get_result(self._h, 0, None, 0, ctypes.byref(n)), "get_result")
        return n.value

    def states(self, first=0, count=None) -> np.ndarray:
        """Final states, complex [count][2][L][L]."""
        if count is None:
            count = self.n_traj - first
        out = np.zeros((count, 2, self.L, self.L, 2))
        check(self._lib.fpsftwa_get_states(self._h, first, count, _dptr(out), out.size), "get_states")
        return out[..., 0] + 1j * out[..., 1]

    def autotune(self, trial_steps: int = 8) -> dict:
        """Time the batched kernels on a shortened run and keep the fastest; returns the choice."""
        k, b, c = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        r = ctypes.c_double(0)
        check(self._lib.fpsftwa_autotune(self._h, trial_steps, ctypes.byref(k), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(r)), "autotune")
        names = {v: n for n, v in KERNELS.items()}
        return {"kernel": names[k.value], "block_threads": b.value, "time_chunk": c.value, "traj_steps_per_s": r.value}

    def adjoint(self, checkpoint_every: int = 50) -> dict:
        """Discrete adjoint of the ensemble double occupancy at the last time point.

        Returns {"J", "grad_u" [L], "grad_t" [L, L], "grad_pulse" [n] (knots) / [1] (Gaussian A0) / None}.
        """
        L = self.L
        spec = _AdjointSpec(0, int(checkpoint_every))
        J = ctypes.c_double(0)
        gu = np.zeros(L)
        gt = np.zeros((L, L))
        npulse = self._n_pulse
        gp = np.zeros(max(npulse, 1))
        check(self._lib.fpsftwa_run_adjoint(self._h, ctypes.byref(spec), ctypes.byref(J), _dptr(gu), _dptr(gt),
                                            _dptr(gp) if npulse else None, npulse), "run_adjoint")
        return {"J": J.value, "grad_u": gu, "grad_t": gt, "grad_pulse": gp[:npulse] if npulse else None}

    def set_graph(self, on: bool) -> None:
        """fpsftwa_set_graph: launch a batched run's chunks as one HIP graph."""
        check(self._lib.fpsftwa_set_graph(self._h, 1 if on else 0), "set_graph")

    def run_noalloc(self, state_ptr: int, state_bytes: int, sums_ptr: int, n_sums: int) -> None:
        """fpsftwa_run_noalloc: the run into caller-owned device buffers (raw device addresses), with the
        library's fixed scratch instead of its own allocations."""
        check(self._lib.fpsftwa_run_noalloc(self._h, ctypes.c_void_p(state_ptr), state_bytes,
                                            ctypes.c_void_p(sums_ptr), n_sums), "run_noalloc")

    def timing(self) -> dict:
        n = ctypes.c_int64(0)
        k = ctypes.c_double(0)
        r = ctypes.c_double(0)
        check(self._lib.fpsftwa_get_timing(self._h, ctypes.byref(n), ctypes.byref(k), ctypes.byref(r)), "timing")
        return {"launches": n.value, "kernel_ms": k.value, "run_ms": r.value}
