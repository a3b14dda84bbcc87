"""Python mirror of the reference's FHSystem / FHfTWASolver interface over the C ABI.

Method names and argument meaning follow include/FHfTWASolver.h:28-41 and include/FHSolver.h:23-47;
parameters and results use the reference's JSON keys (src/FHfTWASolver.cpp:49-60, src/FHSolver.cpp:34-45,
:171-307).  The computation is the gfx950 engine's; this module only marshals arrays.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class FHSystem:
    """The matrices the fTWA path reads from FHSystem (src/FHSystem.cpp:387-391) plus occupations."""
    tM: np.ndarray
    U: np.ndarray | None = None
    V: np.ndarray | None = None
    coulomb_type: int = _lib.COULOMB_ONSITE
    pot: np.ndarray | None = None
    n_up: np.ndarray | None = None
    n_dn: np.ndarray | None = None
    rot_up: np.ndarray | None = None  # [L][L] initial-state rotations (tbGroundState / mfGroundState); None = identity
    rot_dn: np.ndarray | None = None
    core_hole_site: int | None = None  # lattice site linked to the core hole (the hole is site L-1)
    pulse: dict = field(default_factory=dict)  # keys of Engine(...): pulse_type, A0, t0, td, w0, phi, pol, knots, disp

    @property
    def L(self) -> int:
        return int(self.tM.shape[0])

    @classmethod
    def from_json_file(cls, path: str) -> "FHSystem":
        """Build from a simulation file's "system" object with the engine's JSON front end."""
        d = _lib.load_system_json(path)
        ic = _lib.load_initial_state_json(path)
        eye = np.eye(d["L"])
        rot = ic["rot_up"] if not (np.allclose(ic["rot_up"], eye) and np.allclose(ic["rot_dn"], eye)) else None
        pot = d["pot"] if np.any(d["pot"]) else None
        with open(path) as fh:
            ch = json.load(fh).get("system", {}).get("coreHole", {})
        site = int(ch.get("site", 0)) if ch.get("presence", False) else None
        return cls(tM=d["tM"], U=d["U"], V=d["V"], coulomb_type=d["coulomb_type"], pot=pot, n_up=d["n_up"],
                   n_dn=d["n_dn"], rot_up=rot, rot_dn=None if rot is None else ic["rot_dn"], core_hole_site=site)

    def getnSites(self) -> int:
        return self.L

    def computeTBGroundState(self) -> np.ndarray:
        """Eigenvectors of tM (src/FHSystem.cpp:909-937), columns in ascending eigenvalue order."""
        return _lib.ground_state(self.tM)["rot_up"]

    def computeMFGroundState(self, n_up: int | None = None, n_dn: int | None = None):
        """Self-consistent Hartree eigenvectors per spin (src/FHSystem.cpp:940-1019)."""
        g = _lib.ground_state(self.tM, self.U, n_up, n_dn, mean_field=True)
        return g["rot_up"], g["rot_dn"]

    def setGroundStateInitialConditions(self, kind: str, n_up: int | None = None, n_dn: int | None = None) -> None:
        """tbGroundState / mfGroundState initial state: the lowest n_up / n_dn eigenstates occupied and the
        sampled state rotated into that basis (src/FHSystem.cpp:536-595, src/FHfTWASolver.cpp:120-153)."""
        L = self.L
        n_up = int(np.ceil(L / 2)) if n_up is None else int(n_up)
        n_dn = int(np.floor(L / 2)) if n_dn is None else int(n_dn)
        if kind == "tbGroundState":
            ru = rd = self.computeTBGroundState()
        elif kind == "mfGroundState":
            ru, rd = self.computeMFGroundState(n_up, n_dn)
        else:
            raise ValueError(f"unknown eigenbasis initial condition {kind!r}")
        self.n_up = (np.arange(L) < n_up).astype(np.float64)
        self.n_dn = (np.arange(L) < n_dn).astype(np.float64)
        self.rot_up, self.rot_dn = ru.astype(np.complex128), rd.astype(np.complex128)


class FHfTWASolver:
    """Drives one engine handle the way fps.cpp:192-207 drives the reference solver."""

    def __init__(self, fh_system: FHSystem):
        self.system = fh_system
        self.tMax, self.dt, self.nSub = 1.0, 0.01, 1      # include/FHSolver.h:179-181
        self.nTraj = 1                                     # include/FHfTWASolver.h:52
        self.flags = _lib.RES_DIAG | _lib.RES_DIAG2 | _lib.RES_ENERGIES
        self.seed = 0
        self.engine = "fp64"
        self.workName, self.folderName = "", ""
        self.writeJson = False
        self.jsonFile: dict = {}
        self._engine: _lib.Engine | None = None
        self.results: dict = {}
        self.saveTB = self.saveMF = False   # "tbStates" / "mfStates" (src/FHSolver.cpp:41-42)
        self._basis_index: dict = {}

    # ---- setters (src/FHfTWASolver.cpp:20-69, src/FHSolver.cpp:11-112) ----
    def setIntegratorOrder(self, order: int) -> None:
        """Accepted and ignored: the reference always integrates with RK4 "3/8" (src/FHfTWASolver.cpp:90)."""

    def setfTWAParameters(self, nTraj: int) -> None:
        self.nTraj = int(nTraj)

    def setTimeParameters(self, t_max: float, dt: float, nPerTS: int = 1) -> None:
        self.tMax, self.dt, self.nSub = float(t_max), float(dt), int(nPerTS)

    def setMethodParameters(self, param: dict) -> None:
        self.setTimeParameters(param.get("t_max", 2.0), param.get("dt", 0.01), param.get("stepsParTimePoints", 1))
        self.setIntegratorOrder(param.get("int_order", 2))
        self.setfTWAParameters(param.get("n_traj", 10))
        m = param.get("mi355x", {})
        self.seed = int(m.get("seed", self.seed))
        self.engine = m.get("engine", self.engine)

    def setResultsToSave(self, res: dict) -> None:
        f = 0
        if res.get("diagDMatrix", True):
            f |= _lib.RES_DIAG
        if res.get("diag2DMatrix", False):
            f |= _lib.RES_DIAG2
        if res.get("onSiteDoublons", False):
            f |= _lib.RES_DOUBLONS
        if res.get("wholeDMatrix", False):
            f |= _lib.RES_WHOLE
        if res.get("energies", True):
            f |= _lib.RES_ENERGIES
        self.saveTB = bool(res.get("tbStates", False))
        self.saveMF = bool(res.get("mfStates", False))
        self.flags = f

    def setResultFile(self, workName: str, folderName: str) -> None:
        self.workName, self.folderName = workName, folderName

    def setResultsWriter(self, how: str, yn: bool = True) -> None:
        if how == "json":
            self.writeJson = yn

    def setJsonFile(self, file: dict) -> None:
        self.jsonFile = dict(file)

    # ---- run ----
    method = "ftwa"

    def initiateSolver(self) -> None:
        s = self.system
        if self._engine is None:
            self._engine = _lib.Engine(s.tM, coulomb_type=s.coulomb_type, U=s.U, V=s.V, pot=s.pot, **s.pulse)
        self._engine.set_method(self.method)
        L = s.L
        nu = np.zeros(L) if s.n_up is None else s.n_up
        nd = np.zeros(L) if s.n_dn is None else s.n_dn
        self._engine.set_initial_state(nu, nd, s.rot_up, s.rot_dn)
        # bases of the tbStates / mfStates observables (src/FHfTWASolver.cpp:80-87), for the system's
        # particle numbers (the occupied counts of the initial state)
        bases, self._basis_index = [], {}
        n_up, n_dn = int(round(float(np.sum(nu)))), int(round(float(np.sum(nd))))
        if self.saveTB:
            V = s.computeTBGroundState()
            self._basis_index["tb"] = len(bases)
            bases.append(np.stack([V, V]))
        if self.saveMF:
            Vu, Vd = s.computeMFGroundState(n_up, n_dn)
            self._basis_index["mf"] = len(bases)
            bases.append(np.stack([Vu, Vd]))
        self._engine.set_bases(np.stack(bases) if bases else None)
        if s.core_hole_site is not None:  # NiCh = <rho(site, L-1)> (src/FHfTWASolver.cpp:413-417)
            if (self.flags & _lib.RES_WHOLE) and s.core_hole_site != 0:
                raise _lib.FtwaError(_lib.EUNSUPPORTED, "wholeDMatrix with a core hole off site 0 is not supported")
            self._engine.set_recorded_element(s.core_hole_site, L - 1)

    @property
    def engine_handle(self) -> _lib.Engine:
        if self._engine is None:
            self.initiateSolver()
        return self._engine

    def runSimulation(self) -> dict:
        e = self.engine_handle
        e.set_params(self.dt, self.tMax, self.nSub, self.nTraj, seed=self.seed, engine=self.engine)
        e.set_results(self.flags)
        try:
            e.run()
        except _lib.FtwaError as err:
            # the batched GPSR kernel covers small lattices (L <= 32, short-range hopping, no pulse); the plain
            # FP64 engine runs the rest
            if self.method != "gpsr" or err.code != _lib.EUNSUPPORTED or self.engine == "fp64_plain":
                raise
            self.engine = "fp64_plain"
            e.set_params(self.dt, self.tMax, self.nSub, self.nTraj, seed=self.seed, engine=self.engine)
            e.run()
        sim = {"t": e.result("time").tolist()}
        f = self.flags
        if f & _lib.RES_WHOLE:
            for k in ("N1hu", "N1hd"):
                c = e.result(k)
                sim[k] = {"r": c[0].tolist(), "i": c[1].tolist()}
        if f & _lib.RES_DIAG:
            sim["Niiu"] = e.result("Niiu").tolist()
            sim["Niid"] = e.result("Niid").tolist()
        if f & _lib.RES_DIAG2:
            sim["N2iiu"] = e.result("N2iiu").tolist()
            sim["N2iid"] = e.result("N2iid").tolist()
        if f & _lib.RES_DOUBLONS:
            sim["NuNd"] = e.result("NuNd").tolist()
        if self.system.core_hole_site is not None:  # src/FHSolver.cpp:285-291
            sim["NiCh"] = {}
            for k, key in (("u", "N1hu"), ("d", "N1hd")):
                c = e.result(key)
                sim["NiCh"][k] = {"r": c[0].tolist(), "i": c[1].tolist()}
        if self._basis_index:  # src/FHSolver.cpp:242-265
            M = e.result("basis")
            for kind, (ku, kd) in (("tb", ("Mkku", "Mkkd")), ("mf", ("Mukku", "Mukkd"))):
                if kind in self._basis_index:
                    b = self._basis_index[kind]
                    sim[ku], sim[kd] = M[b, 0].tolist(), M[b, 1].tolist()
        if f & _lib.RES_ENERGIES:
            E = e.result("E")
            sim["E"] = {k: E[i].tolist() for i, k in enumerate(("tH", "U", "V", "pe", "tot"))}
        self.results = sim
        if self.writeJson:
            doc = dict(self.jsonFile)
            doc["fpsVersion"] = "0.5.0"
            doc["simulation"] = sim
            with open(os.path.join(self.folderName, self.workName + ".json"), "w") as fh:
                json.dump(doc, fh, indent="\t")
        return sim


class FHHFSolver(FHfTWASolver):
    """Time-dependent Hartree-Fock (include/FHHFSolver.h, src/FHHFSolver.cpp): one deterministic density
    matrix, Fock exchange with V (:575-608), on the plain FP64 engine; results file `<sim>_HF_res`."""

    method = "hf"

    def __init__(self, fh_system: FHSystem):
        super().__init__(fh_system)
        self.nTraj = 1
        self.engine = "fp64_plain"

    def setMethodParameters(self, param: dict) -> None:
        """dt, t_max, stepsParTimePoints, int_order with the defaults of src/FHHFSolver.cpp:47-57."""
        self.setTimeParameters(param.get("t_max", 2.0), param.get("dt", 0.01), param.get("stepsParTimePoints", 1))
        self.setIntegratorOrder(param.get("int_order", 2))


class FHGPSRSolver(FHfTWASolver):
    """Gaussian phase-space representation (include/FHGPSRSolver.h, src/FHGPSRSolver.cpp): stochastic
    trajectories of the normally ordered density matrix with the drift of derivative (:110-133) and the
    diffusion of step_Diffusion (:211-232); same method parameters and defaults as the fTWA solver
    (:48-58); results file `<sim>_GPSR_res`."""

    method = "gpsr"
