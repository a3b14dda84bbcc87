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
    pulse: dict = field(default_factory=dict)  # keys of Engine(...): pulse_type, A0, t0, td, w0, phi, pol, knots, disp

    @property
    def L(self) -> int:
        return int(self.tM.shape[0])

    @classmethod
    def from_json_file(cls, path: str) -> "FHSystem":
        """Build from a simulation file's "system" object with the engine's JSON front end."""
        d = _lib.load_system_json(path)
        pot = d["pot"] if np.any(d["pot"]) else None
        return cls(tM=d["tM"], U=d["U"], V=d["V"], coulomb_type=d["coulomb_type"], pot=pot, n_up=d["n_up"],
                   n_dn=d["n_dn"])

    def getnSites(self) -> int:
        return self.L


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
        if res.get("tbStates", False) or res.get("mfStates", False):
            raise _lib.FtwaError(_lib.EUNSUPPORTED, "tbStates / mfStates are not supported by this build")
        self.flags = f

    def setResultFile(self, workName: str, folderName: str) -> None:
        self.workName, self.folderName = workName, folderName

    def setResultsWriter(self, how: str, yn: bool = True) -> None:
        if how == "json":
            self.writeJson = yn

    def setJsonFile(self, file: dict) -> None:
        self.jsonFile = dict(file)

    # ---- run ----
    def initiateSolver(self) -> None:
        s = self.system
        if self._engine is None:
            self._engine = _lib.Engine(s.tM, coulomb_type=s.coulomb_type, U=s.U, V=s.V, pot=s.pot, **s.pulse)
        L = s.L
        nu = np.zeros(L) if s.n_up is None else s.n_up
        nd = np.zeros(L) if s.n_dn is None else s.n_dn
        self._engine.set_initial_state(nu, nd)

    @property
    def engine_handle(self) -> _lib.Engine:
        if self._engine is None:
            self.initiateSolver()
        return self._engine

    def runSimulation(self) -> dict:
        e = self.engine_handle
        e.set_params(self.dt, self.tMax, self.nSub, self.nTraj, seed=self.seed, engine=self.engine)
        e.set_results(self.flags)
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
