"""GPU parity of the gfx950 engine against the oracle, through the C ABI (include/fps_ftwa.h).

Bars (stated per test): FP64 engines match the oracle per trajectory and per observable to 1e-10
relative on identical samples; FP32 trajectories match to the tolerances written next to each check.
The oracle is used only as the checker, on the same seeded inputs; committed fixtures under
tests/golden/ hold oracle outputs so part of the suite needs no oracle library at all.
"""
import json
import os
import zlib

import numpy as np
import pytest

from fermiphasespace_amd import _lib
from fermiphasespace_amd import lattices as lat

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
FP64_RTOL = 1e-10


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def chain(L, t=-1.0, periodic=False):
    tM = np.zeros((L, L))
    for i in range(L - 1):
        tM[i, i + 1] = tM[i + 1, i] = t
    if periodic and L > 2:
        tM[0, L - 1] = tM[L - 1, 0] = t
    return tM


def engine_run(eng, dt, t_max, n_sub, B, engine, seed=0, noise=None, bases=False):
    eng.set_params(dt, t_max, n_sub, B, seed=seed, engine=engine)
    eng.set_noise(noise)
    eng.run()
    out = {k: eng.result(k) for k in ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e")}
    out["states"] = eng.states()
    out["corner"] = np.vstack([eng.result("N1hu"), eng.result("N1hd")]) * B  # [re u, im u, re d, im d][T]
    if bases:
        out["sum_basis"] = eng.result("sum_basis")
    return out


def oracle_system(O, tM, coulomb_type, U=None, V=None, pot=None, pulse=None):
    pulse = pulse or {}
    return O.System(tM=tM, coulomb_type=coulomb_type, U=U, V=V, pot=pot,
                    pulse_type=pulse.get("pulse_type", 0), A0=pulse.get("A0", 0.0), t0=pulse.get("t0", 0.0),
                    td=pulse.get("td", 1.0), w0=pulse.get("w0", 0.0), phi=pulse.get("phi", 0.0),
                    pol=pulse.get("pol", (0.0, 0.0, 0.0)), knots=pulse.get("knots"), disp=pulse.get("disp"))


def compare(O, tM, n_up, n_dn, *, coulomb_type=1, U=None, V=None, pot=None, pulse=None, dt=0.02, t_max=0.2,
            n_sub=1, B=6, seed=3, engines=("fp64_plain", "fp64", "fp32"), fp32_tol=2e-5, rot=None, bases=None,
            corner=None):
    L = tM.shape[0]
    eng = _lib.Engine(tM, coulomb_type=coulomb_type, U=U, V=V, pot=pot, **(pulse or {}))
    if bases is not None:
        eng.set_bases(bases)
    if corner is not None:
        eng.set_recorded_element(*corner)
    if rot is None:
        eng.set_initial_state(n_up, n_dn)
    else:
        eng.set_initial_state(n_up, n_dn, rot[0], rot[1])
    eng.set_params(dt, t_max, n_sub, B, seed=seed)
    noise = eng.sample_noise(0, B)
    ref = O.run(oracle_system(O, tM, coulomb_type, U, V, pot, pulse), dt, t_max, n_sub, n_up, n_dn,
                noise_arr=noise, keep_states=True, rot_up=None if rot is None else rot[0],
                rot_dn=None if rot is None else rot[1], bases=bases, corner=corner)
    runs = []
    z = int(((tM != 0) & ~np.eye(L, dtype=bool)).sum(axis=1).max())
    stencil = L <= 64 and z <= 8  # where the LDS stencil kernels apply
    for e in engines:
        if e == "fp32":  # every FP32 kernel: full-matrix and upper-triangle (where they fit), the dense path
            # with the library GEMM and with its fused matrix-core kernel (where L allows)
            runs += ([("fp32", "full"), ("fp32", "hermitian")] if stencil else []) + [("fp32", "dense"),
                                                                                     ("fp32", "dense_mfma")]
            if stencil and (pulse or {}).get("pulse_type", 0) == 0 and coulomb_type != 2:
                runs.append(("fp32", "pipelined"))
                if L <= 64 and bases is None:
                    runs.append(("fp32", "sequential"))
            if (64 < L <= 107 and z <= 8 and (pulse or {}).get("pulse_type", 0) == 0 and coulomb_type != 2
                    and bases is None):  # the column-blocked spin-sequential kernel
                runs += [("fp32", "sequential"), ("fp32", "auto")]
        elif e == "fp64":  # upper-triangle (16 and 8 waves) and stencil kernels where they fit, the dense path
            herm_seq = (stencil and L <= 64 and (pulse or {}).get("pulse_type", 0) == 0 and coulomb_type != 2
                        and bases is None)
            runs += ([("fp64", "auto"), ("fp64", "full")] if stencil else []) + [("fp64", "dense")]
            if herm_seq:
                runs += [("fp64", "hermitian"), ("fp64", "hermitian/1024")]
        else:
            runs.append((e, "auto"))
    for e, kern in runs:
        kname, _, blk = kern.partition("/")
        eng.set_tuning(int(blk or 0), 0, kname)
        got = engine_run(eng, dt, t_max, n_sub, B, e, seed=seed, noise=noise, bases=bases is not None)
        tol = fp32_tol if e == "fp32" else FP64_RTOL
        print(f"L={L} {e}: states {rel(got['states'], ref.final_states):.2e}")
        assert rel(got["states"], ref.final_states) < tol, e
        assert rel(got["sum_u"], ref.sum_u) < tol, e
        assert rel(got["sum_d"], ref.sum_d) < tol, e
        assert rel(got["sum2_u"], ref.sum2_u) < tol, e
        assert rel(got["sum_ud"], ref.sum_ud) < tol, e
        assert rel(got["sum_e"], ref.sum_e) < tol, e
        if bases is not None:
            assert rel(got["sum_basis"], ref.sum_basis) < tol, (e, kern)
        ci, cj = (0, L - 1) if corner is None else corner
        if ci != cj:  # the recorded off-diagonal element (wholeDMatrix corner / core-hole NiCh)
            assert np.abs(got["corner"] - ref.sum_corner).max() <= tol * max(np.abs(ref.sum_corner).max(), 1e-3), (e, kern)
    eng.close()


def test_device_noise_matches_oracle(oracle_mod):
    L, B = 9, 40
    eng = _lib.Engine(np.zeros((L, L)), U=np.eye(L))
    eng.set_params(0.1, 0.1, 1, B, seed=0xDEADBEEF12345)
    w = eng.sample_noise(0, B)
    wo = oracle_mod.noise(L, 0, B, 0xDEADBEEF12345)
    assert rel(w, wo) < 1e-14  # same Philox integers; log/sincos differ by at most an ulp or two
    np.testing.assert_array_equal(eng.sample_noise(7, 3), w[7:10])  # counter-based slices


def test_golden_fixtures():
    for name in ("chain6_onsite", "ring6_general"):
        f = np.load(os.path.join(GOLDEN, name + ".npz"))
        ct = int(f["coulomb_type"])
        eng = _lib.Engine(f["tM"], coulomb_type=ct, U=f["U"], V=f["V"] if ct == 2 else None,
                          pot=f["pot"] if np.any(f["pot"]) else None)
        eng.set_initial_state(f["n_up"], f["n_dn"])
        B = int(f["B"])
        for e in ("fp64_plain", "fp64"):
            got = engine_run(eng, float(f["dt"]), float(f["t_max"]), int(f["n_sub"]), B, e, noise=f["noise"])
            assert rel(got["states"], f["final_states"]) < FP64_RTOL, (name, e)
            for k in ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e"):
                assert rel(got[k], f[k]) < FP64_RTOL, (name, e, k)
        eng.close()


def test_chain10_workload_shape(oracle_mod):
    """The 10-site chain workload (t=-1, U=4, Neel, dt=0.01, 10 sub-steps), shortened to t=0.3."""
    L = 10
    n_up = np.array([1.0, 0] * 5)
    compare(oracle_mod, chain(L), n_up, 1 - n_up, U=4.0 * np.eye(L), dt=0.01, t_max=0.3, n_sub=10, B=12)


def test_hubbard8x8_shape(oracle_mod):
    pos, box = lat.grid(8, 8)
    tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=5.3)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    compare(oracle_mod, tM, occ, occ, U=Um, dt=0.005, t_max=0.05, B=8, seed=202)


@pytest.mark.parametrize("L", [2, 5, 24, 36, 48, 54])
def test_site_counts(oracle_mod, L):
    rng = np.random.default_rng(L)
    n_up = (rng.uniform(size=L) < 0.5).astype(float)
    n_dn = (rng.uniform(size=L) < 0.5).astype(float)
    compare(oracle_mod, chain(L, periodic=True), n_up, n_dn, U=3.0 * np.eye(L), dt=0.02, t_max=0.1, B=3,
            engines=("fp64", "fp32"))


@pytest.mark.parametrize("shape", ["chain80", "grid10x10", "chain90_reach4", "chain107", "grid10x10_bonds",
                                   "grid10x10_diag"])
def test_wide_lattices(oracle_mod, shape):
    """64 < L <= 107: the FP32 spin-sequential kernel for L > 64 (automatic and forced) and the dense path
    against the oracle, with site potentials; z = 4 and 8, one hopping value on every neighbour entry
    (grid10x10: z = 4; grid10x10_diag, nearest and diagonal neighbours alike: z = 8) or several (chains,
    whose padded entries are zero; grid10x10_bonds: a different hopping per bond), L at the kernel's LDS
    limit."""
    rng = np.random.default_rng(zlib.crc32(shape.encode()) % 1000)  # stable across processes
    if shape.startswith("grid10x10"):
        pos, box = lat.grid(10, 10)
        tM, _ = lat.system_arrays(pos, box, True)
        if shape == "grid10x10_bonds":
            f = rng.uniform(0.5, 1.5, tM.shape)
            tM = tM * np.triu(f, 1)
            tM = tM + tM.T
        elif shape == "grid10x10_diag":
            tM = tM + lat.system_arrays(pos, box, True, nn=2 ** 0.5)[0]
    elif shape == "chain90_reach4":
        L = 90
        tM = np.zeros((L, L))
        for d in range(1, 5):
            for i in range(L - d):
                tM[i, i + d] = tM[i + d, i] = -1.0 / d
    else:
        tM = chain(int(shape[5:]), periodic=True)
    L = tM.shape[0]
    n_up = (rng.uniform(size=L) < 0.5).astype(float)
    n_dn = (rng.uniform(size=L) < 0.5).astype(float)
    compare(oracle_mod, tM, n_up, n_dn, U=4.0 * np.eye(L), pot=rng.uniform(-1, 1, L), dt=0.02, t_max=0.1, B=3,
            engines=("fp64", "fp32"))


@pytest.mark.parametrize("engine", ["fp32", "fp64"])
@pytest.mark.parametrize("chunk", [0, 3])
def test_in_kernel_sampling_matches_sampling_kernel(engine, chunk):
    """The spin-sequential kernel samples the initial states in its first launch.  The same run fed the
    generator's own noise through the separate sampling kernel (fpsftwa_set_noise) gives bitwise the same
    final states and ensemble sums: 8x8 periodic lattice, 512 trajectories, whole-run and 3-time-point
    launches."""
    pos, box = lat.grid(8, 8)
    tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=4.4)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    occ[5], occ[6] = 0.3, 0.8  # fractional occupations: S_ij takes every branch of its formula
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(occ, 1.0 - occ)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES | _lib.RES_WHOLE)
    eng.set_tuning(0, chunk, "sequential")
    out = []
    for explicit in (False, True):
        eng.set_params(0.01, 0.1, 1, 512, seed=31, engine=engine)
        eng.set_noise(eng.sample_noise(0, 512) if explicit else None)
        eng.run()
        out.append((eng.states(), [eng.result(k) for k in ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e", "N1hu")]))
    eng.close()
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("engine", ["fp32", "fp64"])
@pytest.mark.parametrize("shape,B", [("honeycomb24", 6), ("honeycomb24", 7), ("grid4x4", 9), ("chain5", 8),
                                     ("chain32", 5)])
def test_packed_trajectories_are_bitwise_neutral(oracle_mod, engine, shape, B):
    """L <= 32: the spin-sequential kernel holds 2 trajectories per workgroup (L <= 16: 4), the last
    workgroup ragged where that does not divide B.  Against one trajectory per workgroup (block hint 512)
    it gives bitwise the same final states, ensemble sums, energies and recorded element, with the initial
    states sampled in the kernel and from explicit noise, in whole-run and 3-time-point launches; and it
    matches the oracle on the explicit samples (FP64 1e-10, FP32 2e-5).  Honeycomb: z = 3, padded to 4."""
    if shape == "honeycomb24":
        pos, box = lat.honeycomb(4, 3)
        tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=4.0)
    elif shape == "grid4x4":
        pos, box = lat.grid(4, 4)
        tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=4.0)
    else:
        tM = chain(int(shape[5:]), periodic=True)
        Um = 4.0 * np.eye(tM.shape[0])
    L = tM.shape[0]
    rng = np.random.default_rng(zlib.crc32(shape.encode()))
    n_up = (rng.uniform(size=L) < 0.5).astype(float)
    n_up[1] = 0.3
    n_dn = 1.0 - n_up
    pot = rng.uniform(-1, 1, L)
    eng = _lib.Engine(tM, U=Um, pot=pot)
    eng.set_initial_state(n_up, n_dn)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES | _lib.RES_WHOLE)
    keys = ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e", "N1hu", "N1hd")
    for explicit in (False, True):
        for chunk in (0, 3):
            out = []
            for block in (0, 512):
                eng.set_tuning(block, chunk, "sequential")
                eng.set_params(0.02, 0.12, 1, B, seed=5, engine=engine)
                eng.set_noise(eng.sample_noise(0, B) if explicit else None)
                eng.run()
                out.append((eng.states(), [eng.result(k) for k in keys]))
            assert np.array_equal(out[0][0], out[1][0]), (explicit, chunk)
            for k, a, b in zip(keys, out[0][1], out[1][1]):
                assert np.array_equal(a, b), (explicit, chunk, k)
    eng.close()
    compare(oracle_mod, tM, n_up, n_dn, U=Um, pot=pot, dt=0.02, t_max=0.12, B=B, engines=(engine,))


def test_graph_launch_is_bitwise_neutral():
    """fpsftwa_set_graph: the run's integrator and reduction launches (here 7 time chunks of 3 points)
    replayed as one HIP graph give bitwise the results of direct launches, alone and in fpsftwa_run_many
    with a system that the graph does not cover (dense path)."""
    pos, box = lat.grid(8, 8)
    tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=4.4)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    keys = ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e", "N1hu")
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(occ, occ)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES | _lib.RES_WHOLE)
    eng.set_tuning(0, 3, "auto")
    eng.set_params(0.01, 0.2, 1, 1024, seed=17, engine="fp32")
    out = []
    for graph in (False, True, False):
        eng.set_graph(graph)
        eng.run()
        out.append((eng.states(), [eng.result(k) for k in keys]))
        assert eng.timing()["launches"] == 7
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0])
        for a, b in zip(o[1], out[0][1]):
            assert np.array_equal(a, b)
    other = _lib.Engine(chain(80), U=np.eye(80))  # L = 80, FP64: the dense path, launched directly
    other.set_initial_state(np.array([1.0, 0] * 40), np.array([0.0, 1] * 40))
    other.set_params(0.02, 0.06, 1, 8, seed=3, engine="fp64")
    other.run()
    ref_other = other.result("sum_u")
    eng.set_graph(True)
    other.set_graph(True)
    _lib.run_many([eng, other])
    assert np.array_equal(eng.states(), out[0][0])
    assert np.array_equal(other.result("sum_u"), ref_other)
    eng.close()
    other.close()


def test_wide_kernel_is_chunking_invariant():
    """The column-blocked kernel's result does not depend on the time chunking (bitwise)."""
    pos, box = lat.grid(10, 10)
    tM, Um = lat.system_arrays(pos, box, True)
    L = tM.shape[0]
    nu = np.array(lat.checkerboard(10, 10)["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(nu, nu)
    out = []
    for chunk in (0, 3):
        eng.set_tuning(0, chunk, "sequential")
        eng.set_params(0.01, 0.2, 1, 512, seed=8, engine="fp32")
        eng.run()
        out.append((eng.states(), eng.result("sum_u"), eng.result("sum_e")))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    eng.close()


def test_general_coulomb_and_potential(oracle_mod):
    L = 12
    rng = np.random.default_rng(5)
    U = rng.uniform(0.2, 1.0, (L, L))
    U = (U + U.T) / 2
    V = rng.uniform(0.0, 0.5, (L, L))
    V = (V + V.T) / 2
    np.fill_diagonal(V, 0)
    pot = rng.uniform(-1, 1, L)
    n_up = np.array([1.0, 0] * 6)
    compare(oracle_mod, chain(L, periodic=True), n_up, n_up[::-1].copy(), coulomb_type=2, U=U, V=V, pot=pot, B=5)


def test_null_coulomb_and_fractional_occupations(oracle_mod):
    L = 7
    n = np.linspace(0.1, 0.9, L)
    compare(oracle_mod, chain(L), n, n[::-1].copy(), coulomb_type=0, B=4)


@pytest.mark.parametrize("kind", ["gaussian", "knots"])
def test_pulse(oracle_mod, kind):
    L = 8
    pos, box = lat.chain(L)
    tM = chain(L, periodic=True)
    tM[2, 2] = 0.25  # a diagonal hopping also enters the pulse potential (t_ii e^{i A.0})
    d = pos[:, None, :] - pos[None, :, :]
    d[0, L - 1, 0], d[L - 1, 0, 0] = 1.0, -1.0  # minimum image across the ring
    if kind == "gaussian":
        pulse = dict(pulse_type=1, A0=0.5, t0=0.1, td=0.2, w0=1.5, phi=0.3, pol=(0.6, 0.8, 0.0), disp=d)
    else:
        pulse = dict(pulse_type=2, knots=np.array([0.0, 0.8, -0.4, 0.2]), pol=(1.0, 0.0, 0.0), disp=d)
    n_up = np.array([1.0, 0] * 4)
    compare(oracle_mod, tM, n_up, 1 - n_up, U=2.0 * np.eye(L), pulse=pulse, dt=0.02, t_max=0.3, n_sub=2, B=4)


def test_rotated_initial_state(oracle_mod):
    L = 6
    rng = np.random.default_rng(1)
    Q = []
    for _ in range(2):
        q, _ = np.linalg.qr(rng.normal(size=(L, L)) + 1j * rng.normal(size=(L, L)))
        Q.append(q)
    n_up = np.array([1.0, 1, 1, 0, 0, 0])
    compare(oracle_mod, chain(L), n_up, n_up, U=4.0 * np.eye(L), B=4, rot=Q, engines=("fp64_plain", "fp64"))


@pytest.mark.parametrize("kind", ["tbGroundState", "mfGroundState"])
def test_eigenbasis_initial_state(oracle_mod, kind):
    """tbGroundState / mfGroundState samples rotated into the eigenbasis (src/FHfTWASolver.cpp:120-153): every
    engine and kernel against the oracle fed LAPACK's eigenvectors in the same sign gauge, to 1e-10 (FP64)."""
    O = oracle_mod
    L = 8
    tM = chain(L) + np.diag(np.linspace(-0.2, 0.3, L))
    U = 3.0 * np.eye(L)
    Vu, Vd, _, _ = O.ground_state(tM, U, 5, 3, kind == "mfGroundState")
    n_up = (np.arange(L) < 5).astype(float)
    n_dn = (np.arange(L) < 3).astype(float)
    compare(O, tM, n_up, n_dn, U=U, B=8, rot=(Vu, Vd), t_max=0.3)


@pytest.mark.parametrize("kind", ["tbGroundState", "mfGroundState"])
def test_eigenbasis_initial_state_json(tmp_path, oracle_mod, kind):
    """The JSON front end builds the eigenbasis itself: its results file matches the oracle run on LAPACK's
    basis (same gauge) and the generator's noise, to 1e-10 relative."""
    O = oracle_mod
    pos, _ = lat.chain(8)
    B, dt, t_max, seed = 16, 0.02, 0.4, 7
    sim = lat.simulation(kind, pos, None, False, U=3.0, ic={"type": kind, "nPartUp": 5, "nPartDown": 3}, dt=dt,
                         t_max=t_max, n_traj=B, mi355x={"seed": seed, "engine": "fp64"})
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    doc = json.load(open(tmp_path / f"{kind}_fTWA_res.json"))
    out = doc["simulation"]
    d = _lib.load_system_json(path)
    Vu, Vd, it, res, Eu, Ed = O.ground_state(d["tM"], d["U"], 5, 3, kind == "mfGroundState", energies=True)
    check_ground_state_blocks(doc, kind, Vu, Vd, it, res, Eu, Ed)
    eng = _lib.Engine(d["tM"], U=d["U"])
    eng.set_initial_state(d["n_up"], d["n_dn"], Vu, Vd)
    eng.set_params(dt, t_max, 1, B, seed=seed, engine="fp64")
    noise = eng.sample_noise(0, B)
    eng.close()
    ref = O.normalise(O.run(oracle_system(O, d["tM"], 1, d["U"]), dt, t_max, 1, d["n_up"], d["n_dn"],
                            noise_arr=noise, rot_up=Vu, rot_dn=Vd), B)
    for k in ("Niiu", "Niid", "NuNd"):
        assert rel(np.asarray(out[k]), ref[k]) < FP64_RTOL, k
    assert rel(np.asarray(out["E"]["tot"]), ref["E"][4]) < FP64_RTOL


def check_ground_state_blocks(doc, kind, Vu, Vd, it, res, Eu, Ed):
    """The results file's eigenbasis echo (src/FHSolver.cpp:177-185, src/FHSystem.cpp:1143-1230) against
    LAPACK's: energies to 1e-12, rotation matrices (rows of V, columns the eigenvectors) in the build's
    sign gauge to 1e-10, the mean-field iteration count exactly and its residue to 1e-9."""
    if kind == "tbGroundState":
        assert "mfGroundState" not in doc
        tb = doc["tbGroundState"]
        np.testing.assert_allclose(tb["energyVector"], Eu, atol=1e-12)
        np.testing.assert_allclose(np.array(tb["rotationMatrix"]), Vu, atol=1e-10)
        return
    assert "tbGroundState" not in doc
    mf = doc["mfGroundState"]
    np.testing.assert_allclose(mf["energyVectorUp"], Eu, atol=1e-12)
    np.testing.assert_allclose(mf["energyVectorDown"], Ed, atol=1e-12)
    np.testing.assert_allclose(np.array(mf["rotationMatrixUp"]), Vu, atol=1e-10)
    np.testing.assert_allclose(np.array(mf["rotationMatrixDown"]), Vd, atol=1e-10)
    assert mf["iterationsNeeded"] == it
    assert abs(mf["residueComputation"] - res) < 1e-9


def test_basis_occupations(oracle_mod):
    """tbStates / mfStates occupations diag(V^T rho V) (src/FHfTWASolver.cpp:385-394) on every engine and
    kernel against the oracle's per-trajectory products, to 1e-10 (FP64) / 2e-5 (FP32): two real bases,
    different per spin, on the stencil-sized chain, the dense path and a pulse."""
    O = oracle_mod
    L = 8
    tM = chain(L) + np.diag(np.linspace(-0.2, 0.3, L))
    U = 3.0 * np.eye(L)
    n_up = (np.arange(L) < 5).astype(float)
    n_dn = (np.arange(L) < 3).astype(float)
    Vt, _, _, _ = O.ground_state(tM, None, 5, 3, False)
    Vu, Vd, _, _ = O.ground_state(tM, U, 5, 3, True)
    bases = np.stack([np.stack([Vt, Vt]), np.stack([Vu, Vd])])
    compare(O, tM, n_up, n_dn, U=U, B=8, rot=(Vu, Vd), t_max=0.3, bases=bases)
    rng = np.random.default_rng(3)
    Q = np.stack([np.stack([np.linalg.qr(rng.normal(size=(L, L)))[0] for _ in range(2)])])
    pos, _ = lat.chain(L)
    disp = (pos[:, None, :] - pos[None, :, :])
    pulse = {"pulse_type": _lib.PULSE_GAUSSIAN, "A0": 0.7, "t0": 0.1, "td": 0.2, "w0": 3.0, "phi": 0.2,
             "pol": (1.0, 0.0, 0.0), "disp": disp}
    compare(O, tM, n_up, 1 - n_up, U=U, B=5, t_max=0.2, pulse=pulse, bases=Q)


def test_basis_occupations_trace_at_full_size():
    """8x8 periodic Hubbard, FP32 upper-triangle kernel with the density records: the occupations of any
    orthonormal basis sum to the trace, so sum_k M_k(t) = sum_i N_ii(t) for both spins at every time point
    (size-independent; 1e-9 relative, both sides from the same FP32 states summed in FP64)."""
    pos, box = lat.grid(8, 8)
    tM, U = lat.system_arrays(pos, box, True, t=-1.0, U=5.3)
    n_up = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    n_dn = n_up
    V = np.linalg.eigh(tM)[1]
    eng = _lib.Engine(tM, U=U)
    eng.set_initial_state(n_up, n_dn)
    eng.set_bases(np.stack([np.stack([V, V])]))
    eng.set_tuning(0, 0, "hermitian")
    eng.set_params(0.005, 0.1, 1, 4096, seed=9, engine="fp32")
    eng.run()
    M = eng.result("sum_basis")[0]
    for sg, k in ((0, "sum_u"), (1, "sum_d")):
        tr = eng.result(k).sum(axis=0)
        assert np.abs(M[sg].sum(axis=0) - tr).max() <= 1e-9 * max(1.0, np.abs(tr).max()) + 1e-9 * 4096, sg
    eng.set_bases(None)
    eng.run()
    with pytest.raises(_lib.FtwaError):
        eng.result("basis")
    eng.close()


def test_basis_occupations_json(tmp_path, oracle_mod):
    """"tbStates" / "mfStates" in a simulation file: Mkku/Mkkd/Mukku/Mukkd in the results file match the
    oracle fed LAPACK's bases in the same gauge (1e-10 relative); the mean-field basis is built for the
    IC's particle numbers (nPartUp / nPartDown of the spin wave)."""
    O = oracle_mod
    pos, _ = lat.chain(8)
    B, dt, t_max, seed = 12, 0.02, 0.3, 4
    sim = lat.simulation("states", pos, None, False, U=2.5, ic={"type": "givenDiagonal", "subtype": "spinWave",
                                                                 "nPartUp": 4, "nPartDown": 3},
                         dt=dt, t_max=t_max, n_traj=B, mi355x={"seed": seed, "engine": "fp64"},
                         results={"diagDMatrix": True, "tbStates": True, "mfStates": True})
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    doc = json.load(open(tmp_path / "states_fTWA_res.json"))
    out = doc["simulation"]
    d = _lib.load_system_json(path)
    Vt, _, _, _, Et, _ = O.ground_state(d["tM"], None, 4, 3, False, energies=True)
    Vu, Vd, it, res, Eu, Ed = O.ground_state(d["tM"], d["U"], 4, 3, True, energies=True)
    # the bases computed for tbStates / mfStates are echoed as the reference's eigenbasis blocks
    check_ground_state_blocks({"tbGroundState": doc["tbGroundState"]}, "tbGroundState", Vt, Vt, 0, 0.0, Et, Et)
    check_ground_state_blocks({"mfGroundState": doc["mfGroundState"]}, "mfGroundState", Vu, Vd, it, res, Eu, Ed)
    eng = _lib.Engine(d["tM"], U=d["U"])
    eng.set_initial_state(d["n_up"], d["n_dn"])
    eng.set_params(dt, t_max, 1, B, seed=seed, engine="fp64")
    noise = eng.sample_noise(0, B)
    eng.close()
    ref = O.normalise(O.run(oracle_system(O, d["tM"], 1, d["U"]), dt, t_max, 1, d["n_up"], d["n_dn"],
                            noise_arr=noise, bases=np.stack([np.stack([Vt, Vt]), np.stack([Vu, Vd])])), B)
    for key, b, sg in (("Mkku", 0, 0), ("Mkkd", 0, 1), ("Mukku", 1, 0), ("Mukkd", 1, 1)):
        assert rel(np.asarray(out[key]), ref["M"][b, sg]) < FP64_RTOL, key
    assert rel(np.asarray(out["Niiu"]), ref["Niiu"]) < FP64_RTOL


@pytest.mark.parametrize("engine", ["fp64", "fp32"])
def test_chain10_against_exact_diagonalisation(engine):
    """Physics check at the 10-site chain (t = 1, Neel state) with 65,536 trajectories: fTWA is exact for
    free fermions, so at U = 0 the ensemble densities follow ED within 5 sampling errors up to t = 3; at
    U = 4 fTWA approximates the interacting dynamics and follows ED within 5 sampling errors up to
    t = 0.35 (the systematic departure reaches that bound near t = 0.45)."""
    from oracle import ed
    L, B, dt = 10, 65536, 0.01
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1))
    nu = np.array([1.0, 0] * 5)
    nd = 1 - nu
    for U, t_ok in ((0.0, 3.0), (4.0, 0.35)):
        nt = int(round(t_ok / 0.05)) + 1
        r = ed.evolve_product_state(L, tM, U * np.ones(L), nu, nd, t_ok, nt)
        eng = _lib.Engine(tM, U=U * np.eye(L))
        eng.set_initial_state(nu, nd)
        eng.set_params(dt, t_ok, 1, B, seed=11, engine=engine)
        eng.run()
        idx = np.round(r["t"] / dt).astype(int)
        for spin, key in (("n_up", "Niiu"), ("n_dn", "Niid")):
            got = eng.result(key)[:, idx]
            err = eng.result(key + "_err")[:, idx]
            dev = np.abs(got - r[spin])
            assert np.all(dev <= 5 * err + 1e-9), (U, spin, float(dev.max()), float(err.max()))
        eng.close()


def test_recorded_element(oracle_mod):
    """The per-time-point off-diagonal element (wholeDMatrix corner, core-hole NiCh) at a chosen (i, j), on
    every engine and kernel, against the oracle's sums of the same element."""
    L = 9
    tM = chain(L) + np.diag(np.linspace(-0.2, 0.3, L))
    n_up = (np.arange(L) % 2 == 0).astype(float)
    compare(oracle_mod, tM, n_up, 1 - n_up, U=2.5 * np.eye(L), B=6, t_max=0.3, corner=(2, 5))
    compare(oracle_mod, tM, n_up, 1 - n_up, U=2.5 * np.eye(L), B=6, t_max=0.3, corner=(7, 1))


def test_core_hole_json(tmp_path, oracle_mod):
    """A core-hole system from a simulation file (src/FHSystem.cpp:295-323): the hole is site L, coupled to
    its site by X, with potential eps and interactions G (site) and Uc (hole); the results file's NiCh is
    <rho_sigma(site, L)> (src/FHfTWASolver.cpp:413-417, src/FHSolver.cpp:285-291).  Against the oracle on
    the front end's own matrices and the generator's noise, 1e-10 relative (FP64)."""
    O = oracle_mod
    pos, _ = lat.chain(6)
    B, dt, t_max, seed = 10, 0.02, 0.4, 8
    sim = lat.simulation("ch", pos, None, False, U=3.0, dt=dt, t_max=t_max, n_traj=B,
                         mi355x={"seed": seed, "engine": "fp64"},
                         results={"diagDMatrix": True, "onSiteDoublons": True, "energies": True})
    sim["system"]["coreHole"] = {"presence": True, "site": 2, "eps": -2.5, "X": -0.7, "Uc": 1.5, "G": 0.8,
                                 "IC": [1, 0]}
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    out = json.load(open(tmp_path / "ch_fTWA_res.json"))["simulation"]
    d = _lib.load_system_json(path)
    L = d["L"]
    assert L == 7 and d["coulomb_type"] == _lib.COULOMB_GENERAL
    eng = _lib.Engine(d["tM"], coulomb_type=_lib.COULOMB_GENERAL, U=d["U"], V=d["V"], pot=d["pot"])
    eng.set_initial_state(d["n_up"], d["n_dn"])
    eng.set_params(dt, t_max, 1, B, seed=seed, engine="fp64")
    noise = eng.sample_noise(0, B)
    eng.close()
    s = O.run(oracle_system(O, d["tM"], 2, d["U"], d["V"], d["pot"]), dt, t_max, 1, d["n_up"], d["n_dn"],
              noise_arr=noise, corner=(2, L - 1))
    ref = O.normalise(s, B)
    for k in ("Niiu", "Niid", "NuNd"):
        assert rel(np.asarray(out[k]), ref[k]) < FP64_RTOL, k
    assert rel(np.asarray(out["E"]["tot"]), ref["E"][4]) < FP64_RTOL
    got = np.array([out["NiCh"]["u"]["r"], out["NiCh"]["u"]["i"], out["NiCh"]["d"]["r"], out["NiCh"]["d"]["i"]])
    assert np.abs(got - s.sum_corner / B).max() <= 1e-10 * np.abs(s.sum_corner / B).max()
    assert np.abs(np.asarray(out["NiCh"]["u"]["r"])[-1]) > 1e-4  # the hole is coupled: the element moves


def _hf_system(L=6):
    tM = chain(L) + np.diag(np.linspace(-0.2, 0.3, L))
    dist = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :]).astype(float)
    V = np.where(dist > 0, 1.2 / (1.0 + dist), 0.0)
    return tM, 2.5 * np.eye(L) + V, V


def test_hartree_fock_engine(oracle_mod):
    """Time-dependent Hartree-Fock (src/FHHFSolver.cpp:575-608): with a V interaction the Fock exchange
    runs on the plain FP64 engine and matches the oracle (fock, zero noise) state and sums to 1e-10; with
    an on-site U (no exchange) every engine runs HF and matches too; a batched engine refuses the Fock term."""
    O = oracle_mod
    L, B, dt, t_max = 6, 1, 0.02, 0.4
    tM, U, V = _hf_system(L)
    n_up = np.array([1.0, 0, 1, 1, 0, 0])
    n_dn = np.array([0.0, 1, 1, 0, 1, 0])
    zero = np.zeros((B, 2, L * (L - 1) // 2, 2))
    for coulomb, fock_ok_engines in ((2, ("fp64_plain",)), (1, ("fp64_plain", "fp64", "fp32"))):
        Um = U if coulomb == 2 else 2.5 * np.eye(L)
        eng = _lib.Engine(tM, coulomb_type=coulomb, U=Um, V=V if coulomb == 2 else None)
        eng.set_method("hf")
        eng.set_initial_state(n_up, n_dn)
        ref = O.run(O.System(tM=tM, coulomb_type=coulomb, U=Um, V=V if coulomb == 2 else None, fock=True), dt, t_max,
                    1, n_up, n_dn, noise_arr=zero, keep_states=True)
        for engine in fock_ok_engines:
            got = engine_run(eng, dt, t_max, 1, B, engine, seed=5)
            tol = 2e-5 if engine == "fp32" else FP64_RTOL
            assert rel(got["states"], ref.final_states) < tol, (coulomb, engine)
            assert rel(got["sum_u"], ref.sum_u) < tol and rel(got["sum_e"], ref.sum_e) < tol, (coulomb, engine)
        if coulomb == 2:
            eng.set_params(dt, t_max, 1, B, seed=5, engine="fp64")
            with pytest.raises(_lib.FtwaError) as e:
                eng.run()
            assert e.value.code == _lib.EUNSUPPORTED
        eng.close()


def test_hartree_fock_json(tmp_path, oracle_mod):
    """"method": "HF" in a simulation file (fps.cpp:88-101): <sim>_HF_res.json with the HF observables of
    the single state, against the oracle's Hartree-Fock run on the front end's matrices (1e-10)."""
    O = oracle_mod
    pos, _ = lat.chain(6)
    dt, t_max = 0.02, 0.4
    sim = lat.simulation("hf", pos, None, False, U=2.5, dt=dt, t_max=t_max,
                         coulomb={"type": "distanceInteraction", "distances": [0.0, 1.0, 2.0],
                                  "energies": [2.5, 0.6, 0.3]},
                         results={"diagDMatrix": True, "onSiteDoublons": True, "energies": True})
    sim["method"] = "HF"
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    out = json.load(open(tmp_path / "hf_HF_res.json"))
    assert out["method"] == "HF" and "HF" in out["mi355x"]["method"]
    d = _lib.load_system_json(path)
    L = d["L"]
    s = O.run(O.System(tM=d["tM"], coulomb_type=2, U=d["U"], V=d["V"], fock=True), dt, t_max, 1, d["n_up"],
              d["n_dn"], noise_arr=np.zeros((1, 2, L * (L - 1) // 2, 2)))
    ref = O.normalise(s, 1)
    sm = out["simulation"]
    for k in ("Niiu", "Niid", "NuNd"):
        assert rel(np.asarray(sm[k]), ref[k]) < FP64_RTOL, k
    assert rel(np.asarray(sm["E"]["tot"]), ref["E"][4]) < FP64_RTOL
    np.testing.assert_allclose(np.asarray(sm["NuNd"]), np.asarray(sm["Niiu"]) * np.asarray(sm["Niid"]), atol=1e-12)


def test_few_trajectories_and_single_time_point(oracle_mod):
    L = 4
    n_up = np.array([1.0, 0, 1, 0])
    compare(oracle_mod, chain(L), n_up, 1 - n_up, U=np.eye(L), B=1, t_max=0.0)
    compare(oracle_mod, chain(L), n_up, 1 - n_up, U=np.eye(L), B=3, t_max=0.06)


def test_tuning_and_chunking_are_bitwise_invariant():
    """Block shape and time chunking change the schedule, not the arithmetic of any element."""
    pos, box = lat.grid(8, 8)
    tM, Um = lat.system_arrays(pos, box, True, U=4.0)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(occ, occ)
    ref = None
    for block, chunk, kern in [(0, 0, "full"), (512, 1, "full"), (1024, 7, "full"), (512, 3, "full")]:
        eng.set_tuning(block, chunk, kern)
        got = engine_run(eng, 0.005, 0.1, 1, 64, "fp32", seed=9)
        if ref is None:
            ref = got
        else:
            for k in ref:
                np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{block} {chunk} {k}")
    again = engine_run(eng, 0.005, 0.1, 1, 64, "fp32", seed=9)
    np.testing.assert_array_equal(again["sum_u"], ref["sum_u"])  # deterministic run to run
    # the upper-triangle kernel: chunking changes nothing either, and it agrees with the full kernel
    href = None
    for chunk, block in ((0, 0), (1, 832), (6, 512), (3, 1024)):
        eng.set_tuning(block, chunk, "hermitian")
        got = engine_run(eng, 0.005, 0.1, 1, 64, "fp32", seed=9)
        if href is None:
            href = got
        else:
            for k in href:
                np.testing.assert_array_equal(got[k], href[k], err_msg=f"hermitian {chunk} {k}")
    assert rel(href["states"], ref["states"]) < 1e-5
    assert np.max(np.abs(href["states"] - np.conj(np.transpose(href["states"], (0, 1, 3, 2))))) == 0
    # the spin-pipelined kernel performs the same operations in the same order: bitwise equal
    for chunk in (0, 1, 6):
        eng.set_tuning(0, chunk, "pipelined")
        got = engine_run(eng, 0.005, 0.1, 1, 64, "fp32", seed=9)
        for k in href:
            np.testing.assert_array_equal(got[k], href[k], err_msg=f"pipelined {chunk} {k}")
    # FP64 upper-triangle kernel: 16 or 8 waves and any chunking give the same bits; it agrees with the
    # FP64 full-matrix kernel to rounding and keeps rho exactly Hermitian
    dref = None
    for chunk, block in ((0, 0), (1, 512), (5, 1024), (2, 512)):
        eng.set_tuning(block, chunk, "hermitian")
        got = engine_run(eng, 0.005, 0.1, 1, 64, "fp64", seed=9)
        if dref is None:
            dref = got
        else:
            for k in dref:
                np.testing.assert_array_equal(got[k], dref[k], err_msg=f"fp64 hermitian {chunk} {block} {k}")
    eng.set_tuning(0, 0, "full")
    full64 = engine_run(eng, 0.005, 0.1, 1, 64, "fp64", seed=9)
    assert rel(dref["states"], full64["states"]) < 1e-12
    assert np.max(np.abs(dref["states"] - np.conj(np.transpose(dref["states"], (0, 1, 3, 2))))) == 0


def test_fp32_tracks_fp64_on_identical_samples():
    pos, box = lat.grid(8, 8)
    tM, Um = lat.system_arrays(pos, box, True, U=4.0)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(occ, occ)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
    res = {}
    for e in ("fp32", "fp64"):
        eng.set_params(0.005, 0.5, 1, 256, seed=4, engine=e)
        eng.run()
        res[e] = {k: eng.result(k) for k in ("Niiu", "NuNd", "E")}
    for k in ("Niiu", "NuNd"):
        assert np.max(np.abs(res["fp32"][k] - res["fp64"][k])) < 1e-5, k
    assert np.max(np.abs(res["fp32"]["E"] - res["fp64"]["E"])) < 1e-4


def test_full_batch_invariants():
    """65,536 trajectories of the 8x8 workload (short): exact t=0 values and conservation laws."""
    sim, pos = lat.workload_hubbard8x8(n_traj=65536)
    pos, box = lat.grid(8, 8)
    U = sim["system"]["coulombInteraction"]["onSiteInteraction"]
    tM, Um = lat.system_arrays(pos, box, True, U=U)
    occ = np.array(sim["system"]["IC"]["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, U=Um)
    eng.set_initial_state(occ, occ)
    eng.set_params(0.005, 0.1, 1, 65536, seed=202, engine="fp32")
    eng.run()
    Niiu, Niid, NuNd, E = (eng.result(k) for k in ("Niiu", "Niid", "NuNd", "E"))
    np.testing.assert_array_equal(Niiu[:, 0], occ)
    np.testing.assert_array_equal(NuNd[:, 0], occ * occ)
    np.testing.assert_allclose(Niiu.sum(axis=0), 32.0, atol=2e-4)  # particle number per spin
    np.testing.assert_allclose(Niid.sum(axis=0), 32.0, atol=2e-4)
    np.testing.assert_allclose(E[4], E[4, 0], atol=1e-3 * max(1.0, abs(E[4, 0])))  # mean-field energy
    err = eng.result("Niiu_err")
    assert np.all(err[:, 0] == 0) and np.all(err[:, -1] > 0)  # sigma/sqrt(B) of sampled densities
    t = eng.timing()
    assert t["launches"] >= 1 and t["kernel_ms"] > 0


def test_json_run_writes_reference_format(tmp_path):
    sim, pos = lat.workload_chain10(n_traj=20, seed=5, engine="fp64")
    sim["parameters"]["t_max"] = 0.2
    sim["results"]["diag2DMatrix"] = True
    sim["results"]["wholeDMatrix"] = True
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    out = json.load(open(tmp_path / "chain10_fTWA_res.json"))
    assert out["fpsVersion"] == "0.5.0" and out["method"] == "fTWA" and out["system"] == sim["system"]
    assert "tbGroundState" not in out and "mfGroundState" not in out  # no eigenbasis computed for a Neel IC
    s = out["simulation"]
    T = 21
    assert len(s["t"]) == T and s["t"][1] == pytest.approx(0.01)
    for k in ("Niiu", "Niid", "N2iiu", "N2iid", "NuNd"):
        assert np.asarray(s[k]).shape == (10, T), k
    assert set(s["E"]) == {"tH", "U", "V", "pe", "tot"} and len(s["E"]["tot"]) == T
    assert set(s["N1hu"]) == {"r", "i"}
    # the same run through the C ABI directly gives the same numbers
    d = _lib.load_system_json(path)
    eng = _lib.Engine(d["tM"], U=d["U"])
    eng.set_initial_state(d["n_up"], d["n_dn"])
    eng.set_params(0.01, 0.2, 10, 20, seed=5, engine="fp64")
    eng.run()
    np.testing.assert_allclose(np.asarray(s["Niiu"]), eng.result("Niiu"), rtol=0, atol=0)
    np.testing.assert_allclose(np.asarray(s["NuNd"]), eng.result("NuNd"), rtol=0, atol=0)


def test_cli_binary_writes_the_results_file(tmp_path):
    """The fps_mi355x binary (fps.cpp's fTWA branch) on a simulation file: exit 0, the reference's
    results-file name next to the simulation, and the same numbers as the in-process front end (same
    seed, bitwise)."""
    import subprocess
    sim, pos = lat.workload_chain10(n_traj=20, seed=5, engine="fp64")
    sim["parameters"]["t_max"] = 0.2
    path = lat.write_simulation(str(tmp_path / "cli"), sim, pos)
    exe = os.path.join(ROOT, "fermiphasespace_amd", "fps_mi355x")
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Program closing" in r.stdout
    cli = json.load(open(tmp_path / "cli" / "chain10_fTWA_res.json"))
    path2 = lat.write_simulation(str(tmp_path / "lib"), sim, pos)
    _lib.run_json_file(path2)
    ref = json.load(open(tmp_path / "lib" / "chain10_fTWA_res.json"))
    assert cli["system"] == ref["system"] and cli["fpsVersion"] == ref["fpsVersion"]
    for k in ("t", "Niiu", "Niid", "NuNd"):
        assert cli["simulation"][k] == ref["simulation"][k], k
    assert cli["simulation"]["E"] == ref["simulation"]["E"]


def test_beyond_stencil_capacity_runs_dense(oracle_mod):
    """L = 80 exceeds the stencil kernels' LDS capacity: the batched engines take the dense GEMM path."""
    L = 80
    n_up = np.array([1.0, 0] * 40)
    compare(oracle_mod, chain(L), n_up, 1 - n_up, U=np.eye(L), B=2, t_max=0.04)


def test_long_range_hopping_general_coulomb_and_pulse_dense(oracle_mod):
    """All-to-all hopping (more than 8 neighbours per site), distance Coulomb U and V, site potentials
    and a Gaussian pulse on a 2D patch: the long-range workload's structure at a small L."""
    pos, box = lat.grid(5, 4)
    L = pos.shape[0]
    d = pos[:, None, :] - pos[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    tM = np.where((r > 0) & (r <= 2.5), -np.exp(-(r - 1.0) / 0.3), 0.0)
    V = np.where(r > 0, 1.5 / (1.0 + r), 0.0)
    U = 3.0 * np.eye(L) + V
    pot = np.linspace(-1.0, 1.0, L)
    pulse = dict(pulse_type=1, A0=0.5, t0=0.05, td=0.1, w0=1.5, phi=0.0, pol=(2 ** -0.5, 2 ** -0.5, 0.0), disp=d)
    n_up = np.array([1.0, 0.0] * (L // 2))
    compare(oracle_mod, tM, n_up, 1 - n_up, coulomb_type=2, U=U, V=V, pot=pot, pulse=pulse, dt=0.01, t_max=0.1,
            B=4)


@pytest.mark.parametrize("nx,ny,with_pulse", [(4, 4, True), (8, 4, False), (8, 4, True)])
def test_long_range_matrix_core_split_diagonal(oracle_mod, nx, ny, with_pulse):
    """Long-range hopping where K (2L with a pulse, L without) is a multiple of 32, so the two halves of K
    that the fused matrix-core kernel gives a diagonal tile pair are equal (L = 16 with a pulse, L = 32
    with and without): the halves must still be added.  Seven trajectories (a partial group of slices)."""
    pos, box = lat.grid(nx, ny)
    L = pos.shape[0]
    d = pos[:, None, :] - pos[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    tM = np.where((r > 0) & (r <= 2.5), -np.exp(-(r - 1.0) / 0.3), 0.0)
    pulse = (dict(pulse_type=1, A0=0.5, t0=0.05, td=0.1, w0=1.5, phi=0.0, pol=(2 ** -0.5, 2 ** -0.5, 0.0), disp=d)
             if with_pulse else None)
    n_up = np.array([1.0, 0.0] * (L // 2))
    compare(oracle_mod, tM, n_up, 1 - n_up, U=3.0 * np.eye(L), pot=np.linspace(-0.5, 0.5, L), pulse=pulse,
            dt=0.01, t_max=0.05, B=7, engines=("fp32",))


def test_long_range_hopping_without_pulse_dense(oracle_mod):
    """Long-range hopping (more than 8 neighbours per site) without a pulse, on-site U, L = 24 (a multiple
    of 8): the dense path's product of the real and imaginary planes with a real hopping matrix, through
    the library GEMM and fused on the matrix cores in FP32 (K = L)."""
    pos, box = lat.grid(6, 4)
    L = pos.shape[0]
    r = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)
    tM = np.where((r > 0) & (r <= 2.3), -np.exp(-(r - 1.0) / 0.4), 0.0)
    n_up = np.array([1.0, 0.0] * (L // 2))
    compare(oracle_mod, tM, n_up, 1 - n_up, U=2.5 * np.eye(L), pot=np.linspace(-0.5, 0.5, L), dt=0.01, t_max=0.1,
            B=5)


def _long_range_coulomb_system(nx, ny, with_pulse):
    """Long-range hopping to 2.5a and a distance Coulomb V = 1/(1 + r) with on-site U = 3."""
    pos, box = lat.grid(nx, ny)
    L = pos.shape[0]
    d = pos[:, None, :] - pos[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    tM = np.where((r > 0) & (r <= 2.5), -np.exp(-(r - 1.0) / 0.3), 0.0)
    V = np.where(r > 0, 1.0 / (1.0 + r), 0.0)
    pulse = (dict(pulse_type=1, A0=0.5, t0=0.05, td=0.1, w0=1.5, phi=0.0, pol=(2 ** -0.5, 2 ** -0.5, 0.0), disp=d)
             if with_pulse else None)
    n_up = np.array([1.0, 0.0] * (L // 2))
    return tM, 3.0 * np.eye(L), V, pulse, n_up


@pytest.mark.parametrize("with_pulse", [True, False])
def test_long_range_general_coulomb_substeps(oracle_mod, with_pulse):
    """The dense path with a distance Coulomb table and two RK sub-steps per time point: the fused
    matrix-core kernel's Hartree fields, Coulomb energies and hopping-energy partial sums (formed in the
    last stage of the last sub-step) against the oracle, with nine trajectories (a partial Hartree
    workgroup of 8 and a partial group of slices); L = 32, K = 64 with a pulse and 32 without."""
    tM, U, V, pulse, n_up = _long_range_coulomb_system(8, 4, with_pulse)
    compare(oracle_mod, tM, n_up, 1 - n_up, coulomb_type=2, U=U, V=V, pot=np.linspace(-0.3, 0.3, tM.shape[0]),
            pulse=pulse, dt=0.01, t_max=0.06, n_sub=2, B=9, engines=("fp64", "fp32"))


def test_dense_matrix_core_chunking():
    """The fused matrix-core kernel across launches of 1, 2 and 3 time points and in one launch: the
    state, the diagonal records and the doublons are bitwise the same (the first record of a launch reads
    the stage input instead of the previous launch's partial sums, so the energies agree to rounding)."""
    tM, U, V, pulse, n_up = _long_range_coulomb_system(8, 4, True)
    eng = _lib.Engine(tM, coulomb_type=2, U=U, V=V, **pulse)
    eng.set_initial_state(n_up, 1 - n_up)
    ref = None
    for chunk in (0, 1, 2, 3):
        eng.set_tuning(0, chunk, "dense_mfma")
        got = engine_run(eng, 0.01, 0.07, 2, 11, "fp32", seed=5)
        if ref is None:
            ref = got
            continue
        for k in ("states", "sum_u", "sum_d", "sum2_u", "sum_ud", "corner"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"chunk {chunk} {k}")
        assert rel(got["sum_e"], ref["sum_e"]) < 1e-6, chunk
    eng.close()


def test_moire196_shape(oracle_mod):
    """The long-range twisted bilayer (L = 196, 17-65 neighbours per site, dense distance Coulomb U and V,
    Gaussian pulse) through the dense path, two trajectories over a few steps."""
    w = lat.workload_moire196(t0=0.005, td=0.01)
    compare(oracle_mod, w["tM"], w["n_up"], w["n_dn"], coulomb_type=2, U=w["U"], V=w["V"], pulse=w["pulse"],
            dt=0.005, t_max=0.015, B=2, seed=404, engines=("fp64", "fp32"))


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_mixed_batch_lattices(oracle_mod, k):
    """Each lattice of the uneven batch (honeycomb 24, kagome 48, square 100, bcc 54, with site
    potentials) against the oracle on every engine that runs it."""
    w = lat.workload_mixed()[k]
    compare(oracle_mod, w["tM"], w["n_up"], w["n_dn"], U=w["U"], pot=w["pot"], dt=0.01, t_max=0.05, B=3, seed=303)


def test_run_many_matches_sequential_runs():
    """fpsftwa_run_many (the four systems at once, one stream each) gives bitwise the results of
    running them one after another."""
    ws = lat.workload_mixed()
    engines = []
    for w in ws:
        e = _lib.Engine(w["tM"], U=w["U"], pot=w["pot"])
        e.set_initial_state(w["n_up"], w["n_dn"])
        e.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
        e.set_params(0.01, 0.1, 1, 64, seed=303, engine="fp32")
        engines.append(e)
    _lib.run_many(engines)
    together = [{k: e.result(k) for k in ("Niiu", "NuNd", "E")} for e in engines]
    for e, got in zip(engines, together):
        e.run()
        for k in ("Niiu", "NuNd", "E"):
            assert np.array_equal(e.result(k), got[k]), k
        e.close()


def test_autotune_picks_a_kernel_that_reproduces_the_run():
    """The autotuner's choice runs the same FP64 dynamics (1e-10) as the default choice."""
    w = lat.workload_mixed()[1]
    e = _lib.Engine(w["tM"], U=w["U"], pot=w["pot"])
    e.set_initial_state(w["n_up"], w["n_dn"])
    e.set_params(0.01, 0.05, 1, 256, seed=303, engine="fp64")
    e.run()
    ref = e.result("sum_u")
    choice = e.autotune(trial_steps=3)
    assert choice["kernel"] in ("hermitian", "full", "dense", "auto") and choice["traj_steps_per_s"] > 0
    e.run()
    assert rel(e.result("sum_u"), ref) < FP64_RTOL
    e.close()


@pytest.mark.parametrize("engine", ["fp64", "fp32"])
def test_autotune_gemm_solution_on_the_long_range_path(oracle_mod, engine):
    """Long-range hopping with a pulse takes the dense path's GEMM form; the autotuner also picks the
    library's GEMM solution there, or in FP32 the fused matrix-core form when that measures faster (it
    does at this L = 36).  The tuned run still matches the oracle on identical samples (1e-10 FP64,
    2e-5 FP32)."""
    O = oracle_mod
    pos, _ = lat.grid(6, 6)
    L = pos.shape[0]
    d = pos[:, None, :] - pos[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    tM = np.where((r > 0) & (r <= 3.0), -np.exp(-(r - 1.0) / 0.3), 0.0)
    U = 3.0 * np.eye(L)
    pulse = dict(pulse_type=1, A0=0.5, t0=0.05, td=0.1, w0=1.5, phi=0.0, pol=(2 ** -0.5, 2 ** -0.5, 0.0), disp=d)
    n_up = np.array([1.0, 0.0] * (L // 2))
    B, dt, t_max = 64, 0.01, 0.1
    e = _lib.Engine(tM, U=U, **pulse)
    e.set_initial_state(n_up, 1 - n_up)
    e.set_params(dt, t_max, 1, B, seed=21, engine=engine)
    noise = e.sample_noise(0, B)
    e.set_tuning(0, 0, "dense")
    choice = e.autotune(trial_steps=3)
    assert choice["kernel"] in (("dense", "dense_mfma", "auto") if engine == "fp32" else ("dense", "auto"))
    assert choice["traj_steps_per_s"] > 0
    e.set_params(dt, t_max, 1, B, seed=21, engine=engine)
    e.set_noise(noise)
    e.run()
    ref = O.run(oracle_system(O, tM, 1, U, pulse=pulse), dt, t_max, 1, n_up, 1 - n_up, noise_arr=noise,
                keep_states=True)
    tol = FP64_RTOL if engine == "fp64" else 2e-5
    assert rel(e.states(), ref.final_states) < tol
    assert rel(e.result("sum_u"), ref.sum_u) < tol
    e.close()


def test_caller_stream():
    """fpsftwa_set_stream: the same runs on a caller's HIP stream (a torch stream, with other work queued
    ahead of them) give bitwise the same results, on the spin-sequential kernel and on the dense path
    (whose library GEMM follows the stream); NULL returns to the handle's own stream."""
    import torch
    s = torch.cuda.Stream()
    for L, kernel in ((16, "auto"), (70, "dense")):
        tM = chain(L, periodic=True)
        nu = (np.arange(L) % 2 == 0).astype(float)
        eng = _lib.Engine(tM, U=2.0 * np.eye(L))
        eng.set_initial_state(nu, 1 - nu)
        eng.set_tuning(0, 0, kernel)
        eng.set_params(0.01, 0.1, 1, 256, seed=3, engine="fp32")
        eng.run()
        ref = (eng.states(), eng.result("sum_u"), eng.result("sum_e"))
        eng.set_stream(s.cuda_stream)
        with torch.cuda.stream(s):
            x = torch.ones(1 << 22, device="cuda")
            y = x * 2.0
        eng.run()
        got = (eng.states(), eng.result("sum_u"), eng.result("sum_e"))
        torch.cuda.synchronize()
        assert float(y.sum()) == 2.0 * (1 << 22)
        eng.set_stream(None)
        eng.run()
        again = (eng.states(), eng.result("sum_u"), eng.result("sum_e"))
        for a, b, c in zip(ref, got, again):
            assert np.array_equal(a, b) and np.array_equal(a, c), (L, kernel)
        eng.close()


def test_measured_peaks_are_below_the_published_figures():
    """fpsftwa_measure_peaks: each measured rate lies between a tenth of MI355X's published figure and the
    figure itself (5 % allowance for clock spread). A wrong flop or byte count per launch (a factor of 2
    or more) leaves this window: 8 TB/s HBM, 78.6 TFLOP/s non-packed FP32 FMA (256 CUs x 4 SIMDs x 16
    lanes x 2 flop x 2.4 GHz), 78.6 TFLOP/s FP64 vector, 157.3 TFLOP/s FP32 MFMA."""
    pk = _lib.measure_peaks()
    assert pk["copy_bytes"] == 2.0 * (1 << 30)
    for key, published in (("hbm_copy_gbps", 8000.0), ("fp32_fma_tflops", 78.6), ("fp64_fma_tflops", 78.6),
                           ("fp32_mfma_tflops", 157.3)):
        assert 0.1 * published < pk[key] < 1.05 * published, (key, pk[key])
