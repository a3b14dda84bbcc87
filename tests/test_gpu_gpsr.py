"""GPU parity of the GPSR method (FHGPSRSolver) against the oracle, through the C ABI.

Every engine draws the same counter-based diffusion noise as the oracle (seeded), so trajectories can be
compared one by one.  Bars: FP64 engines match the oracle's final states and ensemble sums to 1e-9 relative
(rounding differences of the noise's log / sincos grow mildly along a GPSR trajectory); FP32 to the
tolerances written next to each check.  Physics: the FP64 batched ensemble follows exact diagonalisation
where GPSR is exact, as the oracle does.
"""
import json

import numpy as np
import pytest

from fermiphasespace_amd import _lib
from fermiphasespace_amd import lattices as lat

pytestmark = pytest.mark.gpu
RTOL64 = 1e-9


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def chain(L, t=-1.0, reach=1):
    tM = np.zeros((L, L))
    for d in range(1, reach + 1):
        for i in range(L - d):
            tM[i, i + d] = tM[i + d, i] = t / d
    return tM


def run_engine(eng, dt, t_max, B, engine, seed, bases=False):
    eng.set_params(dt, t_max, 1, B, seed=seed, engine=engine)
    eng.run()
    out = {k: eng.result(k) for k in ("sum_u", "sum_d", "sum2_u", "sum_ud", "sum_e")}
    out["corner"] = np.vstack([eng.result("N1hu"), eng.result("N1hd")]) * B
    out["states"] = eng.states()
    if bases:
        out["sum_basis"] = eng.result("sum_basis")
    return out


def gpsr_compare(O, tM, n_up, n_dn, *, coulomb_type=1, U=None, V=None, pot=None, dt=0.02, t_max=0.4, B=48, seed=9,
                 engines=("fp64_plain", "fp64", "fp32"), fp32_tol=5e-4, rot=None, bases=None):
    L = tM.shape[0]
    eng = _lib.Engine(tM, coulomb_type=coulomb_type, U=U, V=V, pot=pot)
    eng.set_method("gpsr")
    if rot is None:
        eng.set_initial_state(n_up, n_dn)
    else:
        eng.set_initial_state(n_up, n_dn, rot[0], rot[1])
    if bases is not None:
        eng.set_bases(bases)
    ref = O.run(O.System(tM=tM, coulomb_type=coulomb_type, U=U, V=V, pot=pot, gpsr=True, gpsr_seed=seed), dt, t_max,
                1, n_up, n_dn, n_traj=B, seed=seed, keep_states=True, rot_up=None if rot is None else rot[0],
                rot_dn=None if rot is None else rot[1], bases=bases)
    assert np.isfinite(ref.final_states).all()
    for engine in engines:
        got = run_engine(eng, dt, t_max, B, engine, seed, bases=bases is not None)
        tol = fp32_tol if engine == "fp32" else RTOL64
        assert rel(got["states"], ref.final_states) < tol, (engine, rel(got["states"], ref.final_states))
        for k, r in (("sum_u", ref.sum_u), ("sum_d", ref.sum_d), ("sum2_u", ref.sum2_u), ("sum_ud", ref.sum_ud),
                     ("sum_e", ref.sum_e), ("corner", ref.sum_corner)):
            assert rel(got[k], r) < tol, (engine, k, rel(got[k], r))
        if bases is not None:
            assert rel(got["sum_basis"], ref.sum_basis) < tol, engine
    eng.close()


def test_gpsr_onsite_chain(oracle_mod):
    """On-site U with diffusion, a site potential and the spin-wave start, every engine (L = 6: N = 8 kernel)."""
    L = 6
    nu = np.array([1.0, 0, 1, 0, 1, 0])
    gpsr_compare(oracle_mod, chain(L), nu, 1 - nu, U=1.5 * np.eye(L), pot=np.linspace(-0.3, 0.4, L))


def test_gpsr_shapes_and_neighbour_widths(oracle_mod):
    """L = 3 and 13 (N = 8 and 16 kernels) with 4 neighbours, L = 12 with 8 (hopping to distance 4),
    L = 20 (N = 32) on a periodic 4 x 5 grid."""
    O = oracle_mod
    for L, reach in ((3, 1), (13, 2), (12, 4)):
        nu = (np.arange(L) % 2 == 0).astype(float)
        gpsr_compare(O, chain(L, reach=reach), nu, 1 - nu, U=0.8 * np.eye(L), t_max=0.2, B=24)
    pos, box = lat.grid(4, 5)
    tM, _ = lat.system_arrays(pos, box, True)
    L = tM.shape[0]
    nu = np.array(lat.checkerboard(4, 5)["IC_NuDiag"], dtype=float)
    gpsr_compare(O, tM, nu, 1 - nu, U=0.8 * np.eye(L), t_max=0.2, B=24)


def test_gpsr_general_coulomb_is_deterministic_drift(oracle_mod):
    """A distance interaction: GPSR's Delta uses the whole U matrix and no V, and the reference adds no
    diffusion (src/FHGPSRSolver.cpp:248), so every trajectory is the same deterministic one."""
    L = 6
    dist = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :]).astype(float)
    V = np.where(dist > 0, 0.7 / (1.0 + dist), 0.0)
    U = 2.0 * np.eye(L) + V
    nu = np.array([1.0, 1, 0, 1, 0, 0])
    nd = np.array([0.0, 1, 1, 0, 0, 1])
    gpsr_compare(oracle_mod, chain(L), nu, nd, coulomb_type=2, U=U, V=V, B=4)
    eng = _lib.Engine(chain(L), coulomb_type=2, U=U, V=V)
    eng.set_method("gpsr")
    eng.set_initial_state(nu, nd)
    st = run_engine(eng, 0.02, 0.4, 4, "fp64", 1)["states"]
    assert np.abs(st - st[:1]).max() == 0.0
    eng.close()


def test_gpsr_rotated_start_and_basis_occupations(oracle_mod):
    """The eigenbasis start R diag(n) R^dagger and the tbStates occupations Re diag(V^T rho V) of the
    non-Hermitian GPSR state."""
    L = 8
    tM = chain(L)
    V = np.linalg.eigh(tM)[1]
    occ = np.array([1.0, 1, 1, 0, 0, 0, 0, 0])
    bases = np.stack([np.stack([V, V])])
    gpsr_compare(oracle_mod, tM, occ, occ, U=1.2 * np.eye(L), rot=(V.astype(complex), V.astype(complex)),
                 bases=bases, t_max=0.3)


def test_gpsr_divergent_trajectories(oracle_mod):
    """Trajectories that overflow become NaN everywhere from the next step on (src/FHGPSRSolver.cpp:611-615):
    every final state is all finite or all NaN, the same trajectories as in the oracle, and the finite ones
    match it (1e-9 for nine in ten; rounding differences grow on the ones near blowing up)."""
    O = oracle_mod
    L, B, seed = 4, 200, 2
    nu = np.array([1.0, 0, 1, 0])
    ref = O.run(O.System(tM=chain(L), coulomb_type=1, U=2.0 * np.eye(L), gpsr=True, gpsr_seed=seed), 0.02, 1.0, 1,
                nu, 1 - nu, n_traj=B, seed=seed, keep_states=True)
    bad_ref = ~np.isfinite(ref.final_states).reshape(B, -1).all(axis=1)
    assert 0 < bad_ref.sum() < B
    eng = _lib.Engine(chain(L), coulomb_type=1, U=2.0 * np.eye(L))
    eng.set_method("gpsr")
    eng.set_initial_state(nu, 1 - nu)
    for engine in ("fp64_plain", "fp64"):
        got = run_engine(eng, 0.02, 1.0, B, engine, seed)
        fin = np.isfinite(got["states"]).reshape(B, -1)
        assert np.all(fin.all(axis=1) | (~fin).all(axis=1)), engine
        assert np.array_equal(~fin.all(axis=1), bad_ref), engine
        ok = ~bad_ref
        # the surviving trajectories include some close to blowing up, where differences of rounding (FMA
        # contraction) grow fastest: most match to 1e-9, none worse than 1e-3
        d = np.array([rel(g, r) for g, r in zip(got["states"][ok], ref.final_states[ok])])
        assert np.quantile(d, 0.9) < 1e-9 and d.max() < 1e-3, (engine, np.quantile(d, 0.9), d.max())
        assert np.isnan(got["sum_u"][:, -1]).all() and np.isfinite(got["sum_u"][:, 0]).all()
    eng.close()


def test_gpsr_against_exact_diagonalisation():
    """Physics check on the device: the FP64 batched GPSR ensemble (4-site chain, U = 1, Neel start,
    200,000 trajectories) follows exact diagonalisation to t = 0.5 -- densities within 4 standard errors,
    <n_up n_dn> within 1.5e-3 (the ED values are those of the oracle's test_gpsr_against_exact_diagonalisation,
    recomputed here by the ED oracle)."""
    from oracle import ed
    L, U, B, dt, t_end = 4, 1.0, 200000, 0.01, 0.5
    tM = chain(L)
    nu = np.array([1.0, 0, 1, 0])
    r = ed.evolve_product_state(L, tM, U * np.ones(L), nu, 1 - nu, t_end, 6)
    idx = np.round(r["t"] / dt).astype(int)
    eng = _lib.Engine(tM, coulomb_type=1, U=U * np.eye(L))
    eng.set_method("gpsr")
    eng.set_initial_state(nu, 1 - nu)
    eng.set_params(dt, t_end, 1, B, seed=21, engine="fp64")
    eng.run()
    n = eng.result("Niiu")[:, idx]
    err = eng.result("Niiu_err")[:, idx]
    assert np.all(np.abs(n - r["n_up"]) <= 4 * err + 1e-9)
    assert np.all(np.abs(eng.result("NuNd")[:, idx] - r["nund"]) <= 1.5e-3)
    eng.close()


def test_gpsr_json(tmp_path, oracle_mod):
    """"method": "GPSR" in a simulation file (fps.cpp:208-220): <sim>_GPSR_res.json with the plain GPSR means
    (src/FHGPSRSolver.cpp:895-929) of the batched FP64 engine, against the oracle on the front end's matrices
    (1e-9); a system beyond the batched kernel (L = 40) runs on the plain FP64 engine."""
    O = oracle_mod
    pos, _ = lat.chain(6)
    dt, t_max = 0.02, 0.3
    sim = lat.simulation("gp", pos, None, False, U=1.5, dt=dt, t_max=t_max, n_traj=32, mi355x={"seed": 17},
                         results={"diagDMatrix": True, "diag2DMatrix": True, "onSiteDoublons": True,
                                  "energies": True, "wholeDMatrix": True})
    sim["method"] = "GPSR"
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    out = json.load(open(tmp_path / "gp_GPSR_res.json"))
    assert out["method"] == "GPSR" and "GPSR" in out["mi355x"]["method"]
    d = _lib.load_system_json(path)
    s = O.run(O.System(tM=d["tM"], coulomb_type=d["coulomb_type"], U=d["U"], gpsr=True, gpsr_seed=17), dt, t_max, 1,
              d["n_up"], d["n_dn"], n_traj=32, seed=17)
    ref = O.normalise(s, 32, gpsr=True)
    sm = out["simulation"]
    for k in ("Niiu", "Niid", "N2iiu", "N2iid", "NuNd"):
        assert rel(np.asarray(sm[k]), ref[k]) < RTOL64, k
    for i, k in enumerate(("tH", "U", "V", "pe", "tot")):
        assert rel(np.asarray(sm["E"][k]), ref["E"][i]) < RTOL64 or np.abs(ref["E"][i]).max() == 0, k
    np.testing.assert_allclose(np.asarray(sm["N1hu"]["r"]), s.sum_corner[0] / 32, rtol=0, atol=1e-12)
    pos, _ = lat.chain(40)
    sim = lat.simulation("big", pos, None, False, U=1.0, dt=0.02, t_max=0.04, n_traj=4,
                         results={"diagDMatrix": True})
    sim["method"] = "GPSR"
    path = lat.write_simulation(str(tmp_path), sim, pos)
    _lib.run_json_file(path)
    out = json.load(open(tmp_path / "big_GPSR_res.json"))
    assert out["mi355x"]["engine"] == "fp64_plain"
    assert np.isfinite(np.asarray(out["simulation"]["Niiu"])).all()


def test_gpsr_substeps_and_pulse_on_the_plain_engine(oracle_mod):
    """Sub-steps (the diffusion noise of every RK step, counter t * nsub + sub, variance dt / nsub) on every
    engine; a Gaussian pulse (Peierls phase on tM) on the plain FP64 engine, which the batched GPSR kernel
    refuses."""
    O = oracle_mod
    L, B, seed, dt, t_max = 6, 16, 4, 0.04, 0.32
    tM = chain(L)
    nu = np.array([1.0, 0, 1, 0, 1, 0])
    U = 1.2 * np.eye(L)
    ref = O.run(O.System(tM=tM, coulomb_type=1, U=U, gpsr=True, gpsr_seed=seed), dt, t_max, 3, nu, 1 - nu, n_traj=B,
                seed=seed, keep_states=True)
    eng = _lib.Engine(tM, coulomb_type=1, U=U)
    eng.set_method("gpsr")
    eng.set_initial_state(nu, 1 - nu)
    for engine, tol in (("fp64_plain", RTOL64), ("fp64", RTOL64), ("fp32", 5e-4)):
        eng.set_params(dt, t_max, 3, B, seed=seed, engine=engine)
        eng.run()
        assert rel(eng.states(), ref.final_states) < tol, engine
        assert rel(eng.result("sum_ud"), ref.sum_ud) < tol, engine
    eng.close()
    disp = np.zeros((L, L, 3))
    disp[..., 0] = np.arange(L)[:, None] - np.arange(L)[None, :]
    pulse = {"pulse_type": 1, "A0": 0.6, "t0": 0.15, "td": 0.1, "w0": 8.0, "phi": 0.0, "pol": (1.0, 0.0, 0.0),
             "disp": disp}
    ref = O.run(O.System(tM=tM, coulomb_type=1, U=U, gpsr=True, gpsr_seed=seed, **pulse), dt, t_max, 1, nu, 1 - nu,
                n_traj=B, seed=seed, keep_states=True)
    eng = _lib.Engine(tM, coulomb_type=1, U=U, **pulse)
    eng.set_method("gpsr")
    eng.set_initial_state(nu, 1 - nu)
    eng.set_params(dt, t_max, 1, B, seed=seed, engine="fp64_plain")
    eng.run()
    assert rel(eng.states(), ref.final_states) < RTOL64
    assert rel(eng.result("sum_e"), ref.sum_e) < RTOL64
    eng.set_params(dt, t_max, 1, B, seed=seed, engine="fp64")
    with pytest.raises(_lib.FtwaError) as e:
        eng.run()
    assert e.value.code == _lib.EUNSUPPORTED
    eng.close()
