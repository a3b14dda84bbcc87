"""CPU tests of the host side: the C ABI library's exports and the JSON front end.

The JSON front end restates FHSystem(Json, folder) (src/FHSystem.cpp:123-326): positions from .xyz
(src/FHLattice.cpp:646-710), minimum-image distances over +-1 cells (:516-539), hopping shells
(src/FHTunnelingInteraction.cpp:293-312), Coulomb shells (src/FHCoulombInteraction.cpp:240-288),
external potentials (src/FHSystem.cpp:146-169) and initial occupations (:173-293, :536-595).
These tests need no GPU: fpsftwa_load_system_json is host code.
"""
import json
import math
import os
import re
import shutil

import numpy as np
import pytest

from fermiphasespace_amd import _lib
from fermiphasespace_amd import lattices as lat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    text = open(os.path.join(ROOT, "include", "fps_ftwa.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+(fpsftwa_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == syms


def test_cli_without_arguments():
    """The fps_mi355x binary, as fps.cpp without a simulation file: the message and a non-zero exit, before
    any device call."""
    import subprocess
    exe = os.path.join(ROOT, "fermiphasespace_amd", "fps_mi355x")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "Need a json file." in r.stdout


def test_version_and_last_error():
    import ctypes
    lib = _lib.load()
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.fpsftwa_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)
    assert isinstance(_lib.last_error(), str)


def write(tmp_path, sim, pos):
    return lat.write_simulation(str(tmp_path), sim, pos)


def load(tmp_path, sim, pos):
    return _lib.load_system_json(write(tmp_path, sim, pos))


def test_chain_open(tmp_path):
    pos, _ = lat.chain(10)
    d = load(tmp_path, lat.simulation("c10", pos, None, False, t=-1.0, U=4.0), pos)
    assert d["L"] == 10 and d["coulomb_type"] == _lib.COULOMB_ONSITE
    tM = d["tM"]
    assert np.count_nonzero(tM) == 18
    assert np.all(np.diag(tM, 1) == -1.0) and np.array_equal(tM, tM.T)
    np.testing.assert_array_equal(d["U"], 4.0 * np.eye(10))
    np.testing.assert_array_equal(d["V"], 0.0)
    np.testing.assert_array_equal(d["n_up"], [1, 0] * 5)  # spinWave: up on even sites
    np.testing.assert_array_equal(d["n_dn"], [0, 1] * 5)


@pytest.mark.parametrize("n", [2, 3, 8])
def test_periodic_square_min_image(tmp_path, n):
    pos, box = lat.grid(n, n)
    d = load(tmp_path, lat.simulation("g", pos, box, True), pos)
    tM = d["tM"]
    assert np.array_equal(tM, tM.T)
    # each pair is added once, even when two images are at the same distance (n = 2)
    assert set(np.unique(tM)) <= {0.0, -1.0}
    deg = np.count_nonzero(tM, axis=1)
    assert np.all(deg == (2 if n == 2 else 4))
    if n == 8:  # torus: site (x, y) bonds to (x +- 1, y), (x, y +- 1) mod 8
        for x, y in [(0, 0), (7, 3), (4, 7)]:
            i = x + 8 * y
            want = {((x + 1) % 8) + 8 * y, ((x - 1) % 8) + 8 * y, x + 8 * ((y + 1) % 8), x + 8 * ((y - 1) % 8)}
            assert set(np.nonzero(tM[i])[0]) == want


@pytest.mark.parametrize("maker,args,z,L", [(lat.honeycomb, (2, 2), 3, 8), (lat.kagome, (4, 4), 4, 48),
                                            (lat.bcc, (3,), 8, 54), (lat.grid, (10, 10), 4, 100)])
def test_lattice_generators_coordination(tmp_path, maker, args, z, L):
    pos, box = maker(*args)
    d = load(tmp_path, lat.simulation("x", pos, box, True), pos)
    assert d["L"] == L
    assert np.all(np.count_nonzero(d["tM"], axis=1) == z)
    # the Python generator's minimum-image distances agree with the C++ front end's bonds
    tM, _ = lat.system_arrays(pos, box, True)
    np.testing.assert_array_equal(tM, d["tM"])


def test_distance_tables(tmp_path):
    pos, box = lat.grid(4, 4)
    sim = lat.simulation("dt", pos, box, True,
                         tunneling={"type": "distanceInteraction", "distances": [1.0, math.sqrt(2)],
                                    "energies": [-1.0, -0.25]},
                         coulomb={"type": "distanceInteraction", "distances": [0.0, 1.0], "energies": [3.0, 0.5]})
    d = load(tmp_path, sim, pos)
    assert d["coulomb_type"] == _lib.COULOMB_GENERAL
    assert np.count_nonzero(d["tM"] == -1.0) == 16 * 4
    assert np.count_nonzero(d["tM"] == -0.25) == 16 * 4
    np.testing.assert_array_equal(np.diag(d["U"]), 3.0)   # distance 0 enters U's diagonal
    np.testing.assert_array_equal(np.diag(d["V"]), 0.0)   # ... but never V's (fillDiagonal = false)
    assert np.count_nonzero(d["V"] == 0.5) == 16 * 4 and np.count_nonzero(d["U"] == 0.5) == 16 * 4


def test_initial_condition_subtypes(tmp_path):
    pos, _ = lat.chain(6)
    cases = {
        ("spinWave", None): ([1, 0, 1, 0, 1, 0], [0, 1, 0, 1, 0, 1]),
        ("chargeWave", None): ([1, 0, 1, 0, 1, 0], [1, 0, 1, 0, 1, 0]),  # index parity (src/FHSystem.cpp:571-583)
        ("spinWave3", None): ([1, 1, 1, 0, 0, 0], [0, 0, 0, 1, 1, 1]),
        ("chargeWave3", None): ([1, 1, 1, 0, 0, 0], [1, 1, 1, 0, 0, 0]),
        ("spinWave", (2, 1)): ([1, 0, 1, 0, 0, 0], [0, 1, 0, 0, 0, 0]),
        ("givenInJson", None): ([1, 1, 0, 0, 0, 1], [0, 0, 1, 1, 0, 0]),
    }
    for (sub, npart), (nu, nd) in cases.items():
        ic = {"type": "givenDiagonal", "subtype": sub}
        if npart:
            ic["nPartUp"], ic["nPartDown"] = npart
        if sub == "givenInJson":
            ic["IC_NuDiag"], ic["IC_NdDiag"] = nu, [False, 0, True, 1, 0, False]
        d = load(tmp_path, lat.simulation("ic", pos, None, False, ic=ic), pos)
        np.testing.assert_array_equal(d["n_up"], nu, err_msg=sub)
        np.testing.assert_array_equal(d["n_dn"], nd, err_msg=sub)


def test_external_potentials(tmp_path):
    pos, _ = lat.chain(4)
    d = load(tmp_path, lat.simulation("p", pos, None, False,
                                      external={"type": "constant", "diagonal": [0.5, -0.5, 0.25, 1.0]}), pos)
    np.testing.assert_array_equal(d["pot"], [0.5, -0.5, 0.25, 1.0])
    d = load(tmp_path, lat.simulation("p", pos, None, False,
                                      external={"type": "coreHole", "site": 2, "holePotential": -3.0}), pos)
    np.testing.assert_array_equal(d["pot"], [0, 0, -3.0, 0])


def test_reference_sample_with_legacy_keys(tmp_path):
    """grid_10.json (shipped with the reference) uses EnergyT/parametersT keys that FHSystem(Json) never
    reads: zero hopping and a null Coulomb interaction (src/FHSystem.cpp:136-143)."""
    for f in ("grid_10.json", "grid_10.xyz", "sym_grid_10.json"):
        shutil.copy(os.path.join(GOLDEN, "reference_inputs", f), tmp_path / f)
    d = _lib.load_system_json(str(tmp_path / "grid_10.json"))
    assert d["L"] == 10 and d["coulomb_type"] == _lib.COULOMB_NULL
    np.testing.assert_array_equal(d["tM"], 0.0)
    np.testing.assert_array_equal(d["n_up"], [1, 0] * 5)
    d2 = _lib.load_system_json(str(tmp_path / "sym_grid_10.json"))
    np.testing.assert_array_equal(d2["positions"][:, 0], np.arange(10))


def test_front_end_errors(tmp_path):
    pos, _ = lat.chain(4)
    sim = lat.simulation("e", pos, None, False)
    path = write(tmp_path, sim, pos)
    os.remove(tmp_path / "e.xyz")
    with pytest.raises(_lib.FtwaError) as e:
        _lib.load_system_json(path)
    assert "e.xyz" in str(e.value)
    sim = lat.simulation("g", pos, None, False, ic={"type": "groundState"})
    with pytest.raises(_lib.FtwaError) as e:
        load(tmp_path, sim, pos)
    assert e.value.code == _lib.EINVAL and "groundState" in str(e.value)
    bad = tmp_path / "bad.json"
    bad.write_text('{"system": {"lattice": ')
    with pytest.raises(_lib.FtwaError) as e:
        _lib.load_system_json(str(bad))
    assert e.value.code == _lib.EIO


def test_xyz_count_mismatch(tmp_path):
    pos, _ = lat.chain(3)
    path = write(tmp_path, lat.simulation("m", pos, None, False), pos)
    (tmp_path / "m.xyz").write_text("4\n\nC 0 0 0\nC 1 0 0\nC 2 0 0\n")
    with pytest.raises(_lib.FtwaError) as e:
        _lib.load_system_json(path)
    assert "4 sites announced, 3 found" in str(e.value)


def test_configuration_draw_is_reproducible():
    u = lat.uniform_draw(202, 0)
    assert 0 <= u < 1 and u == lat.uniform_draw(202, 0) and u != lat.uniform_draw(203, 0)
    sim, pos = lat.workload_hubbard8x8(n_traj=16)
    U = sim["system"]["coulombInteraction"]["onSiteInteraction"]
    assert 2.0 <= U <= 8.0
    assert sum(sim["system"]["IC"]["IC_NuDiag"]) == 32  # half-filled checkerboard


def test_engine_refuses_without_device():
    """No CPU fallback: creating an engine without a gfx950 device fails loudly."""
    try:
        info = _lib.device_info()
        if "gfx950" in info["name"]:
            pytest.skip("a gfx950 device is present")
    except _lib.FtwaError:
        pass
    tM = np.zeros((2, 2))
    with pytest.raises(_lib.FtwaError) as e:
        _lib.Engine(tM, U=np.eye(2))
    assert e.value.code == _lib.EDEVICE


def test_create_argument_checks():
    with pytest.raises(_lib.FtwaError) as e:
        _lib.Engine(np.zeros((3, 3)), coulomb_type=_lib.COULOMB_GENERAL, U=np.eye(3))
    assert e.value.code == _lib.EINVAL
    with pytest.raises(_lib.FtwaError):
        _lib.Engine(np.zeros((3, 3)), U=np.eye(3), pulse_type=1, A0=1.0)  # pulse without displacements


def _chain_tm(L, t=-1.0, periodic=False):
    tM = t * (np.eye(L, k=1) + np.eye(L, k=-1))
    if periodic:
        tM[0, L - 1] = tM[L - 1, 0] = t
    return tM


def test_ground_state_matches_lapack(oracle_mod):
    """fpsftwa_ground_state (Jacobi) against numpy's LAPACK eigh in the same sign gauge: random symmetric
    matrices, chains, and mean-field iterations that converge (iteration counts equal too)."""
    O = oracle_mod
    rng = np.random.default_rng(11)
    for L in (2, 7, 24):
        A = rng.normal(size=(L, L))
        A = A + A.T
        g = _lib.ground_state(A)
        np.testing.assert_allclose(g["rot_up"], O.ground_state(A, None, 1, 1, False)[0], atol=1e-12)
        np.testing.assert_array_equal(g["rot_up"], g["rot_dn"])
        np.testing.assert_allclose(g["energy_up"], np.linalg.eigvalsh(A), atol=1e-12 * np.abs(A).max())
        np.testing.assert_array_equal(g["energy_up"], g["energy_dn"])
    for L, nu, nd in ((6, 3, 3), (6, 4, 2), (8, 5, 3), (9, 5, 4), (16, 9, 7), (30, 16, 14)):
        tM = _chain_tm(L)
        U = 4.0 * np.eye(L)
        g = _lib.ground_state(tM, U, nu, nd, mean_field=True)
        Vu, Vd, it, res, Eu, Ed = O.ground_state(tM, U, nu, nd, True, energies=True)
        assert g["iterations"] == it and abs(g["residue"] - res) < 1e-9, (L, nu, nd)
        np.testing.assert_allclose(g["rot_up"], Vu, atol=1e-10)
        np.testing.assert_allclose(g["rot_dn"], Vd, atol=1e-10)
        np.testing.assert_allclose(g["energy_up"], Eu, atol=1e-12)
        np.testing.assert_allclose(g["energy_dn"], Ed, atol=1e-12)
        np.testing.assert_allclose(_lib.ground_state(tM)["rot_up"], O.ground_state(tM, None, nu, nd, False)[0],
                                   atol=1e-12)


def test_ground_state_degenerate_spectrum(oracle_mod):
    """Periodic ring: cos k levels are doubly degenerate, so any orthonormal basis of a level is valid.  The
    basis is orthogonal, diagonalises tM with LAPACK's eigenvalues, and the occupied projector at a closed
    shell (n = 5 of 10: levels 1 + 2 + 2) equals LAPACK's."""
    O = oracle_mod
    L = 10
    tM = _chain_tm(L, periodic=True)
    V = _lib.ground_state(tM)["rot_up"]
    np.testing.assert_allclose(V.T @ V, np.eye(L), atol=1e-13)
    np.testing.assert_allclose(V.T @ tM @ V, np.diag(np.linalg.eigvalsh(tM)), atol=1e-12)
    Vo = O.ground_state(tM, None, 5, 5, False)[0]
    np.testing.assert_allclose(V[:, :5] @ V[:, :5].T, Vo[:, :5] @ Vo[:, :5].T, atol=1e-12)


def test_ground_state_argument_checks():
    tM = _chain_tm(4)
    with pytest.raises(_lib.FtwaError):
        _lib.ground_state(tM, None, 5, 1)
    with pytest.raises(_lib.FtwaError):
        _lib.ground_state(tM, None, -1, 1)


def test_eigenbasis_initial_conditions_from_json(tmp_path, oracle_mod):
    """IC types tbGroundState / mfGroundState (src/FHSystem.cpp:262-270, :536-595; src/FHfTWASolver.cpp:120-153):
    nPartUp / nPartDown lowest states occupied, rotations = eigenvectors of tM (both spins) or of the
    self-consistent Hartree Hamiltonian per spin; other IC types keep the identity."""
    O = oracle_mod
    pos, _ = lat.chain(7)
    path = write(tmp_path, lat.simulation("tb", pos, None, False, ic={"type": "tbGroundState"}), pos)
    d = _lib.load_initial_state_json(path)
    np.testing.assert_array_equal(d["n_up"], [1, 1, 1, 1, 0, 0, 0])  # nPartUp default ceil(L/2)
    np.testing.assert_array_equal(d["n_dn"], [1, 1, 1, 0, 0, 0, 0])  # nPartDown default floor(L/2)
    tM = _lib.load_system_json(path)["tM"]
    V = O.ground_state(tM, None, 4, 3, False)[0]
    np.testing.assert_allclose(d["rot_up"], V, atol=1e-12)
    np.testing.assert_allclose(d["rot_dn"], V, atol=1e-12)
    path = write(tmp_path, lat.simulation("mf", pos, None, False, U=3.0,
                                          ic={"type": "mfGroundState", "nPartUp": 5, "nPartDown": 2}), pos)
    d = _lib.load_initial_state_json(path)
    np.testing.assert_array_equal(d["n_up"], [1, 1, 1, 1, 1, 0, 0])
    np.testing.assert_array_equal(d["n_dn"], [1, 1, 0, 0, 0, 0, 0])
    Vu, Vd, it, _ = O.ground_state(tM, 3.0 * np.eye(7), 5, 2, True)
    assert 1 < it < 100
    np.testing.assert_allclose(d["rot_up"], Vu, atol=1e-10)
    np.testing.assert_allclose(d["rot_dn"], Vd, atol=1e-10)
    assert np.abs(d["rot_up"].imag).max() == 0
    from fermiphasespace_amd import FHSystem
    fs = FHSystem.from_json_file(path)
    np.testing.assert_allclose(fs.rot_up, Vu, atol=1e-10)
    fs2 = FHSystem(tM=tM, U=3.0 * np.eye(7))
    fs2.setGroundStateInitialConditions("mfGroundState", 5, 2)
    np.testing.assert_allclose(fs2.rot_dn, Vd, atol=1e-10)
    np.testing.assert_array_equal(fs2.n_up, d["n_up"])
    d = _lib.load_initial_state_json(write(tmp_path, lat.simulation("sw", pos, None, False), pos))
    np.testing.assert_array_equal(d["rot_up"], np.eye(7))
    assert FHSystem.from_json_file(str(tmp_path / "sw.json")).rot_up is None
    with pytest.raises(_lib.FtwaError) as e:
        _lib.load_initial_state_json(write(tmp_path, lat.simulation(
            "big", pos, None, False, ic={"type": "tbGroundState", "nPartUp": 8}), pos))
    assert "nPartUp" in str(e.value)


def test_core_hole_system(tmp_path):
    """coreHole (src/FHSystem.cpp:295-323, src/FHCoreHole.cpp:10-26): site L added, tM[site, L] = X,
    U[site, L] = U[L, site] = G/2 and U[L, L] = Uc, V[site, L] = V[L, site] = G/2 (addSite halves the vector
    into the last row and column, src/FHCoulombInteraction.cpp:224-230), potential eps on the hole alone
    (the file's potential is replaced), occupations IC = [up, down]; defaults eps -2, X -1, Uc 1, G 1, [1, 0]."""
    pos, _ = lat.chain(4)
    sim = lat.simulation("ch", pos, None, False, U=3.0, external={"type": "constant", "diagonal": [0.5, 0, 0, 0]})
    sim["system"]["coreHole"] = {"presence": True, "site": 1, "eps": -2.5, "X": -0.7, "Uc": 1.5, "G": 0.8,
                                 "IC": [0.75, 0.25]}
    d = load(tmp_path, sim, pos)
    assert d["L"] == 5 and d["coulomb_type"] == _lib.COULOMB_GENERAL
    tM = np.zeros((5, 5))
    tM[:4, :4] = -(np.eye(4, k=1) + np.eye(4, k=-1))
    tM[1, 4] = tM[4, 1] = -0.7
    np.testing.assert_array_equal(d["tM"], tM)
    U = np.diag([3.0, 3, 3, 3, 1.5])
    U[1, 4] = U[4, 1] = 0.4
    np.testing.assert_allclose(d["U"], U, atol=1e-15)
    V = np.zeros((5, 5))
    V[1, 4] = V[4, 1] = 0.4
    np.testing.assert_allclose(d["V"], V, atol=1e-15)
    np.testing.assert_array_equal(d["pot"], [0, 0, 0, 0, -2.5])
    np.testing.assert_array_equal(d["n_up"], [1, 0, 1, 0, 0.75])
    np.testing.assert_array_equal(d["n_dn"], [0, 1, 0, 1, 0.25])
    np.testing.assert_array_equal(d["positions"][4], d["positions"][1])
    sim["system"]["coreHole"] = {"presence": True}
    sim["system"]["IC"] = {"type": "tbGroundState", "nPartUp": 2, "nPartDown": 2}
    d = load(tmp_path, sim, pos)
    assert d["tM"][0, 4] == -1.0 and d["U"][4, 4] == 1.0 and d["U"][0, 4] == 0.5 and d["pot"][4] == -2.0
    np.testing.assert_array_equal(d["n_up"], [1, 1, 0, 0, 1])
    np.testing.assert_array_equal(d["n_dn"], [1, 1, 0, 0, 0])
    ic = _lib.load_initial_state_json(str(tmp_path / "ch.json"))
    assert ic["rot_up"].shape == (5, 5) and ic["rot_up"][4, 4] == 1 and np.all(ic["rot_up"][4, :4] == 0)
    sim["system"]["coulombInteraction"] = {"type": "distanceInteraction", "distances": [0.0], "energies": [2.0]}
    with pytest.raises(_lib.FtwaError) as e:
        load(tmp_path, sim, pos)
    assert "core hole" in str(e.value)


# ---- the system generator (fps_make_system, makeSitesSystem.cpp) ----

MAKE = os.path.join(ROOT, "fermiphasespace_amd", "fps_make_system")


def make_system(tmp_path, *args):
    """Runs fps_make_system in tmp_path; returns (simulation dict, positions [L][3] read from the .xyz)."""
    import subprocess
    r = subprocess.run([MAKE, *map(str, args)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    name = args[0]
    sim = json.load(open(tmp_path / "systemsFolder" / f"{name}.json"))
    lines = open(tmp_path / "systemsFolder" / f"{name}.xyz").read().split("\n")
    n = int(lines[0])
    assert lines[1] == "" and len(lines) == n + 2  # count, empty comment, one line per site, no final newline
    pos = np.array([[float(x) for x in ln.split("\t")[1:]] for ln in lines[2:]])
    assert all(ln.split("\t")[0] == "C" for ln in lines[2:])
    return sim, pos, r.stdout


def _distances(pos):
    return np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)


def _unique_points(pts, tol=1e-6):
    """Brute-force count of distinct points (greedy clustering, independent of the generator's order)."""
    keep = []
    for p in pts:
        if not any(np.abs(p - q).sum() < tol for q in keep):
            keep.append(p)
    return np.array(keep)


def test_make_system_matches_the_reference_file(tmp_path):
    """`honeyCombs 16` (the lattice behind the reference's systemsFolder/hc16_CH.json, written by
    makeSitesSystem): the same lattice block (box digits included), hopping, Coulomb and potential blocks;
    16 combs of the square-as-possible sequence hold 48 distinct sites with coordination <= 3."""
    sim, pos, out = make_system(tmp_path, "hc16", "honeyCombs", 16)
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_inputs", "hc16_CH.json")))
    for key in ("lattice", "tunnelingInteraction", "coulombInteraction", "externalPotential"):
        assert sim["system"][key] == ref["system"][key], key
    assert sim["method"] == "HF" and sim["sim_name"] == "hc16"
    assert sim["parameters"] == {"dt": 0.01, "int_order": 4, "n_traj": 1000, "stepsPerTimePoints": 1, "t_max": 2.0}
    assert sim["system"]["IC"] == {"type": "givenDiagonal", "subtype": "spinWave"}
    assert "== hc16 system created with 48 sites and 48 particles." in out
    d = _distances(pos)
    np.fill_diagonal(d, np.inf)
    assert len(pos) == 48 and d.min() > 0.99999
    assert ((np.abs(d - 1.0) < 1e-4).sum(axis=1) <= 3).all()


@pytest.mark.parametrize("args,L,nn,zmax", [
    (("grid", 3, 4, 3, 1.5), 36, 1.5, 6),
    (("bcc", 2, 2, 2, 1.0), 27 + 8, math.sqrt(3) / 2, 8),
    (("fcc", 2, 2, 2, 1.0), 27 + 3 * 3 * 2 * 2, 1 / math.sqrt(2), 12),
    (("triangular", 4, 3, 1, 1.0), 12, 1.0, 6),
    (("triangular", 3, 3, 3, 1.0), 27, 1.0, 12),
    (("graphene", 4, 3, 1.0), 24, 1.0, 3),
])
def test_make_system_generators(tmp_path, args, L, nn, zmax):
    """Site counts of the closed forms (bcc: corners + cube centres; fcc: corners + face centres), the
    nearest-neighbour distance of each lattice (grid a, bcc a sqrt3/2, fcc a/sqrt2, triangular and
    tetrahedral a, graphene a) and the coordination bound, with no site repeated."""
    sim, pos, _ = make_system(tmp_path, "s", *args)
    assert len(pos) == L
    d = _distances(pos)
    np.fill_diagonal(d, np.inf)
    assert abs(d.min() - nn) < 1e-5
    z = (np.abs(d - nn) < 1e-4).sum(axis=1)
    assert z.max() == zmax
    assert sim["system"]["lattice"]["latticeType"] == args[0]


def test_make_system_honeycomb_sequence(tmp_path):
    """The comb sequence (0,0), (1,0), (0,1), (1,1), (2,0), ... with shared corners removed: every prefix of
    n combs holds exactly the distinct corners of those combs (brute force over the union), and the box
    is the ceil(sqrt n) x (floor(sqrt n) + ...) parallelogram of the reference."""
    s3 = math.sqrt(3)
    corners = np.array([[-s3 / 2, -0.5, 0], [-s3 / 2, 0.5, 0], [0, 1, 0], [0, -1, 0], [s3 / 2, -0.5, 0],
                        [s3 / 2, 0.5, 0]]) * 0.8
    seq = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (2, 1), (0, 2), (1, 2), (2, 2), (3, 0)]
    for n in (1, 2, 5, 9, 10):
        sim, pos, _ = make_system(tmp_path, f"h{n}", "honeyCombs", n, 0.8)
        pts = np.vstack([corners + np.array([s3 * 0.8 * i + s3 / 2 * 0.8 * j, 1.5 * 0.8 * j, 0]) for i, j in seq[:n]])
        want = _unique_points(pts)
        assert len(pos) == len(want), n
        assert all(np.abs(want - p).sum(axis=1).min() < 1e-5 for p in pos)
        r = math.sqrt(n)
        nx, ny = math.ceil(r), math.floor(r) + (1 if n - math.floor(r) ** 2 > math.floor(r) else 0)
        box = np.array(sim["system"]["lattice"]["latticeMatrix"])
        assert np.allclose(box, [[s3 * 0.8 * nx, s3 / 2 * 0.8 * ny, 0], [0, 1.5 * 0.8 * ny, 0], [0, 0, 8.0]])


def test_make_system_bilayers(tmp_path):
    """moire nx ny a b theta: the package's twisted-bilayer generator (the 196-site workload's) site by site;
    moireHC n a b theta: the second flake is the first rotated by theta about z and lifted by b (a = 1),
    each flake's sites distinct apart from the reference's unchecked first point."""
    sim, pos, _ = make_system(tmp_path, "m", "moire", 3, 2, 1.0, 2.35, 5.0)
    want, box = lat.moire(3, 2, 1.0, 2.35, 5.0)
    assert np.allclose(pos, want, atol=5e-6)
    assert np.allclose(np.array(sim["system"]["lattice"]["latticeMatrix"]), box)
    sim, pos, _ = make_system(tmp_path, "mh", "moireHC", 7, 1.0, 3.0, 10.0)
    half = len(pos) // 2
    assert len(pos) == 2 * half
    l0, l1 = pos[:half], pos[half:]
    c, s = math.cos(math.radians(10)), math.sin(math.radians(10))
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    assert np.allclose(l0 @ R.T + [0, 0, 3.0], l1, atol=5e-6)
    assert len(_unique_points(l0[1:], 1e-4)) == half - 1


def test_make_system_initial_conditions(tmp_path):
    """ic_cw: charge wave (ceil(L/2) of each spin); ic_tbgs / ic_mfgs: the ground-state blocks, whose
    tight-binding energies are the eigenvalues of the nearest-neighbour hopping built from the written
    positions (numpy, to 1e-12); the file loads through the JSON front end.  After a word that is not a
    number, later numbers are not read (the reference's argument loop)."""
    sim, pos, out = make_system(tmp_path, "cw", "graphene", 3, 2, "ic_cw")
    assert sim["system"]["IC"] == {"type": "givenDiagonal", "subtype": "chargeWave"}
    assert "with 12 sites and 12 particles" in out
    s = _lib.load_system_json(str(tmp_path / "systemsFolder" / "cw.json"))
    assert np.array_equal(s["n_up"], s["n_dn"]) and s["n_up"].sum() == 6
    sim, pos, _ = make_system(tmp_path, "tb", "graphene", 3, 2, "ic_tbgs")
    d = _distances(pos)
    tM = np.where(np.abs(d - 1.0) < 1e-4, -1.0, 0.0)
    e = np.linalg.eigvalsh(tM)
    assert np.allclose(sim["system"]["tbGroundState"]["energyVector"], e, atol=1e-12)
    V = np.array(sim["system"]["tbGroundState"]["rotationMatrix"])
    assert np.allclose(V.T @ tM @ V, np.diag(e), atol=1e-10)
    sim, pos, _ = make_system(tmp_path, "mf", "graphene", 3, 2, "ic_mfgs")
    mf = sim["system"]["mfGroundState"]
    assert set(mf) == {"energyVectorUp", "energyVectorDown", "rotationMatrixUp", "rotationMatrixDown",
                       "iterationsNeeded", "residueComputation"}
    assert mf["iterationsNeeded"] >= 1
    sim, pos, _ = make_system(tmp_path, "q", "grid", 3, "x", 4, 2)
    assert len(pos) == 3  # 3 x 1 x 1: the numbers after "x" are not read


def test_make_system_without_arguments():
    """fps_make_system alone prints the usage, as makeSitesSystem does."""
    import subprocess
    r = subprocess.run([MAKE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "./makeSitesSystem <nameOfTheSystem> <geometryOfTheSites>" in r.stdout
