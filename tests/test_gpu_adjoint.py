"""Discrete adjoint (fpsftwa_run_adjoint) on the MI355X, through the C ABI.

Bars:
  * the objective J (mean over sites of NuNd at the last time point) equals the FP64 engine's and the
    oracle's on the same samples to 1e-12 relative;
  * every FP64 gradient entry agrees with central finite differences of the oracle's J (dense-Delta
    restatement of the reference, identical samples, h = 1e-5) to 1e-6 relative to the largest entry;
  * checkpoint spacing does not change a bit of the result;
  * the FP32 adjoint agrees with the FP64 adjoint to 1e-4 relative to the largest entry.
"""
import numpy as np
import pytest

from fermiphasespace_amd import _lib
from fermiphasespace_amd import lattices as lat

pytestmark = pytest.mark.gpu

L = 8
DT, TMAX, NSUB, B, SEED = 0.02, 0.4, 1, 8, 5
KNOTS = np.array([0.0, 0.6, -0.3, 0.4, 0.1])


@pytest.fixture(scope="module")
def oracle_mod():
    from oracle import oracle as O

    return O


def ring():
    tM = np.zeros((L, L))
    for i in range(L):
        tM[i, (i + 1) % L] = tM[(i + 1) % L, i] = -1.0
    tM[0, 0] = 0.3  # a diagonal entry, so its gradient is exercised
    pos, _ = lat.chain(L)
    d = pos[:, None, :] - pos[None, :, :]
    d[0, L - 1, 0], d[L - 1, 0, 0] = 1.0, -1.0  # minimum image across the ring
    return tM, d


def occupations():
    n_up = np.array([1.0, 0.0] * (L // 2))
    n_dn = np.array([0.8, 0.3] * (L // 2))
    return n_up, n_dn


def make_engine(tM, U, pulse):
    eng = _lib.Engine(tM, coulomb_type=_lib.COULOMB_ONSITE, U=np.diag(U), **pulse)
    eng.set_initial_state(*occupations())
    return eng


def knot_pulse(knots, d):
    return dict(pulse_type=_lib.PULSE_KNOTS, knots=knots, pol=(1.0, 0.0, 0.0), disp=d)


def gauss_pulse(A0, d):
    return dict(pulse_type=_lib.PULSE_GAUSSIAN, A0=A0, t0=0.2, td=0.15, w0=3.0, phi=0.2, pol=(1.0, 0.0, 0.0), disp=d)


def oracle_J(O, tM, U, pulse, noise):
    n_up, n_dn = occupations()
    sysO = O.System(tM=tM, coulomb_type=1, U=np.diag(U), pulse_type=pulse["pulse_type"],
                    A0=pulse.get("A0", 0.0), t0=pulse.get("t0", 0.0), td=pulse.get("td", 1.0),
                    w0=pulse.get("w0", 0.0), phi=pulse.get("phi", 0.0), pol=pulse["pol"], knots=pulse.get("knots"),
                    disp=pulse["disp"])
    s = O.run(sysO, DT, TMAX, NSUB, n_up, n_dn, noise_arr=noise)
    return float(O.normalise(s, noise.shape[0])["NuNd"][:, -1].mean())


def adjoint(tM, U, pulse, engine="fp64", ck=3):
    eng = make_engine(tM, U, pulse)
    eng.set_params(DT, TMAX, NSUB, B, seed=SEED, engine=engine)
    noise = eng.sample_noise(0, B)
    res = eng.adjoint(ck)
    eng.close()
    return res, noise


def fd_check(O, tM, U, pulse, res, noise, pulse_key, h=1e-5):
    """Central differences of the oracle's J against every requested gradient entry."""
    checks = []  # (label, fd, adjoint)
    for i in (0, 3):
        Up, Um = U.copy(), U.copy()
        Up[i] += h
        Um[i] -= h
        checks.append((f"U{i}", (oracle_J(O, tM, Up, pulse, noise) - oracle_J(O, tM, Um, pulse, noise)) / (2 * h),
                       res["grad_u"][i]))
    for k, l in ((0, 1), (1, 0), (2, 3), (0, 0)):
        tp, tm = tM.copy(), tM.copy()
        tp[k, l] += h
        tm[k, l] -= h
        checks.append((f"t{k}{l}", (oracle_J(O, tp, U, pulse, noise) - oracle_J(O, tm, U, pulse, noise)) / (2 * h),
                       res["grad_t"][k, l]))
    base = pulse[pulse_key]
    for k in range(np.size(base)):
        pp, pm = dict(pulse), dict(pulse)
        if np.ndim(base):
            vp, vm = base.copy(), base.copy()
            vp[k] += h
            vm[k] -= h
        else:
            vp, vm = base + h, base - h
        pp[pulse_key], pm[pulse_key] = vp, vm
        checks.append((f"{pulse_key}{k}", (oracle_J(O, tM, U, pp, noise) - oracle_J(O, tM, U, pm, noise)) / (2 * h),
                       res["grad_pulse"][k]))
    scale = max(abs(c[1]) for c in checks)
    for name, fd, ad in checks:
        print(f"{name}: fd {fd:+.10e} adjoint {ad:+.10e} diff {abs(fd - ad) / scale:.2e}")
    for name, fd, ad in checks:
        assert abs(fd - ad) <= 1e-6 * scale, name


def test_objective_matches_engine_and_oracle(oracle_mod):
    tM, d = ring()
    U = np.full(L, 2.0)
    pulse = knot_pulse(KNOTS, d)
    res, noise = adjoint(tM, U, pulse)
    eng = make_engine(tM, U, pulse)
    eng.set_params(DT, TMAX, NSUB, B, seed=SEED, engine="fp64")
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS)
    eng.run()
    J_engine = float(eng.result("NuNd")[:, -1].mean())
    eng.close()
    J_oracle = oracle_J(oracle_mod, tM, U, pulse, noise)
    assert abs(res["J"] - J_engine) <= 1e-12 * abs(J_engine)
    assert abs(res["J"] - J_oracle) <= 1e-12 * abs(J_oracle)


def test_gradient_knot_pulse_vs_finite_differences(oracle_mod):
    tM, d = ring()
    U = np.full(L, 2.0)
    pulse = knot_pulse(KNOTS, d)
    res, noise = adjoint(tM, U, pulse)
    fd_check(oracle_mod, tM, U, pulse, res, noise, "knots")


def test_gradient_gaussian_pulse_vs_finite_differences(oracle_mod):
    tM, d = ring()
    U = np.full(L, 3.0)
    pulse = gauss_pulse(0.7, d)
    res, noise = adjoint(tM, U, pulse)
    fd_check(oracle_mod, tM, U, pulse, res, noise, "A0")


def test_checkpoint_spacing_is_bitwise_neutral():
    tM, d = ring()
    U = np.full(L, 2.0)
    pulse = knot_pulse(KNOTS, d)
    ref, _ = adjoint(tM, U, pulse, ck=1)
    for ck in (3, 7, 50):
        got, _ = adjoint(tM, U, pulse, ck=ck)
        assert got["J"] == ref["J"]
        for k in ("grad_u", "grad_t", "grad_pulse"):
            assert np.array_equal(got[k], ref[k]), (ck, k)


def test_fp32_adjoint_close_to_fp64():
    tM, d = ring()
    U = np.full(L, 2.0)
    pulse = knot_pulse(KNOTS, d)
    r64, _ = adjoint(tM, U, pulse, "fp64")
    r32, _ = adjoint(tM, U, pulse, "fp32")
    assert abs(r32["J"] - r64["J"]) <= 1e-5
    for k in ("grad_u", "grad_t", "grad_pulse"):
        scale = np.max(np.abs(r64[k]))
        assert np.max(np.abs(r32[k] - r64[k])) <= 1e-4 * scale, k


def test_no_pulse_and_unsupported_coulomb():
    tM, _ = ring()
    eng = _lib.Engine(tM, coulomb_type=_lib.COULOMB_ONSITE, U=2.0 * np.eye(L))
    eng.set_initial_state(*occupations())
    eng.set_params(DT, TMAX, NSUB, B, seed=SEED, engine="fp64")
    res = eng.adjoint(4)
    assert res["grad_pulse"] is None and np.isfinite(res["grad_u"]).all()
    eng.close()
    eng = _lib.Engine(tM, coulomb_type=_lib.COULOMB_GENERAL, U=2.0 * np.eye(L), V=0.5 * np.ones((L, L)))
    eng.set_initial_state(*occupations())
    eng.set_params(DT, TMAX, NSUB, B, seed=SEED, engine="fp64")
    with pytest.raises(_lib.FtwaError):
        eng.adjoint(4)
    eng.close()


def test_torch_autograd_function():
    """J.backward() through fermiphasespace_amd.autograd gives the adjoint's gradients, and a central
    difference of J in U_0 through the same function agrees to 1e-6."""
    torch = pytest.importorskip("torch")
    from fermiphasespace_amd import autograd as ag

    tM0, d = ring()
    n_up, n_dn = occupations()
    prob = ag.FtwaProblem(n_up=n_up, n_dn=n_dn, dt=DT, t_max=TMAX, n_sub=NSUB, n_traj=B, seed=SEED, engine="fp64",
                          checkpoint_every=3, pol=(1.0, 0.0, 0.0), disp=d)
    U = torch.full((L,), 2.0, dtype=torch.float64, requires_grad=True)
    tM = torch.tensor(tM0, requires_grad=True)
    knots = torch.tensor(KNOTS, requires_grad=True)
    J = ag.double_occupancy(U, tM, knots, prob)
    J.backward()
    ref, _ = adjoint(tM0, np.full(L, 2.0), knot_pulse(KNOTS, d))
    assert J.item() == ref["J"]
    assert np.array_equal(U.grad.numpy(), ref["grad_u"])
    assert np.array_equal(tM.grad.numpy(), ref["grad_t"])
    assert np.array_equal(knots.grad.numpy(), ref["grad_pulse"])
    h = 1e-5
    with torch.no_grad():
        Up, Um = U.detach().clone(), U.detach().clone()
        Up[0] += h
        Um[0] -= h
        fd = (ag.double_occupancy(Up, tM.detach(), knots.detach(), prob)
              - ag.double_occupancy(Um, tM.detach(), knots.detach(), prob)).item() / (2 * h)
    assert abs(fd - U.grad[0].item()) <= 1e-6 * abs(U.grad).max().item()


# ---- the 6x6 pulse workload (L = 36, 200 knots, checkpoints every 50 steps) ----------------------------
C5_DT, C5_TMAX, C5_B, C5_SEED, C5_CK = 0.01, 1.0, 64, 505, 50


def c5_engine(w, U=None, tM=None, knots=None):
    eng = _lib.Engine(w["tM"] if tM is None else tM, coulomb_type=_lib.COULOMB_ONSITE,
                      U=w["U"] if U is None else U, pulse_type=_lib.PULSE_KNOTS,
                      knots=w["knots"] if knots is None else knots, pol=w["pol"], disp=w["disp"])
    eng.set_initial_state(w["n_up"], w["n_dn"])
    return eng


def c5_oracle_J(O, w, noise, U=None, tM=None, knots=None):
    sysO = O.System(tM=w["tM"] if tM is None else tM, coulomb_type=1, U=w["U"] if U is None else U,
                    pulse_type=_lib.PULSE_KNOTS, pol=w["pol"], knots=w["knots"] if knots is None else knots,
                    disp=w["disp"])
    s = O.run(sysO, C5_DT, C5_TMAX, 1, w["n_up"], w["n_dn"], noise_arr=noise)
    return float(O.normalise(s, noise.shape[0])["NuNd"][:, -1].mean())


def test_adjoint_6x6_pulse_workload_vs_finite_differences(oracle_mod):
    """The adjoint workload's own system: 6x6 periodic square lattice (L = 36), t = -1, U = 4, charge-density
    wave, 200 knots, checkpoints every 50 steps over 100 steps (two segments), B = 64, dt = 0.01, t_max = 1.
    The FP64 adjoint against central differences (h = 1e-5) of the oracle's J on the same samples for the
    uniform U, the uniform hopping scale (d/d ln t), two bonds (x and y, both directions) and three knots:
    <= 1e-6 relative to the largest entry.  The FP32 adjoint agrees with the FP64 one to 1e-4 of the largest
    entry of each gradient."""
    O = oracle_mod
    w = lat.workload_adjoint6x6()
    L = w["L"]
    assert L == 36 and w["knots"].size == 200
    eng = c5_engine(w)
    eng.set_params(C5_DT, C5_TMAX, 1, C5_B, seed=C5_SEED, engine="fp64")
    noise = eng.sample_noise(0, C5_B)
    r64 = eng.adjoint(C5_CK)
    eng.set_params(C5_DT, C5_TMAX, 1, C5_B, seed=C5_SEED, engine="fp32")
    r32 = eng.adjoint(C5_CK)
    eng.close()
    J0 = c5_oracle_J(O, w, noise)
    assert abs(r64["J"] - J0) <= 1e-12 * abs(J0)

    h = 1e-5
    checks = []  # (label, central difference, adjoint)

    def fd(**kw_pm):
        plus = {k: v[0] for k, v in kw_pm.items()}
        minus = {k: v[1] for k, v in kw_pm.items()}
        return (c5_oracle_J(O, w, noise, **plus) - c5_oracle_J(O, w, noise, **minus)) / (2 * h)

    eye = np.eye(L)
    checks.append(("U", fd(U=(w["U"] + h * eye, w["U"] - h * eye)), r64["grad_u"].sum()))
    checks.append(("ln t", fd(tM=(w["tM"] * (1 + h), w["tM"] * (1 - h))), float((r64["grad_t"] * w["tM"]).sum())))
    for k, l in ((0, 1), (0, 6)):  # an x bond and a y bond of site 0
        assert w["tM"][k, l] != 0
        e = np.zeros((L, L))
        e[k, l] = e[l, k] = 1.0
        checks.append((f"bond{k}-{l}", fd(tM=(w["tM"] + h * e, w["tM"] - h * e)),
                       r64["grad_t"][k, l] + r64["grad_t"][l, k]))
    for q in (3, 101, 198):
        e = np.zeros(w["knots"].size)
        e[q] = 1.0
        checks.append((f"knot{q}", fd(knots=(w["knots"] + h * e, w["knots"] - h * e)), r64["grad_pulse"][q]))
    scale = max(abs(c[1]) for c in checks)
    for name, f, a in checks:
        print(f"{name}: fd {f:+.10e} adjoint {a:+.10e} diff {abs(f - a) / scale:.2e}", flush=True)
    for name, f, a in checks:
        assert abs(f - a) <= 1e-6 * scale, name

    assert abs(r32["J"] - r64["J"]) <= 1e-5
    for k in ("grad_u", "grad_t", "grad_pulse"):
        sc = np.max(np.abs(r64[k]))
        assert np.max(np.abs(r32[k] - r64[k])) <= 1e-4 * sc, k
