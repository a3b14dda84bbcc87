"""Pins of the CPU oracle (oracle/ftwa_oracle.c) against what the mathematics fixes.

The reference ships no golden vectors for this path and cannot be built here (it needs Armadillo),
so each oracle function is checked against known answers, closed forms and invariants:

  philox4x32_10   Random123 known-answer vectors
  n_time_points   floor(t_max/dt + 1) in FP64, incl. the 0.3/0.1 floor quirk (src/FHSolver.cpp:100)
  noise/sampling  closed-form moments of the fTWA Wigner sample (src/FHfTWASolver.cpp:187-198, :342-362)
  derivative+RK4  atomic limit (tM = 0): each rho_ij is multiplied by the RK4 stability polynomial of
                  -i(h_j - h_i) h per step, with h_sigma = U (n_other) -- pins the Hartree term, its spin,
                  the +1/2 and the sign of the commutator;
                  free fermions: the exact RK4 propagator in the eigenbasis of Delta, the exact
                  e^{i Delta t} rho e^{-i Delta t} to RK4 accuracy, and order-4 convergence
  pulse           A = 0 doubles the hopping (extPot added to tM); a static Peierls dimer closed form
  observables     t = 0 known answers; particle number, energy, Hermiticity and spectrum conservation
"""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy import stats


def rk4_poly(z):
    return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24


def initial_states(L, n_up, n_dn, noise):
    """rho = diag(n - 1/2) + S.W with W_ij = noise_p (i<j), W_ji = conj, S_ij = sqrt((ni+nj-2ninj)/2)."""
    B = noise.shape[0]
    out = np.zeros((B, 2, L, L), dtype=complex)
    iu = np.triu_indices(L, 1)
    for s, n in enumerate((n_up, n_dn)):
        S = np.sqrt((n[:, None] + n[None, :] - 2 * n[:, None] * n[None, :]) / 2)
        for b in range(B):
            w = noise[b, s, :, 0] + 1j * noise[b, s, :, 1]
            M = np.zeros((L, L), dtype=complex)
            M[iu] = w * S[iu]
            M = M + M.conj().T
            M[np.diag_indices(L)] = n - 0.5
            out[b, s] = M
    return out


def chain_tM(L, t=-1.0):
    tM = np.zeros((L, L))
    for i in range(L - 1):
        tM[i, i + 1] = tM[i + 1, i] = t
    return tM


def test_philox_known_answers(oracle_mod):
    O = oracle_mod
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(O.philox4x32_10(ctr, key)) == want


def test_time_grid(oracle_mod):
    O = oracle_mod
    assert O.n_time_points(0.01, 10) == 1001
    assert O.n_time_points(0.005, 20) == 4001
    assert O.n_time_points(0.1, 0.3) == 3  # 0.3/0.1 = 2.9999999999999996 in FP64
    assert O.n_time_points(0.01, 2) == 201
    assert O.n_time_points(0.05, 0) == 1


def test_noise_moments(oracle_mod):
    w = oracle_mod.noise(6, 0, 4000, 7)  # 4000 x 2 x 15 complex values
    re, im = w[..., 0].ravel(), w[..., 1].ravel()
    assert abs(np.mean(re)) < 0.01 and abs(np.mean(im)) < 0.01
    assert abs(np.var(re) - 0.5) < 0.01 and abs(np.var(im) - 0.5) < 0.01
    assert abs(np.corrcoef(re, im)[0, 1]) < 0.01
    # each component is N(0, 1/2)
    assert stats.kstest(re / np.sqrt(0.5), "norm").pvalue > 1e-3
    assert stats.kstest(im / np.sqrt(0.5), "norm").pvalue > 1e-3
    # the two spins and different trajectories are independent streams
    assert abs(np.corrcoef(w[:, 0, :, 0].ravel(), w[:, 1, :, 0].ravel())[0, 1]) < 0.01
    # counter-based: a slice reproduces the same values
    np.testing.assert_array_equal(oracle_mod.noise(6, 100, 5, 7), w[100:105])


def test_sampled_state_moments(oracle_mod):
    O = oracle_mod
    L = 4
    n_up = np.array([1.0, 0.0, 1.0, 0.0])
    n_dn = np.array([1.0, 1.0, 0.0, 0.0])
    B = 20000
    s = O.System(tM=np.zeros((L, L)), coulomb_type=1, U=np.eye(L))
    r = O.run(s, 0.1, 0.0, 1, n_up, n_dn, n_traj=B, seed=11, keep_states=True)
    rho = r.final_states  # T = 1: the sampled initial states
    for sg, n in enumerate((n_up, n_dn)):
        R = rho[:, sg]
        np.testing.assert_allclose(R, np.conj(np.transpose(R, (0, 2, 1))), atol=0)  # Hermitian
        np.testing.assert_array_equal(np.diagonal(R, axis1=1, axis2=2).real, np.broadcast_to(n - 0.5, (B, L)))
        want = (n[:, None] + n[None, :] - 2 * n[:, None] * n[None, :]) / 2
        got = np.mean(np.abs(R) ** 2, axis=0)
        off = ~np.eye(L, dtype=bool)
        np.testing.assert_allclose(got[off], want[off], atol=0.03)
        assert np.all(np.abs(R[:, (want == 0) & off]) == 0)  # no noise between equal occupations
        np.testing.assert_allclose(np.mean(R, axis=0)[off], 0, atol=0.02)
    # up and down noise are independent
    c = np.corrcoef(rho[:, 0, 0, 1].real, rho[:, 1, 0, 2].real)[0, 1]
    assert abs(c) < 0.03


def test_atomic_limit_closed_form(oracle_mod):
    """tM = 0: diag(rho) is constant and rho_ij picks up R(-i(h_j - h_i) dt) per step."""
    O = oracle_mod
    L, B = 4, 3
    U = np.diag([1.0, 2.5, 3.0, 4.25])
    n_up = np.array([1.0, 0.0, 1.0, 0.0])
    n_dn = np.array([0.0, 1.0, 1.0, 0.0])
    noise = O.noise(L, 0, B, 5)
    s = O.System(tM=np.zeros((L, L)), coulomb_type=1, U=U)
    dt, t_max = 0.1, 2.0
    r = O.run(s, dt, t_max, 1, n_up, n_dn, noise_arr=noise, keep_states=True)
    nsteps = r.T - 1
    rho0 = initial_states(L, n_up, n_dn, noise)
    h = [np.diag(U) * n_dn, np.diag(U) * n_up]  # h_up = U (rho_dn + 1/2) = U n_dn
    for sg in range(2):
        z = -1j * (h[sg][None, :] - h[sg][:, None]) * dt
        want = rho0[:, sg] * rk4_poly(z)[None] ** nsteps
        np.testing.assert_allclose(r.final_states[:, sg], want, rtol=0, atol=1e-13)


def _free_expected(tM, pot, rho0, dt, nsteps):
    D = tM + np.diag(pot)
    lam, V = np.linalg.eigh(D)
    rt = V.conj().T @ rho0 @ V
    z = 1j * (lam[:, None] - lam[None, :]) * dt  # d rho/dt = i[D, rho]
    return V @ (rt * rk4_poly(z) ** nsteps) @ V.conj().T, V @ (rt * np.exp(z * nsteps)) @ V.conj().T


def test_free_fermions_closed_form(oracle_mod):
    O = oracle_mod
    L, B = 5, 2
    rng = np.random.default_rng(3)
    tM = chain_tM(L)
    tM[0, 3] = tM[3, 0] = 0.3
    pot = rng.uniform(-1, 1, L)
    n_up = np.array([1.0, 0.0, 1.0, 0.0, 1.0])
    n_dn = 1 - n_up
    noise = O.noise(L, 0, B, 9)
    s = O.System(tM=tM, coulomb_type=0, pot=pot)
    dt, t_max = 0.05, 1.0
    r = O.run(s, dt, t_max, 1, n_up, n_dn, noise_arr=noise, keep_states=True)
    rho0 = initial_states(L, n_up, n_dn, noise)
    for b in range(B):
        for sg in range(2):
            exact_rk4, exact = _free_expected(tM, pot, rho0[b, sg], dt, r.T - 1)
            np.testing.assert_allclose(r.final_states[b, sg], exact_rk4, atol=1e-13)
            np.testing.assert_allclose(r.final_states[b, sg], exact, atol=5e-4)  # RK4 truncation at dt=0.05


def test_rk4_order_four(oracle_mod):
    O = oracle_mod
    L = 4
    tM = chain_tM(L)
    n_up = np.array([1.0, 0.0, 1.0, 0.0])
    noise = O.noise(L, 0, 1, 1)
    s = O.System(tM=tM, coulomb_type=0)
    rho0 = initial_states(L, n_up, 1 - n_up, noise)[0, 0]
    errs = []
    for nsub in (2, 4, 8):
        r = O.run(s, 0.4, 0.4, nsub, n_up, 1 - n_up, noise_arr=noise, keep_states=True)
        _, exact = _free_expected(tM, np.zeros(L), rho0, 0.4 / nsub, nsub)
        errs.append(np.max(np.abs(r.final_states[0, 0] - exact)))
    for a, b in zip(errs, errs[1:]):
        assert 13 < a / b < 19, errs


def test_pulse_zero_amplitude_doubles_hopping(oracle_mod):
    O = oracle_mod
    L, B = 5, 2
    tM = chain_tM(L)
    pos = np.arange(L, dtype=float)[:, None] * np.array([1.0, 0.0, 0.0])
    disp = pos[:, None, :] - pos[None, :, :]
    n_up = np.array([1.0, 0.0, 1.0, 0.0, 1.0])
    noise = O.noise(L, 0, B, 4)
    U = 3.0 * np.eye(L)
    with_pulse = O.System(tM=tM, coulomb_type=1, U=U, pulse_type=1, A0=0.0, t0=1.0, td=1.0, w0=2.0,
                          pol=(1.0, 0.0, 0.0), disp=disp)
    doubled = O.System(tM=2 * tM, coulomb_type=1, U=U)
    a = O.run(with_pulse, 0.05, 0.5, 1, n_up, 1 - n_up, noise_arr=noise, keep_states=True)
    b = O.run(doubled, 0.05, 0.5, 1, n_up, 1 - n_up, noise_arr=noise, keep_states=True)
    np.testing.assert_array_equal(a.final_states, b.final_states)
    # with A = 0 the pulse energy is the hopping energy (extPot = tM)
    np.testing.assert_allclose(a.sum_e[3], a.sum_e[0], rtol=1e-12, atol=1e-12)


def test_static_peierls_dimer(oracle_mod):
    """Two sites, U = 0, constant amplitude a along x: Delta_01 = t (1 + e^{i a (x0 - x1)})."""
    O = oracle_mod
    t, a = -1.0, 0.7
    tM = np.array([[0.0, t], [t, 0.0]])
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    disp = pos[:, None, :] - pos[None, :, :]
    s = O.System(tM=tM, coulomb_type=0, pulse_type=2, knots=np.array([a]), pol=(1.0, 0.0, 0.0), disp=disp)
    n_up = np.array([1.0, 0.0])
    noise = O.noise(2, 0, 1, 2)
    dt = 0.05
    r = O.run(s, dt, 1.0, 1, n_up, 1 - n_up, noise_arr=noise, keep_states=True)
    D = np.array([[0, t * (1 + np.exp(-1j * a))], [t * (1 + np.exp(1j * a)), 0]])
    rho0 = initial_states(2, n_up, 1 - n_up, noise)[0, 0]
    lam, V = np.linalg.eigh(D)
    z = 1j * (lam[:, None] - lam[None, :]) * dt
    want = V @ ((V.conj().T @ rho0 @ V) * rk4_poly(z) ** (r.T - 1)) @ V.conj().T
    np.testing.assert_allclose(r.final_states[0, 0], want, atol=1e-13)
    assert abs(O.pulse_amplitude(s, 0.3, 1.0) - a) == 0


def test_pulse_amplitude_forms(oracle_mod):
    O = oracle_mod
    g = O.System(tM=np.zeros((1, 1)), pulse_type=1, A0=0.5, t0=5.0, td=2.0, w0=1.5, phi=0.2)
    for t in (0.0, 3.0, 5.0, 7.5):
        want = 0.5 * np.exp(-(((t - 5.0) / 2.0) ** 2) / 2) * np.cos(1.5 * (t - 5.0) + 0.2)
        assert abs(O.pulse_amplitude(g, t, 10.0) - want) < 1e-15
    k = O.System(tM=np.zeros((1, 1)), pulse_type=2, knots=np.array([0.0, 1.0, -1.0]))
    assert O.pulse_amplitude(k, 2.5, 10.0) == pytest.approx(0.5)
    assert O.pulse_amplitude(k, 7.5, 10.0) == pytest.approx(0.0)
    assert O.pulse_amplitude(k, 10.0, 10.0) == -1.0


def test_time_zero_observables(oracle_mod):
    O = oracle_mod
    L, B = 6, 50
    n_up = np.array([1.0, 0, 1, 0, 1, 1])
    n_dn = np.array([1.0, 1, 0, 0, 1, 0])
    s = O.System(tM=chain_tM(L), coulomb_type=1, U=4 * np.eye(L))
    r = O.run(s, 0.01, 0.05, 2, n_up, n_dn, n_traj=B, seed=3)
    out = O.normalise(r, B)
    np.testing.assert_array_equal(out["Niiu"][:, 0], n_up)
    np.testing.assert_array_equal(out["Niid"][:, 0], n_dn)
    np.testing.assert_array_equal(out["NuNd"][:, 0], n_up * n_dn)
    np.testing.assert_allclose(out["N2iiu"][:, 0], n_up, atol=1e-15)  # (n - 1/2)^2 + (n - 1/2) + 1/4 = n^2
    # E_U(0) = U sum_i n_up n_dn; E_tH(0) = 0 (noise is off-diagonal with zero mean per trajectory pair? no:
    # per trajectory E_tH = sum_ij N_ij t_ij = 2 t Re sum_bonds rho_ij, nonzero; only its mean vanishes)
    np.testing.assert_allclose(out["E"][1, 0], 4 * np.sum(n_up * n_dn), rtol=1e-14)


def test_conservation_laws(oracle_mod):
    O = oracle_mod
    L, B = 6, 4
    rng = np.random.default_rng(8)
    tM = chain_tM(L)
    Ug = rng.uniform(0.5, 1.5, (L, L))
    Ug = (Ug + Ug.T) / 2
    Vg = rng.uniform(0.0, 0.5, (L, L))
    Vg = (Vg + Vg.T) / 2
    np.fill_diagonal(Vg, 0)
    pot = rng.uniform(-1, 1, L)
    n_up = np.array([1.0, 0, 1, 0, 1, 0])
    noise = O.noise(L, 0, B, 12)
    for s in (O.System(tM=tM, coulomb_type=1, U=4 * np.eye(L)),
              O.System(tM=tM, coulomb_type=2, U=Ug, V=Vg, pot=pot)):
        r = O.run(s, 0.01, 2.0, 1, n_up, 1 - n_up, noise_arr=noise, keep_states=True)
        Nup = r.sum_u.sum(axis=0)
        np.testing.assert_allclose(Nup, Nup[0], atol=1e-11)  # particle number per spin
        E = r.sum_e.sum(axis=0)  # tH + U + V + pe: the mean-field energy of every trajectory
        np.testing.assert_allclose(E, E[0], atol=1e-7 * max(1.0, abs(E[0])))
        rho0 = initial_states(L, n_up, 1 - n_up, noise)
        for b in range(B):
            for sg in range(2):
                R = r.final_states[b, sg]
                np.testing.assert_allclose(R, R.conj().T, atol=1e-13)
                # isospectral flow, up to the RK4 truncation error (~h^4 t)
                np.testing.assert_allclose(np.linalg.eigvalsh(R), np.linalg.eigvalsh(rho0[b, sg]), atol=2e-6)


def test_seeded_equals_explicit_noise(oracle_mod):
    O = oracle_mod
    L, B = 5, 6
    s = O.System(tM=chain_tM(L), coulomb_type=1, U=2 * np.eye(L))
    n_up = np.array([1.0, 0, 1, 0, 1])
    a = O.run(s, 0.02, 0.2, 1, n_up, 1 - n_up, n_traj=B, seed=77, nthreads=2)
    b = O.run(s, 0.02, 0.2, 1, n_up, 1 - n_up, noise_arr=O.noise(L, 0, B, 77), nthreads=3)
    np.testing.assert_allclose(a.sum_u, b.sum_u, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(a.sum_e, b.sum_e, rtol=1e-13, atol=1e-13)


def test_ground_state_closed_forms(oracle_mod):
    """Open chain: eig(tM) has the closed form v_k(i) = sqrt(2/(L+1)) sin(pi k (i+1)/(L+1)) with eigenvalue
    2 t cos(pi k/(L+1)) (ascending for t < 0 with k = 1..L).  Mean field at half filling on the chain keeps
    n = 1/2 everywhere (particle-hole symmetry), so the second iterate reproduces the first: two iterations,
    residue 0, and the same eigenvectors as tight binding.  With U = 0 mean field is tight binding."""
    O = oracle_mod
    L, t = 9, -1.0
    tM = t * (np.eye(L, k=1) + np.eye(L, k=-1))
    V, Vd, it, res, E, Ed = O.ground_state(tM, None, 5, 4, False, energies=True)
    k = np.arange(1, L + 1)
    closed = np.sqrt(2.0 / (L + 1)) * np.sin(np.pi * np.outer(np.arange(1, L + 1), k) / (L + 1))
    np.testing.assert_allclose(np.abs(V), np.abs(closed), atol=1e-13)
    np.testing.assert_allclose(np.diag(V.T @ tM @ V), 2 * t * np.cos(np.pi * k / (L + 1)), atol=1e-13)
    # the echoed energies ("energyVector") are the same closed form, ascending, equal for both spins
    np.testing.assert_allclose(E, 2 * t * np.cos(np.pi * k / (L + 1)), atol=1e-13)
    assert np.array_equal(E, Ed)
    assert np.array_equal(V, Vd) and it == 0
    big = np.argmax(np.abs(V) >= np.abs(V).max(axis=0) / (1 + 1e-8), axis=0)  # first of the (tied) largest
    assert np.all(V[big, np.arange(L)] > 0)
    L = 10
    tM = t * (np.eye(L, k=1) + np.eye(L, k=-1))
    Vu, Vd, it, res, Eu, Ed = O.ground_state(tM, 4.0 * np.eye(L), 5, 5, True, energies=True)
    assert it == 2 and res < 1e-12
    np.testing.assert_allclose(Vu, O.ground_state(tM, None, 5, 5, False)[0], atol=1e-12)
    # n = 1/2 everywhere: the mean-field levels are the tight-binding ones shifted by U/2
    np.testing.assert_allclose(Eu, 2 * t * np.cos(np.pi * np.arange(1, L + 1) / (L + 1)) + 2.0, atol=1e-12)
    np.testing.assert_allclose(Ed, Eu, atol=1e-13)
    np.testing.assert_allclose((Vu[:, :5] ** 2).sum(axis=1), 0.5, atol=1e-13)
    Vu0, Vd0, _, _ = O.ground_state(tM + np.diag(np.linspace(-0.3, 0.2, L)), np.zeros((L, L)), 6, 3, True)
    np.testing.assert_allclose(Vu0, O.ground_state(tM + np.diag(np.linspace(-0.3, 0.2, L)), None, 6, 3, False)[0],
                               atol=1e-12)


def test_ground_state_mean_field_fixed_point(oracle_mod):
    """A converged mean-field basis diagonalises tM + diag(U n_other) built from its own densities (to the
    1e-4 L convergence threshold), and the spins differ when n_up != n_dn."""
    O = oracle_mod
    L = 8
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1))
    U = 4.0 * np.eye(L)
    Vu, Vd, it, res = O.ground_state(tM, U, 5, 3, True)
    assert 1 < it < 100 and res < 1e-4 * L
    nu, nd = (Vu[:, :5] ** 2).sum(axis=1), (Vd[:, :3] ** 2).sum(axis=1)
    for V, other in ((Vu, nd), (Vd, nu)):
        D = V.T @ (tM + np.diag(U @ other)) @ V
        assert np.abs(D - np.diag(np.diag(D))).max() < 1e-2
        assert np.all(np.diff(np.diag(D)) > -1e-2)
    assert np.abs(nu - nd).max() > 0.1


def test_basis_occupations(oracle_mod):
    """tbStates / mfStates sums diag(V^dagger rho V) (src/FHfTWASolver.cpp:385-394).
    (1) V = identity gives the site densities.  (2) Free fermions started in their own eigenbasis keep
    every eigenstate occupation: in that basis RK4 multiplies element (k, l) by a polynomial in
    (E_k - E_l) h, which is 1 on the diagonal, so diag(V^T rho V) = n - 1/2 per trajectory at every time
    point (a transposed V fails this).  (3) A complex unitary basis against the final states."""
    O = oracle_mod
    L, B = 6, 5
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1)) + np.diag(np.linspace(-0.3, 0.4, L))
    U = 2.0 * np.eye(L)
    n_up = np.array([1.0, 0, 1, 1, 0, 0])
    n_dn = np.array([0.0, 1, 0, 1, 1, 0])
    rng = np.random.default_rng(5)
    noise = rng.normal(size=(B, 2, L * (L - 1) // 2, 2)) / np.sqrt(2)
    eye = np.broadcast_to(np.eye(L), (1, 2, L, L))
    r = O.run(O.System(tM=tM, coulomb_type=1, U=U), 0.05, 0.5, 2, n_up, n_dn, noise_arr=noise, bases=eye)
    np.testing.assert_allclose(r.sum_basis[0, 0], r.sum_u, atol=1e-13)
    np.testing.assert_allclose(r.sum_basis[0, 1], r.sum_d, atol=1e-13)

    V = O.ground_state(tM, None, 3, 3, False)[0]
    occ = (np.arange(L) < 3).astype(float)
    r = O.run(O.System(tM=tM, coulomb_type=0), 0.05, 1.0, 2, occ, occ, noise_arr=noise, rot_up=V, rot_dn=V,
              bases=np.stack([np.stack([V, V])]))
    np.testing.assert_allclose(r.sum_basis[0, 0], B * (occ[:, None] - 0.5) * np.ones((1, r.T)), atol=1e-12)
    np.testing.assert_allclose(r.sum_basis[0, 1], B * (occ[:, None] - 0.5) * np.ones((1, r.T)), atol=1e-12)
    assert np.abs(r.sum_u[:, -1] - r.sum_u[:, 0]).max() > 1e-3  # the site densities do move
    Vt = np.stack([np.stack([V.T, V.T])])
    rt = O.run(O.System(tM=tM, coulomb_type=0), 0.05, 1.0, 2, occ, occ, noise_arr=noise, rot_up=V, rot_dn=V, bases=Vt)
    assert np.abs(rt.sum_basis[0, 0, :, -1] - B * (occ - 0.5)).max() > 1e-3

    Q = [np.linalg.qr(rng.normal(size=(L, L)) + 1j * rng.normal(size=(L, L)))[0] for _ in range(2)]
    r = O.run(O.System(tM=tM, coulomb_type=1, U=U), 0.05, 0.3, 1, n_up, n_dn, noise_arr=noise, keep_states=True,
              bases=np.stack([np.stack(Q)]))
    for sg in range(2):
        want = sum(np.real(np.diag(Q[sg].conj().T @ r.final_states[b, sg] @ Q[sg])) for b in range(B))
        np.testing.assert_allclose(r.sum_basis[0, sg, :, -1], want, atol=1e-12)
    M = O.normalise(r, B)["M"]
    np.testing.assert_allclose(M, r.sum_basis / B + 0.5)


@pytest.fixture(scope="module")
def ed_mod():
    from oracle import ed
    return ed


def test_ed_known_answers(ed_mod):
    """The ED reference itself: (1) free fermions follow the one-particle propagator,
    n_i(t) = sum_k |(e^{-i tM t})_ik|^2 n_k(0); (2) the two-site Hubbard dimer from |up, dn>: the singlet
    half couples to the symmetric doublon state with amplitude 2t, so the total double occupancy is
    (1/2)(4t^2/W^2) sin^2(W t), W = sqrt(U^2/4 + 4t^2)."""
    ed = ed_mod
    L = 6
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1)) + np.diag([0.1, -0.2, 0.0, 0.3, 0.0, -0.1])
    occ = np.array([1, 1, 0, 1, 0, 0])
    r = ed.evolve_product_state(L, tM, np.zeros(L), occ, 1 - occ, 1.5, 7)
    for k, t in enumerate(r["t"]):
        P = np.abs(expm(-1j * tM * t)) ** 2
        np.testing.assert_allclose(r["n_up"][:, k], P @ occ, atol=1e-10)
        np.testing.assert_allclose(r["n_dn"][:, k], P @ (1 - occ), atol=1e-10)
    t_hop, U = 0.7, 3.0
    r = ed.evolve_product_state(2, -t_hop * np.array([[0.0, 1.0], [1.0, 0.0]]), [U, U], [1, 0], [0, 1], 2.0, 9)
    W = np.sqrt(U ** 2 / 4 + 4 * t_hop ** 2)
    want = 0.5 * (4 * t_hop ** 2 / W ** 2) * np.sin(W * r["t"]) ** 2
    np.testing.assert_allclose(r["nund"].sum(axis=0), want, atol=1e-10)


def test_ftwa_against_exact_diagonalisation(oracle_mod, ed_mod):
    """Physics check of the algorithm on the 10-site open chain (t = 1, Neel state): fTWA is exact for
    U = 0, so the ensemble densities match ED at every time within the sampling error; at U = 4 it is an
    approximation that matches ED at short times (t <= 0.5) and departs later (0.07 at t = 1)."""
    O, ed = oracle_mod, ed_mod
    L, B, dt = 10, 1024, 0.01
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1))
    nu = np.array([1.0, 0] * 5)
    nd = 1 - nu
    for U, t_ok in ((0.0, 2.0), (4.0, 0.5)):
        r = ed.evolve_product_state(L, tM, U * np.ones(L), nu, nd, t_ok, int(round(t_ok / 0.25)) + 1)
        s = O.run(O.System(tM=tM, coulomb_type=1, U=U * np.eye(L)), dt, t_ok, 1, nu, nd, n_traj=B, seed=7,
                  nthreads=8)
        n = O.normalise(s, B)
        idx = np.round(r["t"] / dt).astype(int)
        for spin, su, s2 in (("n_up", s.sum_u, s.sum2_u), ("n_dn", s.sum_d, s.sum2_d)):
            mean = su[:, idx] / B
            err = np.sqrt(np.maximum(s2[:, idx] / B - mean ** 2, 0) / B)
            got = n["Niiu" if spin == "n_up" else "Niid"][:, idx]
            assert np.all(np.abs(got - r[spin]) <= 5 * err + 1e-9), (U, spin, np.abs(got - r[spin]).max())


def test_hartree_fock_closed_form_and_conservation(oracle_mod):
    """Hartree-Fock (FHHFSolver::derivative, src/FHHFSolver.cpp:575-608): for one particle the Fock term
    cancels the Hartree self-interaction exactly, so with V alone the density follows the free propagator
    |(e^{-i tM t})_i0|^2 (without the Fock term it does not); for several particles TDHF conserves the
    particle number and the HF energy E_tH + E_U + E_V(with exchange) + E_pe."""
    O = oracle_mod
    from scipy.linalg import expm as _expm
    L = 5
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1))
    dist = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :]).astype(float)
    V = np.where(dist > 0, 1.5 / np.maximum(dist, 1.0), 0.0)
    nu = np.array([1.0, 0, 0, 0, 0])
    noise = np.zeros((1, 2, L * (L - 1) // 2, 2))
    t_end = 2.0
    want = np.abs(_expm(-1j * tM * t_end)[:, 0]) ** 2
    s = O.run(O.System(tM=tM, coulomb_type=2, U=np.zeros((L, L)), V=V, fock=True), 0.01, t_end, 1, nu, 0 * nu,
              noise_arr=noise)
    np.testing.assert_allclose(s.sum_u[:, -1] + 0.5, want, atol=1e-7)
    s = O.run(O.System(tM=tM, coulomb_type=2, U=np.zeros((L, L)), V=V, fock=False), 0.01, t_end, 1, nu, 0 * nu,
              noise_arr=noise)
    assert np.abs(s.sum_u[:, -1] + 0.5 - want).max() > 1e-2
    Um = 2.0 * np.eye(L) + V
    n_up = np.array([1.0, 1, 0, 1, 0])
    n_dn = np.array([0.0, 1, 1, 0, 1])
    s = O.run(O.System(tM=tM, coulomb_type=2, U=Um, V=V, pot=np.linspace(-0.5, 0.5, L), fock=True), 0.01, 3.0, 2,
              n_up, n_dn, noise_arr=noise)
    E = s.sum_e.sum(axis=0)
    assert np.abs(E - E[0]).max() < 1e-7 * max(1.0, abs(E[0]))
    np.testing.assert_allclose(s.sum_u.sum(axis=0) + L / 2, n_up.sum(), atol=1e-10)
    s_h = O.run(O.System(tM=tM, coulomb_type=2, U=Um, V=V, pot=np.linspace(-0.5, 0.5, L), fock=False), 0.01, 3.0, 2,
                n_up, n_dn, noise_arr=noise)
    assert np.abs(s_h.sum_e[2, -1] - s.sum_e[2, -1]) > 1e-3  # exchange once coherences have built up



# ---- GPSR (FHGPSRSolver) ---------------------------------------------------------------------------

def test_gpsr_free_fermions_and_initial_state(oracle_mod):
    """Without an interaction GPSR has no diffusion and its drift is the free commutator equation, so the
    normally ordered rho (no 1/2 shift) follows e^{i Delta t} diag(n) e^{-i Delta t}; the tbGroundState-type
    start R diag(n) R^dagger of the tM eigenbasis is stationary and its densities are sum_occ |V_ik|^2."""
    O = oracle_mod
    L = 5
    rng = np.random.default_rng(4)
    tM = chain_tM(L)
    tM[1, 4] = tM[4, 1] = -0.4
    pot = rng.uniform(-1, 1, L)
    n_up = np.array([1.0, 0.0, 1.0, 1.0, 0.0])
    n_dn = np.array([0.0, 1.0, 1.0, 0.0, 0.0])
    dt, t_max = 0.05, 1.0
    for ct in (0, 1):  # null Coulomb, and on-site with U(0,0) = 0 (diffusion constant sqrt(i U0) = 0)
        s = O.System(tM=tM, coulomb_type=ct, U=np.zeros((L, L)), pot=pot, gpsr=True, gpsr_seed=3)
        r = O.run(s, dt, t_max, 1, n_up, n_dn, n_traj=2, seed=3, keep_states=True)
        for sg, n in enumerate((n_up, n_dn)):
            exact_rk4, exact = _free_expected(tM, pot, np.diag(n).astype(complex), dt, r.T - 1)
            for b in range(2):
                np.testing.assert_allclose(r.final_states[b, sg], exact_rk4, atol=1e-13)
            np.testing.assert_allclose(r.final_states[0, sg], exact, atol=5e-4)
        np.testing.assert_allclose(r.sum_u[:, 0] / 2, n_up, atol=0)  # Niiu = rho_ii, no 1/2
    lam, V = np.linalg.eigh(tM)
    R = V.astype(complex)
    occ = np.array([1.0, 1.0, 0.0, 0.0, 0.0])
    s = O.System(tM=tM, coulomb_type=0, gpsr=True)
    r = O.run(s, 0.05, 1.0, 1, occ, occ, n_traj=1, seed=1, rot_up=R, rot_dn=R)
    want = (V[:, :2] ** 2).sum(axis=1)
    np.testing.assert_allclose(r.sum_u, np.repeat(want[:, None], r.T, axis=1), atol=1e-12)
    np.testing.assert_allclose(r.sum_e[0], 2 * lam[:2].sum(), atol=1e-12)  # E_tH = Re tr(rho tM), both spins


def _quarter_diffusion_matrix(ru, rd, U0):
    """A of the reference's diffusion matrix D = [[0, A], [A^T, 0]] as FHGPSRSolver::quarterDiffusionMatrix
    (src/FHGPSRSolver.cpp:337-383) builds it: line index l*L + k is element (k, l) of the column-major rho."""
    L = ru.shape[0]
    I = np.eye(L)
    B1u = np.zeros((L * L, L), complex)
    B2u = np.zeros((L * L, L), complex)
    B1d = np.zeros((L * L, L), complex)
    B2d = np.zeros((L * L, L), complex)
    for l in range(L):
        for k in range(L):
            for p in range(L):
                B1u[l * L + k, p] = ru[p, l] * (I - ru)[k, p]
                B2u[l * L + k, p] = 1j * (I - ru)[p, l] * ru[k, p]
                B1d[l * L + k, p] = rd[p, l] * (I - rd)[k, p]
                B2d[l * L + k, p] = 1j * (I - rd)[p, l] * rd[k, p]
    return 1j * U0 * (B1u @ B1d.T + B2u @ B2d.T)


def test_gpsr_diffusion_second_moments(oracle_mod):
    """The oracle's analytic diffusion increment (FHGPSRSolver::step_Diffusion) has the second moments of the
    diffusion matrix the reference factorises on its rsvd path: E[d rho_up (x) d rho_dn] = A h and
    E[d rho_s (x) d rho_s] = 0 (the zero diagonal blocks of D), for non-Hermitian rho and either sign of U.
    The increment is linear in the 4L real noise variables (Re, Im of xi_1 and xi_2, each of variance h/2),
    so the moments follow exactly from its values at unit noises."""
    O = oracle_mod
    L, h = 3, 0.01
    rng = np.random.default_rng(8)
    ru = rng.normal(size=(L, L)) + 1j * rng.normal(size=(L, L))
    rd = rng.normal(size=(L, L)) + 1j * rng.normal(size=(L, L))
    for U0 in (2.5, -1.5):
        s = O.System(tM=chain_tM(L), coulomb_type=1, U=U0 * np.eye(L), gpsr=True)
        Ju = np.zeros((L * L, 4 * L), complex)  # rows: column-major element index l*L + k
        Jd = np.zeros((L * L, 4 * L), complex)
        for r in range(4 * L):
            xi = np.zeros(2 * L, complex)
            k, part = divmod(r, 2)
            xi[k] = 1.0 if part == 0 else 1j
            du, dd = O.gpsr_diffusion(s, ru, rd, xi.reshape(2, L))
            Ju[:, r] = du.T.reshape(-1)
            Jd[:, r] = dd.T.reshape(-1)
        A = _quarter_diffusion_matrix(ru, rd, U0)
        np.testing.assert_allclose((h / 2) * Ju @ Jd.T, A * h, atol=1e-12 * np.abs(A).max())
        np.testing.assert_allclose((h / 2) * Ju @ Ju.T, 0, atol=1e-12 * np.abs(A).max())
        np.testing.assert_allclose((h / 2) * Jd @ Jd.T, 0, atol=1e-12 * np.abs(A).max())
    s = O.System(tM=chain_tM(L), coulomb_type=2, U=np.eye(L), V=np.eye(L), gpsr=True)
    du, dd = O.gpsr_diffusion(s, ru, rd, np.ones((2, L), complex))
    assert np.abs(du).max() == 0 and np.abs(dd).max() == 0  # diffusion only for an on-site interaction (:248)


def test_gpsr_noise_stream(oracle_mod):
    """The diffusion noise: Box-Muller moments (E xi = 0, E|xi|^2 = 1, Re and Im independent with variance 1/2,
    E xi^2 = 0) and a stream apart from the initial-state noise of the same key."""
    O = oracle_mod
    L = 64
    x = np.concatenate([O.gpsr_noise(L, b, st, 5).ravel() for b in range(20) for st in range(25)])
    n = x.size
    assert abs(x.mean()) < 5 / np.sqrt(n)
    assert abs(np.mean(np.abs(x) ** 2) - 1) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(x ** 2)) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(x.real * x.imag)) < 5 / np.sqrt(2 * n)
    assert stats.kstest(x.real * np.sqrt(2), "norm").pvalue > 1e-3
    w = O.noise(L, 0, 1, 5)[0, :, :, 0].ravel()
    assert not np.intersect1d(np.round(w, 12), np.round(x.real, 12)).size


def test_gpsr_against_exact_diagonalisation(oracle_mod, ed_mod):
    """Physics check: GPSR is an exact representation, so its ensemble means follow exact diagonalisation
    until the trajectories' known instability (4-site chain, U = 1, Neel start, to t = 0.5) -- the densities
    within the sampling error and <n_up n_dn> within 1.5e-3 -- where the drift alone (mean field, no
    diffusion) is off by 3e-3 or more in <n_up n_dn> and by several standard errors in the densities."""
    O, ed = oracle_mod, ed_mod
    L, U, B, dt, t_end = 4, 1.0, 100000, 0.01, 0.5
    tM = chain_tM(L)
    nu = np.array([1.0, 0, 1, 0])
    nd = 1 - nu
    r = ed.evolve_product_state(L, tM, U * np.ones(L), nu, nd, t_end, 6)
    idx = np.round(r["t"] / dt).astype(int)
    s = O.run(O.System(tM=tM, coulomb_type=1, U=U * np.eye(L), gpsr=True, gpsr_seed=11), dt, t_end, 1, nu, nd,
              n_traj=B, seed=11, nthreads=8)
    n = O.normalise(s, B, gpsr=True)
    mean = s.sum_u[:, idx] / B
    err = np.sqrt(np.maximum(s.sum2_u[:, idx] / B - mean ** 2, 0) / B)  # |rho_ii|^2 bounds (Re rho_ii)^2
    assert np.all(np.abs(n["Niiu"][:, idx] - r["n_up"]) <= 4 * err + 1e-9)
    assert np.all(np.abs(n["NuNd"][:, idx] - r["nund"]) <= 1.5e-3)
    mf = O.normalise(O.run(O.System(tM=tM, coulomb_type=2, U=U * np.eye(L), V=np.zeros((L, L)), gpsr=True), dt,
                           t_end, 1, nu, nd, n_traj=1, seed=1), 1, gpsr=True)
    assert np.abs(mf["NuNd"][:, idx[-1]] - r["nund"][:, -1]).max() > 3e-3
    assert np.abs(mf["Niiu"][:, idx[-1]] - r["n_up"][:, -1]).max() > 3 * err[:, -1].max()


def test_gpsr_divergent_trajectories_become_nan(oracle_mod):
    """FHGPSRSolver::step (src/FHGPSRSolver.cpp:611-615): a trajectory holding a NaN or Inf is set to NaN
    everywhere, so its final state is either finite or all NaN, and the ensemble sums turn NaN from then on."""
    O = oracle_mod
    L = 4
    s = O.System(tM=chain_tM(L), coulomb_type=1, U=2.0 * np.eye(L), gpsr=True, gpsr_seed=2)
    nu = np.array([1.0, 0, 1, 0])
    r = O.run(s, 0.02, 1.0, 1, nu, 1 - nu, n_traj=200, seed=2, keep_states=True)
    fin = np.isfinite(r.final_states).reshape(200, -1)
    assert np.all(fin.all(axis=1) | (~fin).all(axis=1))
    assert (~fin).all(axis=1).any() and fin.all(axis=1).any()
    assert np.isnan(r.sum_u[:, -1]).all() and np.isfinite(r.sum_u[:, 0]).all()
