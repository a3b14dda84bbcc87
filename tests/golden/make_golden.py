"""Writes tests/golden/*.npz: inputs and oracle outputs for the GPU parity tests.

The oracle (oracle/ftwa_oracle.c) is pinned by tests/test_oracle.py against closed forms; these
fixtures freeze its outputs on small systems so the GPU tests can check the engine without building
or loading the oracle.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def chain(L, t=-1.0, periodic=False):
    tM = np.zeros((L, L))
    for i in range(L - 1):
        tM[i, i + 1] = tM[i + 1, i] = t
    if periodic:
        tM[0, L - 1] = tM[L - 1, 0] = t
    return tM


def fixture(name, system, n_up, n_dn, dt, t_max, n_sub, B, seed, **extra):
    L = system.L
    noise = O.noise(L, 0, B, seed)
    r = O.run(system, dt, t_max, n_sub, n_up, n_dn, noise_arr=noise, keep_states=True)
    arrays = dict(tM=system.tM, U=system.U if system.U is not None else np.zeros((L, L)),
                  V=system.V if system.V is not None else np.zeros((L, L)),
                  pot=system.pot if system.pot is not None else np.zeros(L),
                  coulomb_type=system.coulomb_type, n_up=n_up, n_dn=n_dn, dt=dt, t_max=t_max, n_sub=n_sub, B=B,
                  seed=seed, noise=noise, sum_u=r.sum_u, sum_d=r.sum_d, sum2_u=r.sum2_u, sum_ud=r.sum_ud,
                  sum_e=r.sum_e, sum_corner=r.sum_corner, final_states=r.final_states, **extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", name, "T =", r.T)


def main():
    L = 6
    n_up = np.array([1.0, 0, 1, 0, 1, 0])
    fixture("chain6_onsite", O.System(tM=chain(L), coulomb_type=1, U=4.0 * np.eye(L)), n_up, 1 - n_up,
            0.02, 0.4, 2, 8, 101)
    rng = np.random.default_rng(2024)
    Ug = rng.uniform(0.2, 1.0, (L, L))
    Ug = (Ug + Ug.T) / 2
    Vg = rng.uniform(0.0, 0.4, (L, L))
    Vg = (Vg + Vg.T) / 2
    np.fill_diagonal(Vg, 0)
    pot = rng.uniform(-1, 1, L)
    fixture("ring6_general", O.System(tM=chain(L, periodic=True), coulomb_type=2, U=Ug, V=Vg, pot=pot),
            n_up, np.array([0.0, 1, 1, 0, 0, 1]), 0.02, 0.4, 1, 6, 7)


if __name__ == "__main__":
    main()
