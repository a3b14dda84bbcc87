"""fpsftwa_run_noalloc (no device allocation: caller-owned state and sums, the library's fixed scratch)
against fpsftwa_run, through the C ABI.

Bar: bitwise-equal ensemble sums and final states (the same samples, kernels and fixed-order sums; the
time chunking differs, which the kernels and the reduction are invariant to), on two calls in a row on
one handle, the second with fewer trajectories, for the FP32 and FP64 batched engines on the 8x8
periodic workload (65,536 trajectories first: several launches per run within the 256 MiB scratch).
Requests beyond the scratch or outside the entry point's scope fail with an error and run nothing."""
import numpy as np
import pytest

from fermiphasespace_amd import _lib
from fermiphasespace_amd import lattices as lat

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DT, TMAX, SEED = 0.005, 0.1, 202
L = 64
Q = 6 * L + 12


def engine():
    pos, box = lat.grid(8, 8)
    tM, U = lat.system_arrays(pos, box, True, t=-1.0, U=4.4)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, coulomb_type=_lib.COULOMB_ONSITE, U=U)
    eng.set_initial_state(occ, occ)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES | _lib.RES_WHOLE)
    return eng


def check_against_run(eng, kind, B, sums_buf):
    """One fpsftwa_run and one fpsftwa_run_noalloc of the same settings: bitwise-equal sums and states."""
    eng.set_params(DT, TMAX, 1, B, seed=SEED, engine=kind)
    eng.run()
    T = eng.result("sum_u").shape[1]
    ref = {k: eng.result(k) for k in ("sum_u", "sum_d", "sum2_u", "sum2_d", "sum_ud", "sum_e", "N1hu", "N1hd")}
    ref_states = eng.states()
    real = torch.float32 if kind == "fp32" else torch.float64
    state = torch.empty(B * 2 * L * L * 2, dtype=real, device="cuda")
    sums_buf.fill_(float("nan"))
    torch.cuda.synchronize()
    print(f"run_noalloc {kind} B={B} T={T}", flush=True)
    eng.run_noalloc(state.data_ptr(), state.numel() * state.element_size(), sums_buf.data_ptr(), sums_buf.numel())
    acc = sums_buf[: T * Q * 2].cpu().numpy().reshape(T, Q, 2)
    assert np.isfinite(acc).all()  # every entry written
    S = acc[..., 0] + acc[..., 1]  # (sum, compensation) pairs, as fpsftwa_get_result forms them
    for m, k in enumerate(("sum_u", "sum_d", "sum2_u", "sum2_d", "sum_ud")):
        assert np.array_equal(S[:, m:6 * L:6].T, ref[k]), (kind, B, k)
    assert np.array_equal(S[:, 6 * L:6 * L + 4].T, ref["sum_e"]), (kind, B)
    assert np.array_equal(S[:, 6 * L + 8] / B, ref["N1hu"][0]), (kind, B)
    assert np.array_equal(S[:, 6 * L + 11] / B, ref["N1hd"][1]), (kind, B)
    st = state.cpu().numpy().astype(np.float64).reshape(B, 2, L, L, 2)
    assert np.array_equal(st[..., 0] + 1j * st[..., 1], ref_states), (kind, B)
    del state


@pytest.mark.parametrize("kind", ["fp32", "fp64"])
def test_noalloc_matches_run_two_calls(kind):
    eng = engine()
    T = int(np.floor(TMAX / DT + 1))
    sums = torch.empty(T * Q * 2, dtype=torch.float64, device="cuda")
    check_against_run(eng, kind, 65536, sums)
    check_against_run(eng, kind, 4096, sums)  # the second call is smaller; the scratch keeps the first's data
    eng.close()


def test_noalloc_rejects_what_it_cannot_run():
    eng = engine()
    T = int(np.floor(TMAX / DT + 1))
    sums = torch.empty(T * Q * 2, dtype=torch.float64, device="cuda")
    small = torch.empty(16, dtype=torch.float64, device="cuda")
    # one time point of FP64 records for 300,000 trajectories (326 MB) exceeds the 256 MiB scratch
    eng.set_params(DT, TMAX, 1, 300000, seed=SEED, engine="fp64")
    with pytest.raises(_lib.FtwaError, match="scratch"):
        eng.run_noalloc(small.data_ptr(), 128, sums.data_ptr(), sums.numel())
    # buffers smaller than the run needs
    eng.set_params(DT, TMAX, 1, 64, seed=SEED, engine="fp32")
    with pytest.raises(_lib.FtwaError, match="state holds"):
        eng.run_noalloc(small.data_ptr(), 128, sums.data_ptr(), sums.numel())
    with pytest.raises(_lib.FtwaError, match="sums holds"):
        eng.run_noalloc(small.data_ptr(), 128, small.data_ptr(), 16)
    # outside the entry point's scope
    eng.set_params(DT, TMAX, 1, 64, seed=SEED, engine="fp64_plain")
    with pytest.raises(_lib.FtwaError):
        eng.run_noalloc(small.data_ptr(), 128, sums.data_ptr(), sums.numel())
    eng.close()
