"""CPU-only AddressSanitizer / UndefinedBehaviorSanitizer build of the host code that reads user input:
the JSON reader and writer, the system builder and the eigensolver (fermiphasespace_amd/csrc/json.cpp,
system.cpp, eigen.cpp), driven by tests/sanitize/host_check.cpp over the reference's sample inputs and
simulation files for every lattice generator, IC type, interaction table, pulse and core hole the front end
reads, plus truncated texts of each.  Any sanitizer report makes the driver exit non-zero."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fermiphasespace_amd import lattices as lat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fermiphasespace_amd", "csrc")


def simulation_files(tmp_path):
    files = []
    for f in ("grid_10.json", "grid_10.xyz", "sym_grid_10.json"):
        shutil.copy(os.path.join(ROOT, "tests", "golden", "reference_inputs", f), tmp_path / f)
    files += [str(tmp_path / "grid_10.json"), str(tmp_path / "sym_grid_10.json")]
    sim, pos = lat.workload_hubbard8x8(n_traj=4)
    files.append(lat.write_simulation(str(tmp_path), sim, pos))
    sim, pos = lat.workload_chain10(n_traj=4)
    files.append(lat.write_simulation(str(tmp_path), sim, pos))
    for name, (pos, box), ic in (
            ("honey", lat.honeycomb(4, 3), {"type": "tbGroundState", "nPartUp": 12, "nPartDown": 12}),
            ("kagome", lat.kagome(4, 4), {"type": "mfGroundState", "nPartUp": 24, "nPartDown": 20}),
            ("bcc", lat.bcc(3), {"type": "givenDiagonal", "subtype": "chargeWave"}),
            ("moire", lat.moire(3, 3), {"type": "givenDiagonal", "subtype": "spinWave"})):
        sim = lat.simulation(name, pos, box if name != "moire" else None, name != "moire", ic=ic,
                             coulomb={"type": "distanceInteraction", "distances": [0.0, 1.0],
                                      "energies": [4.0, 0.5]} if name == "bcc" else None,
                             tunneling={"type": "distanceInteraction", "distances": [1.0, 2.0],
                                        "energies": [-1.0, -0.2]} if name == "kagome" else None,
                             external={"type": "pulse", "A0": 0.5, "t0": 1.0, "td": 0.5, "w0": 2.0,
                                       "polarisation": [1.0, 0.0, 0.0]} if name == "honey" else None)
        files.append(lat.write_simulation(str(tmp_path), sim, pos))
    pos, _ = lat.chain(6)
    sim = lat.simulation("ch", pos, None, False, U=3.0)
    sim["system"]["coreHole"] = {"presence": True, "site": 1, "eps": -2.5, "X": -0.7, "Uc": 1.5, "G": 0.8}
    files.append(lat.write_simulation(str(tmp_path), sim, pos))
    bad = tmp_path / "bad.json"
    bad.write_text('{"method": "fTWA", "system": {"lattice": {"positionsFile": "missing.xyz"}, "IC": [1, 2,')
    files.append(str(bad))
    return files


def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_check")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
           os.path.join(CSRC, "json.cpp"), os.path.join(CSRC, "system.cpp"), os.path.join(CSRC, "eigen.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    files = simulation_files(tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, *files], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert f"{len(files)} files, {len(files) - 1} parsed, {len(files) - 1} systems" in out, out[-2000:]
    assert "0 failures" in out and "bad.json: does not parse" in out, out[-2000:]
