"""Lattice generators and simulation-file writers for the fTWA workloads.

Writes the reference's input format: an .xyz positions file (site count, comment, "name x y z",
read by src/FHLattice.cpp:646-710) and a simulation .json with "method", "parameters", "sim_name",
"system" {lattice, tunnelingInteraction, coulombInteraction, externalPotential, IC} and "results"
(src/FHSystem.cpp:123-326, src/FHfTWASolver.cpp:49-60, src/FHSolver.cpp:34-45).  The periodic box goes in
"latticeMatrix" whose columns are the lattice vectors (src/FHLattice.cpp:133-145, :516-539).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

# ---- site positions ----------------------------------------------------------------------


def chain(n: int, a: float = 1.0):
    pos = np.zeros((n, 3))
    pos[:, 0] = a * np.arange(n)
    return pos, np.diag([n * a, 1.0, 1.0])


def grid(nx: int, ny: int, nz: int = 1, a: float = 1.0):
    """Simple cubic/square grid, site index x + nx*(y + ny*z) (as FHLattice(x, y, z), src/FHLattice.cpp:66-97)."""
    pos = np.array([[x, y, z] for z in range(nz) for y in range(ny) for x in range(nx)], dtype=float) * a
    return pos, np.diag([nx * a, ny * a, max(nz, 1) * a])


def honeycomb(nx: int, ny: int, a: float = 1.0):
    """Honeycomb with nearest-neighbour distance a: 2 sites per rectangular-ish cell, periodic box."""
    a1 = np.array([1.5 * a, math.sqrt(3) / 2 * a, 0.0])
    a2 = np.array([1.5 * a, -math.sqrt(3) / 2 * a, 0.0])
    basis = [np.zeros(3), np.array([a, 0.0, 0.0])]
    pos = [i * a1 + j * a2 + b for j in range(ny) for i in range(nx) for b in basis]
    box = np.column_stack([nx * a1, ny * a2, [0, 0, 1.0]])
    return np.array(pos), box


def kagome(nx: int, ny: int, a: float = 1.0):
    """Kagome with nearest-neighbour distance a (3 sites per triangular cell of side 2a)."""
    a1 = np.array([2 * a, 0.0, 0.0])
    a2 = np.array([a, math.sqrt(3) * a, 0.0])
    basis = [np.zeros(3), a1 / 2, a2 / 2]
    pos = [i * a1 + j * a2 + b for j in range(ny) for i in range(nx) for b in basis]
    box = np.column_stack([nx * a1, ny * a2, [0, 0, 1.0]])
    return np.array(pos), box


def bcc(n: int, a: float | None = None):
    """Body-centred cubic, n^3 cubic cells of side 2/sqrt(3) so the nearest-neighbour distance is 1."""
    c = 2.0 / math.sqrt(3.0) if a is None else a
    pos = []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                pos.append([x * c, y * c, z * c])
                pos.append([(x + 0.5) * c, (y + 0.5) * c, (z + 0.5) * c])
    return np.array(pos), np.diag([n * c] * 3)


def graphene(nx: int, ny: int, a: float = 1.0):
    """Graphene patch of nx x ny two-site cells, carbon distance a (makeGraphene, src/FHLattice.cpp:1093-1130):
    cell (i, j) holds (0, -a/2) and (0, a/2) shifted by i*(sqrt3 a, 0) + j*(sqrt3 a/2, 3a/2)."""
    s3 = math.sqrt(3.0)
    pos = []
    for j in range(ny):
        for i in range(nx):
            x = s3 * a * i + s3 / 2 * a * j
            y = 1.5 * a * j
            pos += [[x, y - 0.5 * a, 0.0], [x, y + 0.5 * a, 0.0]]
    box = np.array([[s3 * a * nx, s3 / 2 * a * ny, 0.0], [0.0, 1.5 * a * ny, 0.0], [0.0, 0.0, 10.0 * a]])
    return np.array(pos), box


def moire(nx: int, ny: int, a: float = 1.0, b: float = 2.35, theta: float = 5.0):
    """Twisted bilayer (makeMoireGraphene, src/FHLattice.cpp:1272-1307): two graphene patches centred on
    the middle of their cell box, the second rotated by theta degrees about z and lifted by b.  Open
    boundaries (the twist breaks the cell's periodicity)."""
    p0, box = graphene(nx, ny, a)
    mid = (box[:, 0] + box[:, 1]) / 2.0
    layers = []
    for layer in range(2):
        c, s_ = math.cos(layer * math.radians(theta)), math.sin(layer * math.radians(theta))
        R = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
        q = (p0 - mid) @ R.T
        q[:, 2] = layer * b
        layers.append(q)
    return np.vstack(layers), box


def workload_moire196(A0=0.5, td=2.0, w0=1.5, t0=5.0, U=3.0, a=1.0):
    """Long-range twisted bilayer, 2 x 7x7 graphene cells (L = 196): t_ij = -exp(-(r_ij - a)/0.3a) for
    r_ij <= 4a, V_ij = U/(1 + r_ij/a) (on-site U), Gaussian pulse, half filling (spinWave by site index)."""
    pos, _ = moire(7, 7, a)
    L = len(pos)
    d = pos[:, None, :] - pos[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    off = ~np.eye(L, dtype=bool)
    tM = np.where(off & (r <= 4.0 * a + 1e-9), -np.exp(-(r - a) / (0.3 * a)), 0.0)
    V = np.where(off, U / (1.0 + r / a), 0.0)
    Umat = np.diag(np.full(L, U)) + V
    n_up = np.array([1.0 if i % 2 == 0 else 0.0 for i in range(L)])
    pulse = dict(pulse_type=1, A0=A0, t0=t0, td=td, w0=w0, phi=0.0, pol=(2 ** -0.5, 2 ** -0.5, 0.0), disp=d)
    return {"tM": tM, "U": Umat, "V": V, "n_up": n_up, "n_dn": 1.0 - n_up, "pulse": pulse, "pos": pos, "L": L}


def workload_mixed(seed=303, t=-1.0, U=4.0):
    """Uneven batch of four lattices: honeycomb L=24 (z=3), kagome 4x4 L=48 (z=4), 10x10 square L=100
    (z=4), bcc 3x3x3 L=54 (z=8), all periodic, with site potentials Uniform[-1, 1] drawn from the seed
    (counter stream 1000*k + 1 + i for site i of system k), spinWave initial state (up on even sites,
    down on odd, src/FHSystem.cpp:559-570) and trajectory counts 32,768 / 16,384 / 4,096 / 8,192."""
    specs = [("honeycomb24", honeycomb(4, 3), 32768), ("kagome48", kagome(4, 4), 16384),
             ("square100", grid(10, 10), 4096), ("bcc54", bcc(3), 8192)]
    out = []
    for k, (name, (pos, box), B) in enumerate(specs):
        L = len(pos)
        tM, Um = system_arrays(pos, box, True, t=t, U=U)
        pot = np.array([-1.0 + 2.0 * uniform_draw(seed, 1000 * k + 1 + i) for i in range(L)])
        n_up = np.array([1.0 if i % 2 == 0 else 0.0 for i in range(L)])
        out.append({"name": name, "L": L, "tM": tM, "U": Um, "pot": pot, "n_up": n_up, "n_dn": 1.0 - n_up,
                    "n_traj": B, "z": int((tM != 0).sum(axis=1).max())})
    return out


def min_image_distances(pos: np.ndarray, box: np.ndarray | None):
    """Minimum-image distances over +-1 cells (src/FHLattice.cpp:516-539)."""
    d = pos[:, None, :] - pos[None, :, :]
    best = np.linalg.norm(d, axis=-1)
    if box is not None:
        for x in (-1, 0, 1):
            for y in (-1, 0, 1):
                for z in (-1, 0, 1):
                    sh = x * box[:, 0] + y * box[:, 1] + z * box[:, 2]
                    best = np.minimum(best, np.linalg.norm(d + sh, axis=-1))
    return best


# ---- counter-based draws (Philox4x32-10, Salmon et al. SC'11) ------------------------------
_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def _philox(ctr, key):
    c = [int(x) & 0xFFFFFFFF for x in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF, ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF,
             p0 & 0xFFFFFFFF]
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return c


def uniform_draw(seed: int, index: int) -> float:
    """The index-th configuration uniform in [0, 1) of a seed: counter (index, 2, ~0, ~0), which the
    trajectory noise (spin field 0 or 1) never uses."""
    x = _philox([index, 2, 0xFFFFFFFF, 0xFFFFFFFF], [seed & 0xFFFFFFFF, seed >> 32])
    return ((x[0] >> 5) * 67108864.0 + (x[1] >> 6)) / 9007199254740992.0


# ---- simulation files ------------------------------------------------------------------------


def write_xyz(path: str, pos: np.ndarray, name: str = "C") -> None:
    with open(path, "w") as f:
        f.write(f"{len(pos)}\n\n")
        for p in pos:
            f.write(f"{name}\t{p[0]:.17g}\t{p[1]:.17g}\t{p[2]:.17g}\n")


def simulation(name: str, pos, box, periodic: bool, *, t=-1.0, U=4.0, ic=None, dt=0.01, t_max=1.0, n_sub=1,
               n_traj=10, results=None, external=None, tunneling=None, coulomb=None, mi355x=None) -> dict:
    lattice = {"positionsFile": name + ".xyz", "periodic": bool(periodic)}
    if box is not None:
        lattice["latticeMatrix"] = [[float(v) for v in row] for row in np.asarray(box)]
    sysj = {
        "lattice": lattice,
        "tunnelingInteraction": tunneling or {"type": "nearestNeighbors", "neighborDistance": 1.0,
                                              "neighborInteraction": float(t)},
        "coulombInteraction": coulomb or {"type": "onSite", "onSiteInteraction": float(U)},
        "IC": ic or {"type": "givenDiagonal", "subtype": "spinWave"},
    }
    if external is not None:
        sysj["externalPotential"] = external
    params = {"dt": dt, "t_max": t_max, "stepsParTimePoints": n_sub, "n_traj": n_traj, "int_order": 4}
    if mi355x:
        params["mi355x"] = mi355x
    return {"method": "fTWA", "sim_name": name, "parameters": params, "system": sysj,
            "results": results or {"diagDMatrix": True, "onSiteDoublons": True, "energies": True}}


def write_simulation(folder: str, sim: dict, pos) -> str:
    os.makedirs(folder, exist_ok=True)
    write_xyz(os.path.join(folder, sim["system"]["lattice"]["positionsFile"]), np.asarray(pos))
    path = os.path.join(folder, sim["sim_name"] + ".json")
    with open(path, "w") as f:
        json.dump(sim, f, indent="\t")
    return path


def checkerboard(nx: int, ny: int):
    """Half-filled charge-density wave on a square grid: both spins on (x+y) even."""
    occ = [1 if (x + y) % 2 == 0 else 0 for y in range(ny) for x in range(nx)]
    return {"type": "givenDiagonal", "subtype": "givenInJson", "IC_NuDiag": occ, "IC_NdDiag": occ}


# ---- the named workloads ------------------------------------------------------------------


def workload_chain10(n_traj=1000, seed=101, engine="fp64"):
    """10-site open chain (grid_10 positions), t=-1, U=4, Neel, RK4 dt=0.01, 10 sub-steps, t_max=10."""
    pos, _ = chain(10)
    sim = simulation("chain10", pos, None, False, t=-1.0, U=4.0, dt=0.01, t_max=10.0, n_sub=10, n_traj=n_traj,
                     mi355x={"seed": seed, "engine": engine},
                     results={"diagDMatrix": True, "onSiteDoublons": True, "energies": True})
    return sim, pos


def workload_hubbard8x8(n_traj=65536, seed=202, engine="fp32", t_max=20.0, dt=0.005):
    """8x8 periodic square lattice, t=-1, U ~ Uniform[2,8] drawn from the seed, half-filled CDW."""
    pos, box = grid(8, 8)
    U = 2.0 + 6.0 * uniform_draw(seed, 0)
    sim = simulation("hubbard8x8", pos, box, True, t=-1.0, U=U, ic=checkerboard(8, 8), dt=dt, t_max=t_max,
                     n_traj=n_traj, mi355x={"seed": seed, "engine": engine},
                     results={"diagDMatrix": True, "onSiteDoublons": True, "energies": True})
    return sim, pos


def workload_adjoint6x6(n_knots=200, amp=0.3):
    """6x6 periodic square lattice (L = 36), t=-1, U=4, charge-density wave up / complement down, knot pulse
    a(t) = 0.3 sin(3 s) on n_knots equally spaced knots over [0, t_max], polarisation (1, 1, 0)/sqrt 2,
    minimum-image displacements on the torus."""
    pos, box = grid(6, 6)
    tM, Um = system_arrays(pos, box, True, t=-1.0, U=4.0)
    d = pos[:, None, :] - pos[None, :, :]
    d = d - np.round(d / 6.0) * 6.0
    d[..., 2] = 0.0
    n_up = np.array(checkerboard(6, 6)["IC_NuDiag"], dtype=float)
    knots = amp * np.sin(np.linspace(0.0, 3.0, n_knots))
    return {"tM": tM, "U": Um, "n_up": n_up, "n_dn": 1.0 - n_up, "knots": knots, "disp": d,
            "pol": (2 ** -0.5, 2 ** -0.5, 0.0), "L": 36}


def system_arrays(pos, box, periodic: bool, t=-1.0, U=4.0, nn=1.0, same_point=1e-4):
    """tM (nearest neighbours at distance nn), on-site U, and min-image displacements, as arrays."""
    D = min_image_distances(pos, box if periodic else None)
    L = len(pos)
    tM = np.zeros((L, L))
    iu = np.triu_indices(L, 1)
    mask = (D[iu] - nn) ** 2 < same_point
    tM[iu[0][mask], iu[1][mask]] = t
    tM = tM + tM.T
    Umat = np.diag(np.full(L, float(U)))
    return tM, Umat
