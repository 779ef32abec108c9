"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

1. P1 unit square n x n, E = 10 GPa, nu = 0.2, bottom u_y = 0, corner
   u_x = 0, top u_y = 1e-3 (uniaxial patch test).
2. P2 unit square n x n with circular inclusions (radius U[0.01, 0.04], gap
   0.002, 35 % area), matrix 12 GPa / 0.3, inclusions 60 GPa / 0.2, BCs as 1.
3. Tet4 unit cube n^3 (6-tet voxel split), E = 30 GPa, nu = 0.2, initial
   damage a Gaussian random field (correlation length 0.05) mapped to
   [0, 0.3], bottom clamped, top u_z = steps * 2e-5.
4. Periodic P1 RVEs with inclusions (radius U[0.02, 0.08], area fraction
   U[0.20, 0.45]), 12/60 GPa, three macro strains (exx, eyy, gxy) = 1e-3.

event_square: the event-driven damage case at mesh size n (P1 unit square,
matrix StiffnessAndFracture(10 GPa, 0.2, NonLocalMazars(1e-4, 10 GPa, 0.2,
Gf 100, radius)) around an elastic 40 GPa disc of radius 0.18 at the
centre, u_x = 0 left, u_y = 0 bottom, u_x = load right).
"""
import numpy as np

import amie_gpu as ag

MACRO3 = np.array([[1e-3, 0., 0.], [0., 1e-3, 0.], [0., 0., 1e-3]])


def uniaxial_bcs_2d(coords, top=1e-3):
    x, y = coords[:, 0], coords[:, 1]
    bottom = np.nonzero(np.isclose(y, 0.))[0]
    topn = np.nonzero(np.isclose(y, 1.))[0]
    corner = np.nonzero(np.isclose(x, 0.) & np.isclose(y, 0.))[0]
    dofs = np.concatenate([2 * bottom + 1, 2 * topn + 1, 2 * corner])
    vals = np.concatenate([np.zeros(len(bottom)), np.full(len(topn), top), np.zeros(len(corner))])
    return dofs.astype(np.int64), vals


def config1(n=128):
    coords, conn = ag.structured_mesh(ag.P1, n)
    C = ag.cauchy_green(10e9, 0.2)[None]
    dofs, vals = uniaxial_bcs_2d(coords)
    return ag.Problem(ag.P1, coords, conn, C, bc_dofs=dofs, bc_vals=vals)


def composite_cmats():
    return np.stack([ag.cauchy_green(12e9, 0.3), ag.cauchy_green(60e9, 0.2)])


def config2(n=400, seed=2, fraction=0.35):
    coords, conn = ag.structured_mesh(ag.P2, n)
    xyr, frac = ag.inclusions_2d(seed, fraction, 0.01, 0.04, 0.002)
    phase = ag.phase_by_centroid(ag.P2, conn, coords, xyr)
    dofs, vals = uniaxial_bcs_2d(coords)
    return ag.Problem(ag.P2, coords, conn, composite_cmats(), phase=phase, bc_dofs=dofs, bc_vals=vals)


def config3(n=64, seed=3, step=1, du=2e-5):
    coords, conn = ag.structured_mesh(ag.TET4, n)
    C = ag.cauchy_green(30e9, 0.2, dim=3)[None]
    d0 = ag.damage_field_3d(seed, n, 0.05, 0.0, 0.3)
    z = coords[:, 2]
    bottom = np.nonzero(np.isclose(z, 0.))[0]
    top = np.nonzero(np.isclose(z, 1.))[0]
    dofs = np.concatenate([3 * bottom, 3 * bottom + 1, 3 * bottom + 2, 3 * top + 2])
    vals = np.concatenate([np.zeros(3 * len(bottom)), np.full(len(top), step * du)])
    return ag.Problem(ag.TET4, coords, conn, C, damage=d0, bc_dofs=dofs.astype(np.int64), bc_vals=vals)


def rve(n, seed, fraction, rmin=0.02, rmax=0.08, gap=0.0, macro=MACRO3):
    coords, conn = ag.structured_mesh(ag.P1, n)
    xyr, frac = ag.inclusions_2d(seed, fraction, rmin, rmax, gap)
    phase = ag.phase_by_centroid(ag.P1, conn, coords, xyr)
    dofmap = ag.periodic_dofmap(n)
    # the periodic fluctuation is fixed at DOF node 0 (removes the translations)
    return ag.Problem(ag.P1, coords, conn, composite_cmats(), phase=phase, dofmap=dofmap, ndofnodes=n * n,
                      bc_dofs=np.array([0, 1], dtype=np.int64), bc_vals=np.zeros(2), macro=macro,
                      nrhs=macro.shape[0])


def config4(nrve=512, seed=4, nmin=32, nmax=256):
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=nrve)
    fr = rng.uniform(0.20, 0.45, size=nrve)
    seeds = rng.integers(1, 2 ** 62, size=nrve)
    return [rve(int(n), int(s), float(f)) for n, s, f in zip(ns, seeds, fr)]


def rve3d(n, seed, fraction=0.3, macro=None):
    """Periodic Tet4 RVE: phases by a seeded voxel field thresholded at the
    requested fraction; macro (nrhs x 6, nrhs 1 or 3) defaults to one load
    case exx = 1e-3."""
    coords, conn = ag.structured_mesh(ag.TET4, n)
    field = ag.damage_field_3d(seed, n, 0.08, 0.0, 1.0)
    thr = np.quantile(field, 1.0 - fraction)
    phase = (field >= thr).astype(np.int32)
    m = n + 1
    i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    dofmap = (((i % n) * n + (j % n)) * n + (k % n)).ravel().astype(np.int32)
    C = np.stack([ag.cauchy_green(12e9, 0.3, dim=3), ag.cauchy_green(60e9, 0.2, dim=3)])
    macro = np.array([[1e-3, 0, 0, 0, 0, 0]], dtype=np.float64) if macro is None else np.atleast_2d(macro)
    return ag.Problem(ag.TET4, coords, conn, C, phase=phase, dofmap=dofmap, ndofnodes=n ** 3,
                      bc_dofs=np.array([0, 1, 2], dtype=np.int64), bc_vals=np.zeros(3), macro=macro,
                      nrhs=macro.shape[0])


def config5(rate=20.0, duration=10.0, seed=5, n2d=(64, 512), n3d=(16, 48), ndamage=32, damage_keys=4):
    """Served solve stream: Poisson arrivals at `rate` req/s over `duration`
    seconds, one third each of 2D P1 RVEs (n ~ U{64..512}, 30 % inclusions,
    3 macro-strain load cases), 3D Tet4 RVEs (n ~ U{16..48}, 1 load case) and
    damage steps of the 32^3 damage case on `damage_keys` persistent states.
    Returns [(t_seconds, kind, problem, key)], key != 0 for damage steps."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    damage = {}
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        kind = int(rng.integers(0, 3))
        if kind == 0:
            n = int(rng.integers(n2d[0], n2d[1] + 1))
            out.append((t, "rve2d", rve(n, int(rng.integers(1, 2 ** 62)), 0.30), 0))
        elif kind == 1:
            n = int(rng.integers(n3d[0], n3d[1] + 1))
            out.append((t, "rve3d", rve3d(n, int(rng.integers(1, 2 ** 62))), 0))
        else:
            key = int(rng.integers(1, damage_keys + 1))
            if key not in damage:
                damage[key] = config3(ndamage, seed=300 + key)
            out.append((t, "damage", damage[key], key))
    return out


def event_square(n, radius=None):
    """The event-driven damage case on a structured P1 n x n mesh: (problem,
    right-edge DOFs, the criterion and NonLocalMazars tables, radius).  The
    characteristic radius defaults to 3 cells (3/n)."""
    coords, conn = ag.structured_mesh(ag.P1, n)
    phase = ag.phase_by_centroid(ag.P1, conn, coords, np.array([[0.5, 0.5, 0.18]]))
    cm = np.stack([ag.cauchy_green(1e10, 0.2), ag.cauchy_green(4e10, 0.2)])
    x, y = coords[:, 0], coords[:, 1]
    left = np.nonzero(np.isclose(x, 0.))[0]
    bottom = np.nonzero(np.isclose(y, 0.))[0]
    right = np.nonzero(np.isclose(x, 1.))[0]
    dofs = np.concatenate([2 * left, 2 * bottom + 1, 2 * right]).astype(np.int64)
    vals = np.zeros(len(dofs))
    prob = ag.Problem(ag.P1, coords, conn, cm, phase=phase, bc_dofs=dofs, bc_vals=vals)
    crit = np.array([1, 0], dtype=np.int32)
    mazars = np.array([[1e-4, 1e10, 0.2, 100.], [0., 0., 0., 1.]])
    return prob, dofs, len(left) + len(bottom), crit, mazars, (3. / n if radius is None else radius)
