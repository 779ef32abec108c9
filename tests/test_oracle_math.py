"""Pins the CPU oracle where no compiled-reference golden vector reaches, by
mathematics that fixes the answer independently of the oracle's own code:
closed-form element stiffness (constant-strain triangle / tetrahedron, exact
integration of the quadratic triangle), rigid-body null spaces, brute-force
pattern and dense assembly, closed-form Dirichlet elimination, library
linear algebra (numpy, scipy direct solves, eigvalsh), exact linear
displacement fields, the exact plane-stress patch solution, volume-weighted
homogeneous stress, periodic condensation as P^T K P and the homogeneous RVE
that must reproduce C.  Each check is chosen so that a dropped term, a wrong
sign or index, or a transposed operand in the oracle fails it.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import amie_gpu as ag
import configs
import oracle as orc

P1, P2, TET4 = 1, 2, 3


def rng(seed=0):
    return np.random.default_rng(seed)


def cst_stiffness(X, C):
    """Constant-strain triangle: K = A B^T C B (engineering shear)."""
    (x1, y1), (x2, y2), (x3, y3) = X
    det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    b = np.array([y2 - y3, y3 - y1, y1 - y2]) / det
    c = np.array([x3 - x2, x1 - x3, x2 - x1]) / det
    B = np.zeros((3, 6))
    B[0, 0::2] = b
    B[1, 1::2] = c
    B[2, 0::2] = c
    B[2, 1::2] = b
    return 0.5 * abs(det) * B.T @ C @ B


def tet_stiffness(X, C):
    """Linear tetrahedron: K = V B^T C B, Voigt [xx, yy, zz, yz, xz, xy]."""
    M = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(M)[1:, :]          # column i = grad N_i
    B = np.zeros((6, 12))
    for i in range(4):
        gx, gy, gz = G[:, i]
        B[:, 3 * i:3 * i + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz], [0, gz, gy], [gz, 0, gx], [gy, gx, 0]]
    return abs(np.linalg.det(M)) / 6. * B.T @ C @ B


def p2_stiffness(X, C):
    """Quadratic triangle, nodes [v0, m01, v1, m12, v2, m20]: the integrand
    B^T C B is quadratic on a straight-sided element, so the edge-midpoint
    rule (exact for degree 2) integrates it exactly."""
    V = X[[0, 2, 4]]
    T = np.array([[1, 1, 1], [V[0, 0], V[1, 0], V[2, 0]], [V[0, 1], V[1, 1], V[2, 1]]])
    area = 0.5 * abs(np.linalg.det(T))
    Ti = np.linalg.inv(T)               # L = Ti @ [1, x, y]
    dL = Ti[:, 1:]                      # dL_i/dx, dL_i/dy
    K = np.zeros((12, 12))
    for bary in ([0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]):
        L = np.array(bary)
        g = np.zeros((6, 2))
        for slot, i in ((0, 0), (2, 1), (4, 2)):
            g[slot] = (4 * L[i] - 1) * dL[i]
        for slot, (i, j) in ((1, (0, 1)), (3, (1, 2)), (5, (2, 0))):
            g[slot] = 4 * (L[i] * dL[j] + L[j] * dL[i])
        B = np.zeros((3, 12))
        B[0, 0::2] = g[:, 0]
        B[1, 1::2] = g[:, 1]
        B[2, 0::2] = g[:, 1]
        B[2, 1::2] = g[:, 0]
        K += area / 3. * B.T @ C @ B
    return K


def rigid_modes(X, dim):
    n = X.shape[0]
    if dim == 2:
        modes = [np.tile([1., 0.], n), np.tile([0., 1.], n), np.column_stack([-X[:, 1], X[:, 0]]).ravel()]
    else:
        modes = [np.tile(e, n) for e in np.eye(3)]
        x, y, z = X.T
        modes += [np.column_stack([0 * x, -z, y]).ravel(), np.column_stack([z, 0 * x, -x]).ravel(),
                  np.column_stack([-y, x, 0 * x]).ravel()]
    return np.array(modes)


def oriented(etype, X):
    return X[orc.orient(etype, X)]


# ------------------------------------------------------------ element level
@pytest.mark.parametrize("seed", range(4))
def test_p1_element_matrix_closed_form(seed):
    X = oriented(P1, rng(seed).uniform(0, 1, (3, 2)))
    C = orc.cauchy_green(rng(seed).uniform(1e9, 9e10), rng(seed + 9).uniform(0.05, 0.45))
    K = orc.element_matrix(P1, X, C)
    np.testing.assert_allclose(K, cst_stiffness(X, C), rtol=1e-12, atol=1e-12 * np.abs(K).max())


@pytest.mark.parametrize("seed", range(4))
def test_tet_element_matrix_closed_form(seed):
    X = oriented(TET4, rng(seed).uniform(0, 1, (4, 3)))
    C = orc.cauchy_green(rng(seed).uniform(1e9, 9e10), rng(seed + 9).uniform(0.05, 0.45), dim=3)
    K = orc.element_matrix(TET4, X, C)
    np.testing.assert_allclose(K, tet_stiffness(X, C), rtol=1e-11, atol=1e-11 * np.abs(K).max())


@pytest.mark.parametrize("seed", range(4))
def test_p2_element_matrix_exact_integral(seed):
    V = rng(seed).uniform(0, 1, (3, 2))
    X = np.array([V[0], (V[0] + V[1]) / 2, V[1], (V[1] + V[2]) / 2, V[2], (V[2] + V[0]) / 2])
    X = oriented(P2, X)
    C = orc.cauchy_green(3e10, 0.25)
    K = orc.element_matrix(P2, X, C)
    np.testing.assert_allclose(K, p2_stiffness(X, C), rtol=1e-10, atol=1e-10 * np.abs(K).max())


@pytest.mark.parametrize("etype", [P1, P2, TET4])
def test_element_matrix_rigid_null_space_and_spd(etype):
    dim = 3 if etype == TET4 else 2
    r = rng(etype)
    if etype == P2:
        V = r.uniform(0, 1, (3, 2))
        X = np.array([V[0], (V[0] + V[1]) / 2, V[1], (V[1] + V[2]) / 2, V[2], (V[2] + V[0]) / 2])
    else:
        X = r.uniform(0, 1, (4 if dim == 3 else 3, dim))
    X = oriented(etype, X)
    K = orc.element_matrix(etype, X, orc.cauchy_green(2e10, 0.3, dim=dim))
    R = rigid_modes(X, dim)
    assert np.abs(K @ R.T).max() <= 1e-12 * np.abs(K).max() * np.abs(R).max()
    np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())
    ev = np.linalg.eigvalsh(K)
    nrig = 3 if dim == 2 else 6
    assert np.all(ev[:nrig] > -1e-10 * ev[-1]) and np.all(ev[nrig:] > 1e-8 * ev[-1])


def test_orientation_is_counter_clockwise_for_triangles():
    X = np.array([[0., 0.], [0., 1.], [1., 0.]])   # clockwise
    for etype, pts in ((P1, X), (P2, np.array([X[0], (X[0] + X[1]) / 2, X[1], (X[1] + X[2]) / 2, X[2],
                                               (X[2] + X[0]) / 2]))):
        Y = oriented(etype, pts)
        v = Y[[0, 2, 4]] if etype == P2 else Y
        area2 = (v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[2, 0] - v[0, 0]) * (v[1, 1] - v[0, 1])
        assert area2 > 0
        if etype == P2:   # mid-nodes stay on their edges
            for m, (i, j) in ((1, (0, 2)), (3, (2, 4)), (5, (4, 0))):
                np.testing.assert_array_equal(Y[m], (Y[i] + Y[j]) / 2)


def test_orientation_of_tetrahedra_is_consistent():
    X = np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    signs = set()
    for Y in (X, X[[1, 0, 2, 3]]):
        Z = oriented(TET4, Y)
        signs.add(np.sign(np.linalg.det(Z[1:] - Z[0])))
    assert len(signs) == 1


# ------------------------------------------------------------ global level
def tiny_mesh(etype, n, seed=0):
    coords, conn = ag.structured_mesh(etype, n)
    return coords, conn


@pytest.mark.parametrize("etype,n", [(P1, 3), (P2, 2), (TET4, 2)])
def test_pattern_brute_force(etype, n):
    coords, conn = tiny_mesh(etype, n)
    nn = coords.shape[0]
    pairs = set((i, i) for i in range(nn))
    for e in conn:
        pairs.update(itertools.product(e.tolist(), e.tolist()))
    rs, ci = orc.pattern(etype, conn, nn)
    exp_rs = np.bincount([p[0] for p in pairs], minlength=nn)
    np.testing.assert_array_equal(rs, exp_rs)
    np.testing.assert_array_equal(ci, [p[1] for p in sorted(pairs)])


@pytest.mark.parametrize("etype,n", [(P1, 3), (P2, 2), (TET4, 2)])
def test_assembly_dense_brute_force_and_rigid_modes(etype, n):
    coords, conn = tiny_mesh(etype, n)
    dim = coords.shape[1]
    C = orc.cauchy_green(1e10, 0.2, dim=dim)[None]
    A = orc.assemble(etype, conn, coords, C)
    nn = coords.shape[0]
    K = np.zeros((nn * dim, nn * dim))
    for e in conn:
        Ke = orc.element_matrix(etype, coords[e], C[0])
        dofs = (e[:, None] * dim + np.arange(dim)).ravel()
        K[np.ix_(dofs, dofs)] += Ke
    D = A.to_dense()
    np.testing.assert_allclose(D, K, rtol=0, atol=1e-14 * np.abs(K).max())
    R = rigid_modes(coords, dim)
    assert np.abs(D @ R.T).max() <= 1e-12 * np.abs(D).max() * np.abs(R).max()


def test_dirichlet_elimination_closed_form():
    coords, conn = tiny_mesh(P1, 3)
    C = orc.cauchy_green(1e10, 0.2)[None]
    A = orc.assemble(P1, conn, coords, C)
    K = A.to_dense()
    b = rng(3).uniform(-1, 1, A.n)
    dofs = np.array([5, 0, 9, 1], dtype=np.int64)
    vals = np.array([2e-3, -1e-3, 0.5e-3, 0.])
    rhs = b.copy()
    orc.apply_dirichlet(A, rhs, dofs, vals)
    Kx, bx = K.copy(), b.copy()
    free = np.setdiff1d(np.arange(A.n), dofs)
    bx[free] -= K[np.ix_(free, dofs)] @ vals
    Kx[:, dofs] = 0.
    Kx[dofs, :] = 0.
    Kx[dofs, dofs] = 1.
    bx[dofs] = vals
    np.testing.assert_allclose(A.to_dense(), Kx, rtol=0, atol=1e-15 * np.abs(K).max())
    # the lift sums products of size |K| |g|: rounding is relative to those
    np.testing.assert_allclose(rhs, bx, rtol=1e-14, atol=1e-14 * np.abs(K).max() * np.abs(vals).max())


def test_spmv_dot_and_inverses_vs_numpy():
    p = configs.config3(2)
    A = orc.assemble(TET4, p.conn, p.coords, p.cmats, damage=p.damage)
    D = A.to_dense()
    x = rng(4).uniform(-1, 1, A.n)
    bb = rng(5).uniform(-1, 1, A.n)
    np.testing.assert_allclose(orc.spmv(A, x, bb), D @ x - bb, rtol=0, atol=1e-13 * np.abs(D).max())
    assert abs(orc.dot(x, bb) - float(np.dot(x, bb))) <= 1e-14 * np.abs(x).sum()
    np.testing.assert_allclose(orc.inverse_diagonal(A), 1. / np.diag(D), rtol=1e-15)


def test_pcg_vs_direct_solve_and_block_jacobi():
    p = configs.config2(4)
    A = orc.assemble(P2, p.conn, p.coords, p.cmats, phase=p.phase)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p.bc_dofs, p.bc_vals)
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    for precond in (0, 1):
        x, nit, rr, ok = orc.pcg(A, b, rtol=1e-13, maxit=100000, precond=precond)
        assert ok
        assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
        assert rr <= 1e-13


def test_pcg_plane_stress_patch_solution():
    p = configs.config1(6)
    A = orc.assemble(P1, p.conn, p.coords, p.cmats)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p.bc_dofs, p.bc_vals)
    x, nit, rr, ok = orc.pcg(A, b, rtol=1e-14, maxit=100000)
    assert ok
    nu = 0.2
    X = p.coords
    exact = np.column_stack([-nu * 1e-3 * X[:, 0], 1e-3 * X[:, 1]]).ravel()
    assert np.abs(x - exact).max() <= 1e-12


@pytest.mark.parametrize("etype", [P1, P2, TET4])
def test_element_strain_of_linear_field(etype):
    coords, conn = tiny_mesh(etype, 2)
    dim = coords.shape[1]
    G = rng(etype).uniform(-1e-3, 1e-3, (dim, dim))
    t = rng(etype + 3).uniform(-1, 1, dim)
    u = (coords @ G.T + t).ravel()
    eps = orc.element_strain(etype, conn, coords, u)
    if dim == 2:
        exp = [G[0, 0], G[1, 1], G[0, 1] + G[1, 0]]
    else:
        exp = [G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0], G[0, 1] + G[1, 0]]
    np.testing.assert_allclose(eps, np.tile(exp, (conn.shape[0], 1)), rtol=0, atol=1e-15)


def test_element_strain_dofmap_adds_the_macro_strain():
    p = configs.rve(4, 7, 0.3)
    E = np.array([1e-3, -2e-3, 5e-4])
    eps = orc.element_strain_dofmap(P1, p.conn, p.dofmap, p.coords, np.zeros(p.ndofnodes * 2), E)
    np.testing.assert_allclose(eps, np.tile(E, (p.conn.shape[0], 1)), rtol=0, atol=1e-18)


# ------------------------------------------------------------ damage law
def principal(strain):
    if strain.shape[0] == 3:
        T = np.array([[strain[0], strain[2] / 2], [strain[2] / 2, strain[1]]])
    else:
        xx, yy, zz, yz, xz, xy = strain
        T = np.array([[xx, xy / 2, xz / 2], [xy / 2, yy, yz / 2], [xz / 2, yz / 2, zz]])
    return np.linalg.eigvalsh(T)


@pytest.mark.parametrize("s", [3, 6])
def test_damage_equivalent_strain_and_softening(s):
    r = rng(s)
    k0, kf = 1e-4, 1e-3
    strain = r.normal(0, 4e-4, (400, s))
    kappa = np.zeros(400)
    d = np.zeros(400)
    orc.damage_update(strain, kappa, d, kappa0=k0, kappaf=kf, dmax=1.0)
    eq = np.array([np.sqrt(np.sum(np.maximum(principal(e), 0) ** 2)) for e in strain])
    np.testing.assert_allclose(kappa, eq, rtol=1e-12, atol=1e-20)
    below = eq <= k0
    assert np.all(d[below] == 0.)
    # exponential softening: the uniaxial stress (1 - d) E kappa equals
    # E kappa0 at onset and decays with slope -1/kappaf in log scale
    ab = ~below
    np.testing.assert_allclose(np.log((1 - d[ab]) * kappa[ab] / k0), -(kappa[ab] - k0) / kf, rtol=1e-9, atol=1e-12)


def test_damage_is_irreversible_and_bounded():
    strain = np.array([[8e-4, 0., 0.], [0., 3e-3, 0.], [-5e-4, -5e-4, 0.]])
    kappa, d = np.zeros(3), np.zeros(3)
    d0 = np.array([0.2, 0.0, 0.4])
    orc.damage_update(strain, kappa, d, d0=d0, dmax=0.99)
    d1, k1 = d.copy(), kappa.copy()
    assert d1[0] > 0.2 and d1[2] == 0.4 and d1[1] <= 0.99
    orc.damage_update(strain * 0.1, kappa, d, d0=d0, dmax=0.99)   # unloading
    np.testing.assert_array_equal(d, d1)
    np.testing.assert_array_equal(kappa, k1)


# ------------------------------------------------------------ homogenisation
def test_average_stress_is_volume_weighted():
    p = configs.rve(6, 11, 0.35)
    E = np.array([1e-3, 2e-3, -1e-3])
    eps = np.tile(E, (p.conn.shape[0], 1))
    s, vol = orc.average_stress(P1, p.conn, p.coords, p.cmats, p.phase, eps)
    X = p.coords[p.conn]
    areas = 0.5 * np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
                         - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    exp = (areas[:, None] * (p.cmats[p.phase] @ E)).sum(0) / areas.sum()
    assert abs(vol - 1.0) <= 1e-14
    np.testing.assert_allclose(s, exp, rtol=1e-13)


def test_periodic_condensation_is_pt_k_p():
    n = 3
    p = configs.rve(n, 5, 0.3)
    A = orc.assemble_periodic(P1, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
    full = orc.assemble(P1, p.conn, p.coords, p.cmats, phase=p.phase).to_dense()
    nn = p.coords.shape[0]
    P = np.zeros((2 * nn, 2 * p.ndofnodes))
    for i in range(nn):
        for a in range(2):
            P[2 * i + a, 2 * p.dofmap[i] + a] = 1.
    np.testing.assert_allclose(A.to_dense(), P.T @ full @ P, rtol=0, atol=1e-14 * np.abs(full).max())


def test_homogeneous_rve_reproduces_c():
    n = 5
    p = configs.rve(n, 5, 0.3)
    C = p.cmats[0]
    cm = np.stack([C, C])
    A = orc.assemble_periodic(P1, p.conn, p.dofmap, p.coords, cm, p.phase)
    scale = np.abs(A.to_dense()).max() * 1e-3
    ceff = np.zeros((3, 3))
    for lc in range(3):
        f = orc.macro_rhs(P1, p.conn, p.dofmap, p.coords, cm, p.phase, p.macro[lc], A.nrows)
        # a linear field is in equilibrium in a homogeneous periodic cell
        assert np.abs(f).max() <= 1e-12 * scale
        eps = orc.element_strain_dofmap(P1, p.conn, p.dofmap, p.coords, np.zeros(A.n), p.macro[lc])
        s, vol = orc.average_stress(P1, p.conn, p.coords, cm, p.phase, eps)
        ceff[lc] = s / 1e-3
    np.testing.assert_allclose(ceff, C, rtol=1e-12, atol=1e-12 * np.abs(C).max())


# ---------------------------------------------------------------- non-local strains
def _mesh(etype, n):
    coords, conn = ag.structured_mesh(etype, n)
    return coords, conn


def _measures(etype, conn, coords):
    X = coords[conn]
    if etype == P1:
        a, b, c = X[:, 0], X[:, 1], X[:, 2]
        return 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    u, v, w = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]
    return np.abs(np.einsum("ij,ij->i", u, np.cross(v, w))) / 6.


@pytest.mark.parametrize("etype,n", [(P1, 12), (orc.TET4, 4)])
def test_smoothed_strain_constant_per_phase(etype, n):
    """Each phase's own constant field comes back: weights are normalised
    and neighbours of the other phase are left out."""
    coords, conn = _mesh(etype, n)
    ne = conn.shape[0]
    s = 3 if etype == P1 else 6
    phase = (np.arange(ne) % 3 == 0).astype(np.int32)
    vals = np.array([np.linspace(1, 2, s), np.linspace(-3, 5, s)]) * 1e-4
    eps = vals[phase]
    out = orc.smoothed_strain(etype, conn, coords, eps, 0.05, phase=phase)
    np.testing.assert_allclose(out, eps, rtol=1e-14, atol=1e-19)


@pytest.mark.parametrize("kernel", [orc.QUARTIC_COMPACT, orc.GAUSSIAN_NONCOMPACT])
@pytest.mark.parametrize("etype,n", [(P1, 10), (orc.TET4, 3)])
def test_smoothed_strain_wide_kernel_is_the_phase_average(etype, n, kernel):
    """A radius far above the mesh takes every same-phase element with a
    cache weight that tends to its measure, counted times its measure again
    (getAverageField returns measure * weight): the average weighted by the
    squared measures."""
    coords, conn = _mesh(etype, n)
    ne = conn.shape[0]
    s = 3 if etype == P1 else 6
    rng = np.random.default_rng(7)
    eps = rng.normal(size=(ne, s))
    phase = (rng.random(ne) < 0.3).astype(np.int32)
    out = orc.smoothed_strain(etype, conn, coords, eps, 1e3, kernel=kernel, phase=phase)
    meas = _measures(etype, conn, coords)
    for ph in (0, 1):
        m = phase == ph
        avg = (meas[m, None] ** 2 * eps[m]).sum(0) / (meas[m] ** 2).sum()
        np.testing.assert_allclose(out[m], np.tile(avg, (m.sum(), 1)), rtol=0, atol=1e-5 * np.abs(eps).max())


@pytest.mark.parametrize("etype,n", [(P1, 6), (orc.TET4, 2)])
def test_smoothed_strain_lone_elements(etype, n):
    """An element with no same-phase neighbour keeps its own strain (its own
    weight is positive: its vertices lie within R = 1.1 r of the centroid)."""
    coords, conn = _mesh(etype, n)
    ne = conn.shape[0]
    s = 3 if etype == P1 else 6
    eps = np.random.default_rng(3).normal(size=(ne, s))
    out = orc.smoothed_strain(etype, conn, coords, eps, 0.01, phase=np.arange(ne, dtype=np.int32))
    np.testing.assert_allclose(out, eps, rtol=4e-16, atol=0)   # w * eps / w: one rounding


def test_smoothed_strain_fractured_as_excluded_phase():
    """Leaving fractured elements (damage >= threshold) out is the same as
    putting them in a phase of their own, for every other element."""
    coords, conn = _mesh(P1, 10)
    ne = conn.shape[0]
    rng = np.random.default_rng(11)
    eps = rng.normal(size=(ne, 3))
    frac = rng.random(ne) < 0.2
    d = np.where(frac, 0.95, 0.1)
    a = orc.smoothed_strain(P1, conn, coords, eps, 0.12, damage=d, threshold=0.9)
    b = orc.smoothed_strain(P1, conn, coords, eps, 0.12, phase=frac.astype(np.int32))
    np.testing.assert_allclose(a[~frac], b[~frac], rtol=1e-14, atol=0)


def test_smoothed_strain_locus_radius():
    """Two triangles: the second counts exactly when its centroid lies within
    2.5 R + 1.5 r of the first's, with r the circumradius (here from
    r = abc / (4 area)) and R = max(radius, 1.1 r); with the Gaussian kernel
    its weight is positive, so the inclusion shows."""
    tri = np.array([[0., 0.], [1., 0.], [0.3, 0.8]])
    a, b, c = (np.linalg.norm(tri[i] - tri[j]) for i, j in ((0, 1), (1, 2), (2, 0)))
    area = 0.5 * abs((tri[1, 0] - tri[0, 0]) * (tri[2, 1] - tri[0, 1]) - (tri[1, 1] - tri[0, 1]) * (tri[2, 0] - tri[0, 0]))
    r = a * b * c / (4. * area)
    radius = 0.2
    R = max(radius, 1.1 * r)
    rho = 2.5 * R + 1.5 * r
    cen = tri.mean(0)
    for dist, inside in ((rho * (1 - 1e-9), True), (rho * (1 + 1e-9), False)):
        shift = np.array([dist, 0.])
        coords = np.vstack([tri, tri + shift])
        conn = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
        eps = np.array([[0., 0., 0.], [1., 1., 1.]])
        out = orc.smoothed_strain(P1, conn, coords, eps, radius, kernel=orc.GAUSSIAN_NONCOMPACT)
        assert (out[0, 0] > 0) == inside, (dist, rho, out[0])
        assert np.allclose(np.linalg.norm((tri + shift).mean(0) - cen), dist)


def test_gauss_strain_of_a_quadratic_field_is_exact():
    """P2 reproduces a quadratic displacement field exactly, so the strain at
    each of the 4 Gauss points (0.2, 0.2), (0.6, 0.2), (0.2, 0.6), centroid,
    mapped to the physical triangle (x = eta X0 + (1 - xi - eta) X2 + xi X4,
    straight sides) is the analytic gradient there: a swapped or shifted
    Gauss point, a transposed Jacobian or a lost shear term fails."""
    coords, conn = ag.structured_mesh(ag.P2, 4)
    x, y = coords[:, 0], coords[:, 1]
    a, b, c, d, e, f = 0.7, -0.4, 0.3, 0.2, 0.5, -0.6
    u = np.zeros(2 * len(x))
    u[0::2] = 1e-3 * (a * x * x + b * x * y + c * y * y)
    u[1::2] = 1e-3 * (d * x * x + e * x * y + f * y * y)
    eps, _ = orc.gauss_fields(P2, conn, coords, ag.cauchy_green(1e10, 0.2)[None], u)
    G = np.array([[0.2, 0.2], [0.6, 0.2], [0.2, 0.6], [1 / 3, 1 / 3]])
    X = coords[conn]
    for g, (xi, eta) in enumerate(G):
        P = eta * X[:, 0] + (1 - xi - eta) * X[:, 2] + xi * X[:, 4]
        px, py = P[:, 0], P[:, 1]
        exx = 1e-3 * (2 * a * px + b * py)
        eyy = 1e-3 * (e * px + 2 * f * py)
        gxy = 1e-3 * (b * px + 2 * c * py + 2 * d * px + e * py)
        np.testing.assert_allclose(eps[:, g, 0], exx, rtol=0, atol=1e-12)
        np.testing.assert_allclose(eps[:, g, 1], eyy, rtol=0, atol=1e-12)
        np.testing.assert_allclose(eps[:, g, 2], gxy, rtol=0, atol=1e-12)


@pytest.mark.parametrize("etype", [P1, P2, TET4])
def test_gauss_stress_of_uniaxial_stress_state(etype):
    """A displacement field whose strain is that of uniaxial stress sigma_xx
    = s (plane stress: e_yy = -nu e_xx; 3D: e_yy = e_zz = -nu e_xx) gives
    stress (s, 0, 0 ...) at every Gauss point, scaled by (1 - d) of the
    element."""
    E, nu, s = 1e10, 0.25, 2e6
    coords, conn = _mesh(etype, 3)
    exx = s / E
    u = np.zeros(coords.size)
    dim = coords.shape[1]
    u[0::dim] = exx * coords[:, 0]
    u[1::dim] = -nu * exx * coords[:, 1]
    if dim == 3:
        u[2::3] = -nu * exx * coords[:, 2]
    C = ag.cauchy_green(E, nu, dim=dim)[None]
    d = np.linspace(0., 0.5, conn.shape[0])
    _, sig = orc.gauss_fields(etype, conn, coords, C, u, damage=d)
    expect = np.zeros(sig.shape[2])
    expect[0] = s
    want = np.broadcast_to((1 - d)[:, None, None] * expect[None, None, :], sig.shape)
    np.testing.assert_allclose(sig, want, rtol=0, atol=1e-6 * s)
