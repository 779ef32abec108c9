"""Pins the CPU oracle against the compiled reference's golden vectors.

The fixtures in tests/golden/ref were written by oracle/ref_harness.cpp linked
against the reference AMIE sources (oracle/Makefile.ref), with one OpenMP
thread; they hold the reference's own outputs for element matrices,
sparsity pattern, Kahan scatter, Dirichlet elimination, SpMV, preconditioners,
dot products and ConjugateGradient::solve.  The reference-held fixture
examples/test/test_cauchy_green_base is restated in test_cauchy_green_fixture;
the elastic behaviour fixture (examples/test/test_stiffness.ini and
check_behaviour_test_stiffness_base) is read from golden/ref_fixtures by
test_stiffness_behaviour_fixture.
"""
import os

import numpy as np
import pytest

import oracle as orc

REF = os.path.join(os.path.dirname(__file__), "golden", "ref")


@pytest.fixture(autouse=True)
def _one_thread():
    # the golden vectors were written with one OpenMP thread
    with orc.one_thread():
        yield


def g(name):
    return np.load(os.path.join(REF, name + ".npy"))


def test_cauchy_green_golden():
    np.testing.assert_array_equal(orc.cauchy_green(1e10, 0.2), g("C_ps_E1e10_nu0.2"))
    np.testing.assert_array_equal(orc.cauchy_green(12e9, 0.3), g("C_ps_E12e9_nu0.3"))
    np.testing.assert_array_equal(orc.cauchy_green(60e9, 0.2), g("C_ps_E60e9_nu0.2"))
    np.testing.assert_array_equal(orc.cauchy_green(30e9, 0.2, dim=3), g("C_3d_E30e9_nu0.2"))


def test_cauchy_green_fixture():
    # examples/test/test_cauchy_green_base (6 significant digits): plane stress
    # and plane strain rows for E = 1e10, nu = 0.2.
    ps = orc.cauchy_green(1e10, 0.2)
    assert ps[0, 0] == pytest.approx(1.04167e10, rel=5e-6)
    assert ps[0, 1] == pytest.approx(2.08333e9, rel=5e-6)
    assert ps[2, 2] == pytest.approx(8.33333e9, rel=5e-6)
    pe = orc.cauchy_green(1e10, 0.2, plane_strain=True)
    assert pe[0, 0] == pytest.approx(1.11111e10, rel=5e-6)
    assert pe[0, 1] == pytest.approx(2.77778e9, rel=5e-6)
    assert pe[2, 2] == pytest.approx(8.33333e9, rel=5e-6)


@pytest.mark.parametrize("n", [1, 3])
def test_stiffness_behaviour_fixture(n):
    """End to end through the oracle (assembly, Dirichlet elimination, the
    reference CG with its default eps, strains, volume-averaged stress) on
    the reference's behaviour test; n = 1 is its 2-element sample."""
    import amie_gpu as ag   # host-side structured mesh (oriented connectivity)
    import stiffness_fixture as sf
    E, nu, smax, expected = sf.load()
    coords, conn = ag.structured_mesh(ag.P1, n)
    xy, dofs, vals, f = sf.sample(coords, traction=smax)
    C = orc.cauchy_green(E, nu)[None]
    A = orc.assemble(orc.P1, conn, xy, C)
    orc.apply_dirichlet(A, f, dofs, vals)
    x, nit, ok = orc.cg_reference(A, f, eps=1e-10)
    assert ok
    eps = orc.element_strain(orc.P1, conn, xy, x)
    sig, vol = orc.average_stress(orc.P1, conn, xy, C, None, eps)
    assert vol == pytest.approx(1e-4, rel=1e-12)
    # the base file's columns (6 significant digits): eps_yy * 1e3, sigma_yy / 1e6, damage * 100
    assert eps[:, 1].mean() * 1e3 == pytest.approx(expected[1], rel=5e-6)
    assert sig[1] / 1e6 == pytest.approx(expected[2], rel=5e-6)
    assert expected[3] == 0.   # Stiffness has no damage
    # closed form of the uniform uniaxial state beside it
    np.testing.assert_allclose(eps, np.tile([-nu * smax / E, smax / E, 0.], (len(eps), 1)), atol=1e-9 * smax / E)
    np.testing.assert_allclose(sig, [0., smax, 0.], atol=1e-8 * smax)


@pytest.mark.parametrize("tag,etype,E,nu", [
    ("p1_a", orc.P1, 1e10, 0.2), ("p1_cw", orc.P1, 12e9, 0.3), ("p1_unit", orc.P1, 1e10, 0.2),
    ("tet_a", orc.TET4, 30e9, 0.2), ("tet_b", orc.TET4, 30e9, 0.2)])
def test_element_matrix_golden(tag, etype, E, nu):
    xy = g("Ke_" + tag + ("_xy" if etype != orc.TET4 else "_xyz"))
    perm = orc.orient(etype, xy)
    np.testing.assert_array_equal(perm, g("Ke_" + tag + "_order"))
    C = orc.cauchy_green(E, nu, dim=orc.DIM[etype])
    K = orc.element_matrix(etype, xy[perm], C)
    Kref = g("Ke_" + tag)
    scale = np.abs(Kref).max()
    assert np.abs(K - Kref).max() <= 4e-16 * scale


@pytest.mark.parametrize("tag,E,nu", [("p2_a", 60e9, 0.2), ("p2_unit", 12e9, 0.3)])
def test_p2_element_matrix_golden(tag, E, nu):
    xy = g("Ke_" + tag + "_xy")  # already in [v0, m01, v1, m12, v2, m20] order
    np.testing.assert_array_equal(orc.orient(orc.P2, xy), np.arange(6))
    K = orc.element_matrix(orc.P2, xy, orc.cauchy_green(E, nu))
    Kref = g("Ke_" + tag)
    assert np.abs(K - Kref).max() <= 4e-16 * np.abs(Kref).max()


P1_SYSTEMS = ["sys_p1_4x3", "sys_p1_8x8", "sys_p1_8x8_nos", "sys_p1_12x10_mix"]
TET_SYSTEMS = ["sys_tet_2", "sys_tet_3"]


def load_system(tag):
    if tag.startswith("sys_tet"):
        etype = orc.TET4
        coords = g(tag + "_xyz")
        E, nu = g(tag + "_mat")
        cm = orc.cauchy_green(E, nu, dim=3)[None]
        phase = np.zeros(len(g(tag + "_conn")), dtype=np.int32)
    else:
        etype = orc.P1
        coords = g(tag + "_xy")
        E0, nu0, E1, nu1 = g(tag + "_mat")
        cm = np.stack([orc.cauchy_green(E0, nu0), orc.cauchy_green(E1, nu1)])
        phase = g(tag + "_phase").astype(np.int32)
    return etype, g(tag + "_conn").astype(np.int32), coords, cm, phase


@pytest.mark.parametrize("tag", P1_SYSTEMS + TET_SYSTEMS)
def test_pattern_bit_exact(tag):
    etype, conn, coords, cm, phase = load_system(tag)
    rs, ci = orc.pattern(etype, conn, coords.shape[0])
    np.testing.assert_array_equal(rs, g(tag + "_row_size"))
    np.testing.assert_array_equal(ci, g(tag + "_column_index"))


@pytest.mark.parametrize("tag", P1_SYSTEMS)
def test_raw_assembly_golden(tag):
    etype, conn, coords, cm, phase = load_system(tag)
    A = orc.assemble(etype, conn, coords, cm, phase)
    ref = g(tag + "_raw_array")
    assert np.abs(A.array - ref).max() <= 1e-15 * np.abs(ref).max()


@pytest.mark.parametrize("tag", P1_SYSTEMS + TET_SYSTEMS)
def test_assembly_with_bcs_golden(tag):
    etype, conn, coords, cm, phase = load_system(tag)
    A = orc.assemble(etype, conn, coords, cm, phase)
    rhs = np.zeros(A.n)
    orc.apply_dirichlet(A, rhs, g(tag + "_bc_dof"), g(tag + "_bc_val"))
    ref = g(tag + "_array")
    assert np.abs(A.array - ref).max() <= 1e-15 * np.abs(ref).max()
    fref = g(tag + "_forces")
    assert np.abs(rhs - fref).max() <= 1e-14 * np.abs(fref).max()


@pytest.mark.parametrize("tag", P1_SYSTEMS + TET_SYSTEMS)
def test_spmv_golden(tag):
    A = orc.BlockCSR(2 if "p1" in tag else 3, g(tag + "_row_size"), g(tag + "_column_index"), g(tag + "_array"))
    b = g(tag + "_forces")
    x = g(tag + "_spmv_x")
    y = orc.spmv(A, x, b)
    np.testing.assert_array_equal(y, g(tag + "_spmv_Axmb"))
    if "p1" in tag:
        np.testing.assert_array_equal(orc.spmv(A, x), g(tag + "_spmv_Ax"))
        assert orc.dot(x, g(tag + "_spmv_Ax")) == g(tag + "_dot_x_Ax")[0]
        np.testing.assert_array_equal(orc.inverse_diagonal(A) * x, g(tag + "_invdiag_x"))
        blk = orc.inverse_2x2_diagonal(A)
        xx = x.reshape(-1, 2)
        t = np.stack([blk[:, 0] * xx[:, 0] + blk[:, 1] * xx[:, 1], blk[:, 2] * xx[:, 0] + blk[:, 3] * xx[:, 1]], 1)
        np.testing.assert_allclose(t.ravel(), g(tag + "_inv2x2_x"), rtol=0, atol=1e-25)


@pytest.mark.parametrize("tag", P1_SYSTEMS + TET_SYSTEMS)
def test_cg_reference_golden(tag):
    A = orc.BlockCSR(2 if "p1" in tag else 3, g(tag + "_row_size"), g(tag + "_column_index"), g(tag + "_array"))
    b = g(tag + "_forces")
    ok_ref, nit_ref, nssor = g(tag + "_cg_info")
    x, nit, ok = orc.cg_reference(A, b, np.zeros(A.n), eps=1e-10, nssor=int(nssor))
    assert ok == ok_ref
    assert nit == nit_ref
    # identical iteration count; the solution agrees to a few ulps of its
    # largest entry (the first q = A p of each restart goes through another
    # expression path in the reference)
    xref = g(tag + "_cg_x")
    assert np.abs(x - xref).max() <= 1e-12 * np.abs(xref).max()


@pytest.mark.parametrize("tag,kernel", [("smooth_quartic", orc.QUARTIC_COMPACT), ("smooth_gauss", orc.GAUSSIAN_NONCOMPACT)])
def test_smoothed_strain_vs_reference_caches(tag, kernel):
    """Non-local strains against the neighbourhoods and weights the compiled
    reference's Mesh::generateCache builds (mesher/mesh.h:546-640, with the
    locus and kernel of FractureCriterion::initialiseCache) on a Delaunay
    mesh of 150 seeded points, composed as Mesh::getSmoothedFields composes
    them (mesher/mesh.h:1440-1466): each neighbour i counts with
    v_i = area_i * w_i, the value ElementState::getAverageField returns
    (elements/integrable_entity.cpp:2026-2133), and the smoothed strain is
    sum v_i eps_i / sum v_i, for a random strain field."""
    xy, conn = g(tag + "_xy"), g(tag + "_conn").astype(np.int32)
    ptr, idx, w = g(tag + "_ptr"), g(tag + "_idx"), g(tag + "_w")
    radius = float(g(tag + "_radius")[0])
    eps = np.random.default_rng(5).normal(size=(conn.shape[0], 3))
    t = xy[conn]
    area = 0.5 * np.abs((t[:, 1, 0] - t[:, 0, 0]) * (t[:, 2, 1] - t[:, 0, 1]) -
                        (t[:, 1, 1] - t[:, 0, 1]) * (t[:, 2, 0] - t[:, 0, 0]))
    ref = np.zeros_like(eps)
    for e in range(conn.shape[0]):
        sl = slice(ptr[e], ptr[e + 1])
        v = area[idx[sl]] * w[sl]
        ref[e] = (v[:, None] * eps[idx[sl]]).sum(0) / v.sum()
    out = orc.smoothed_strain(orc.P1, conn, xy, eps, radius, kernel=kernel)
    np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-12)



def relabel_vertices(conn6, xy, seed):
    """P1 input of a reference P2 mesh: the vertex columns, with the vertices
    given random labels (first-touch numbering must not depend on them)."""
    verts = np.unique(conn6[:, ::2])
    lab = np.full(xy.shape[0], -1, dtype=np.int64)
    lab[verts] = np.random.default_rng(seed).permutation(len(verts))
    vxy = np.empty((len(verts), 2))
    vxy[lab[verts]] = xy[verts]
    return lab[conn6[:, ::2]].astype(np.int32), vxy


@pytest.mark.parametrize("tag", ["ingest_p1", "ingest_tet"])
def test_renumber_vs_reference_meshes(tag):
    """Meshes FeatureTree made itself (RectangularFeature / Sample3D, its
    sampling and Delaunay mesher, then FeatureTree::renumber): with the nodes
    relabelled at random, the oracle's first-touch numbering gives back the
    reference's ids, element by element."""
    conn, coords = g(tag + "_conn"), g(tag + "_coords")
    n = coords.shape[0]
    perm = np.random.default_rng(11).permutation(n)
    ids, count = orc.renumber(perm[conn].astype(np.int32), n)
    assert count == n
    np.testing.assert_array_equal(ids[perm[conn]], conn)


@pytest.mark.parametrize("tag", ["ingest_p2", "ingest_p2_inc"])
def test_midpoints_and_renumber_vs_reference_meshes(tag):
    """FeatureTree with QUADRATIC elements (its stitch() adds the mid-edge
    nodes through DelaunayTree::addSharedNodes, then renumber()), on a box and
    on a box with an inclusion (projection on boundaries off): from the
    vertex-only triangles with random vertex labels, add_midpoints then
    renumber gives the reference's P2 connectivity and, bit for bit, its node
    coordinates."""
    conn6, xy = g(tag + "_conn"), g(tag + "_coords")
    conn3, vxy = relabel_vertices(conn6, xy, 12)
    c6, mid = orc.add_midpoints(vxy, conn3)
    allxy = np.vstack([vxy, mid])
    ids, count = orc.renumber(c6, allxy.shape[0])
    assert count == xy.shape[0] == allxy.shape[0]
    np.testing.assert_array_equal(ids[c6], conn6)
    out = np.empty_like(allxy)
    out[ids] = allxy
    np.testing.assert_array_equal(out, xy)


def test_midpoints_brute_force():
    """Small random triangle soups: one mid-edge node per distinct edge, made
    by the first triangle (in order) that has the edge, shared by every later
    one, at the exact half-sum of the edge's ends."""
    rng = np.random.default_rng(3)
    for trial in range(20):
        nv = int(rng.integers(3, 12))
        conn3 = np.array([rng.choice(nv, 3, replace=False) for _ in range(int(rng.integers(1, 15)))], dtype=np.int32)
        xy = rng.normal(size=(nv, 2))
        c6, mid = orc.add_midpoints(xy, conn3)
        first = {}
        for t, tri in enumerate(conn3):
            for s in range(3):
                a, b = int(tri[s]), int(tri[(s + 1) % 3])
                e = (min(a, b), max(a, b))
                if e not in first:
                    first[e] = nv + len(first)
                    np.testing.assert_array_equal(mid[first[e] - nv], xy[a] * 0.5 + xy[b] * 0.5)
                assert c6[t, 2 * s] == a and c6[t, 2 * s + 1] == first[e]
        assert len(mid) == len(first)


def test_voxel_sampling_vs_reference_files(tmp_path):
    """The compiled reference's VoxelWriter after a FeatureTree solve of a
    unit cube (clamped left face, sheared right face; tests/golden/ref/
    voxels_tet_*): from its tetrahedra in mesh order and their stresses, the
    oracle's voxel ownership and the product's file writer give the six
    VWFT_STRESS files byte for byte."""
    import amie_gpu as ag
    conn, xyz, sig = g("voxels_tet_conn").astype(np.int32), g("voxels_tet_coords"), g("voxels_tet_stress")
    own = orc.voxel_owner(xyz, conn, np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0]), 24)
    vals = np.where(own[None, :] >= 0, sig[own].T, -1e9)
    ag.write_voxels(tmp_path / "v", vals, 24)
    for c in range(6):
        got = (tmp_path / f"v_24_{c}.vox").read_bytes()
        assert got == open(os.path.join(REF, f"voxels_tet_24_{c}.vox"), "rb").read(), c
