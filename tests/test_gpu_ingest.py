"""GPU parity of the mesh input and output kernels through the C ABI:
amie_gpu_renumber and amie_gpu_add_midpoints against meshes the compiled
reference's FeatureTree made and numbered itself (tests/golden/ref/ingest_*),
amie_gpu_voxelize against the files the reference's VoxelWriter wrote
(tests/golden/ref/voxels_tet_*), all written by oracle/ref_harness.cpp, and
against the oracle's serial restatements on larger and irregular inputs.
Integer results, the new nodes' coordinates and the voxel files are
compared bit for bit."""
import os

import numpy as np
import pytest

import amie_gpu as ag
import oracle as orc

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "golden", "ref")


def g(name):
    return np.load(os.path.join(REF, name + ".npy"))


@pytest.fixture(scope="module")
def ctx():
    c = ag.Context(0)
    yield c
    c.close()


def relabel_vertices(conn6, xy, seed):
    verts = np.unique(conn6[:, ::2])
    lab = np.full(xy.shape[0], -1, dtype=np.int64)
    lab[verts] = np.random.default_rng(seed).permutation(len(verts))
    vxy = np.empty((len(verts), 2))
    vxy[lab[verts]] = xy[verts]
    return lab[conn6[:, ::2]].astype(np.int32), vxy


@pytest.mark.parametrize("tag", ["ingest_p1", "ingest_tet"])
def test_renumber_vs_reference_meshes(ctx, tag):
    print("case", tag, flush=True)
    conn, coords = g(tag + "_conn"), g(tag + "_coords")
    n = coords.shape[0]
    perm = np.random.default_rng(11).permutation(n)
    ids, count = ag.renumber_gpu(ctx, perm[conn], n)
    assert count == n
    np.testing.assert_array_equal(ids[perm[conn]], conn)


@pytest.mark.parametrize("tag", ["ingest_p2", "ingest_p2_inc"])
def test_midpoints_and_renumber_vs_reference_meshes(ctx, tag):
    print("case", tag, flush=True)
    conn6, xy = g(tag + "_conn"), g(tag + "_coords")
    conn3, vxy = relabel_vertices(conn6, xy, 12)
    c6, mid = ag.add_midpoints_gpu(ctx, vxy, conn3)
    allxy = np.vstack([vxy, mid])
    ids, count = ag.renumber_gpu(ctx, c6, allxy.shape[0])
    assert count == xy.shape[0] == allxy.shape[0]
    np.testing.assert_array_equal(ids[c6], conn6)
    out = np.empty_like(allxy)
    out[ids] = allxy
    np.testing.assert_array_equal(out, xy)


@pytest.mark.parametrize("n", [1, 7, 257])
def test_structured_shuffled_vs_oracle(ctx, n):
    """A structured P1 mesh with its nodes relabelled and its triangles
    shuffled (no locality left), sizes from one cell to 132 k triangles."""
    print("case", n, flush=True)
    coords, conn = ag.structured_mesh(ag.P1, n)
    rng = np.random.default_rng(n)
    perm = rng.permutation(coords.shape[0])
    xy = np.empty_like(coords)
    xy[perm] = coords
    conn3 = perm[conn][rng.permutation(conn.shape[0])].astype(np.int32)
    c6, mid = ag.add_midpoints_gpu(ctx, xy, conn3)
    o6, omid = orc.add_midpoints(xy, conn3)
    np.testing.assert_array_equal(c6, o6)
    np.testing.assert_array_equal(mid, omid)
    nall = xy.shape[0] + mid.shape[0]
    ids, count = ag.renumber_gpu(ctx, c6, nall)
    oids, ocount = orc.renumber(c6, nall)
    assert count == ocount == nall
    np.testing.assert_array_equal(ids, oids)


def test_triangle_soup_vs_oracle(ctx):
    """Random triangles over few nodes: edges shared by many triangles (runs
    longer than two after the sort), repeated triangles, and nodes no
    element names (id -1)."""
    rng = np.random.default_rng(7)
    nv = 60
    conn3 = np.array([rng.choice(nv - 5, 3, replace=False) for _ in range(5000)], dtype=np.int32)
    xy = rng.normal(size=(nv, 2))
    c6, mid = ag.add_midpoints_gpu(ctx, xy, conn3)
    o6, omid = orc.add_midpoints(xy, conn3)
    np.testing.assert_array_equal(c6, o6)
    np.testing.assert_array_equal(mid, omid)
    ids, count = ag.renumber_gpu(ctx, conn3, nv)
    oids, ocount = orc.renumber(conn3, nv)
    assert count == ocount
    np.testing.assert_array_equal(ids, oids)
    assert (ids[nv - 5:] == -1).all()


def test_out_of_range_connectivity_is_an_error(ctx):
    """Checked on the host before any kernel indexes by it."""
    conn = np.array([[0, 1, 5]], dtype=np.int32)
    with pytest.raises(ag.AmieError, match="outside"):
        ag.renumber_gpu(ctx, conn, 3)
    with pytest.raises(ag.AmieError, match="outside"):
        ag.add_midpoints_gpu(ctx, np.zeros((3, 2)), conn)


def test_voxel_sampling_vs_reference_files(ctx, tmp_path):
    """Device voxel ownership on the reference's mesh and stresses: the same
    owners as the oracle, and the reference VoxelWriter's six files byte for
    byte once written."""
    conn, xyz, sig = g("voxels_tet_conn").astype(np.int32), g("voxels_tet_coords"), g("voxels_tet_stress")
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    vals = ag.voxelize_gpu(ctx, xyz, conn, sig, box, 24)
    own = orc.voxel_owner(xyz, conn, box, 24)
    np.testing.assert_array_equal(vals, np.where(own[None, :] >= 0, sig[own].T, -1e9))
    ag.write_voxels(tmp_path / "v", vals, 24)
    for c in range(6):
        got = (tmp_path / f"v_24_{c}.vox").read_bytes()
        assert got == open(os.path.join(REF, f"voxels_tet_24_{c}.vox"), "rb").read(), c


def test_voxel_sampling_larger_mesh_vs_oracle(ctx):
    """A structured Tet4 mesh (30^3 cells) with a random field on a 64^3 grid:
    device owners and values equal the oracle's; every voxel is owned."""
    xyz, conn = ag.structured_mesh(ag.TET4, 30)
    field = np.random.default_rng(4).normal(size=(conn.shape[0], 2))
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    vals = ag.voxelize_gpu(ctx, xyz, conn, field, box, 64)
    own = orc.voxel_owner(xyz, conn, box, 64)
    assert (own >= 0).all()
    np.testing.assert_array_equal(vals, field[own].T)


def test_voxel_sampling_ignores_tetrahedra_outside_the_grid(ctx):
    """Tetrahedra far outside the grid (coordinates near +-1e300, where the
    reference's index range would not fit an int) own no voxel and leave
    the others' owners as they are: their index ranges are clamped to the
    grid before the loops."""
    xyz, conn = ag.structured_mesh(ag.TET4, 12)
    field = np.random.default_rng(5).normal(size=(conn.shape[0], 1))
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    own = orc.voxel_owner(xyz, conn, box, 24)
    far = np.array([[1e300, 0, 0], [1e300, 1, 0], [1e300, 0, 1], [1.1e300, 0, 0],
                    [-1e300, -1e300, -1e300], [-1e300, -1e300, -0.9e300], [-1e300, -0.9e300, -1e300],
                    [-0.9e300, -1e300, -1e300]])
    xyz2 = np.vstack([xyz, far])
    n0 = xyz.shape[0]
    conn2 = np.vstack([conn, [[n0, n0 + 1, n0 + 2, n0 + 3], [n0 + 4, n0 + 5, n0 + 6, n0 + 7]]]).astype(np.int32)
    field2 = np.vstack([field, [[7.0], [8.0]]])
    vals = ag.voxelize_gpu(ctx, xyz2, conn2, field2, box, 24)
    assert (own >= 0).all()
    np.testing.assert_array_equal(vals, field[own].T)


def test_voxel_grid_must_be_cubic(ctx):
    conn = np.array([[0, 1, 2, 3]], dtype=np.int32)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    with pytest.raises(ag.AmieError, match="cubic"):
        check = ag.lib().amie_gpu_voxelize
        vals = np.zeros(2 * 3 * 4)
        field = np.zeros(1)
        box = np.array([0.5, 0.5, 0.5, 1, 1, 1.0])
        ag.check(check(ctx.h, 4, xyz.ctypes.data, 1, conn.ctypes.data, 1, field.ctypes.data, box.ctypes.data, 2, 3, 4,
                       vals.ctypes.data, None))


def test_many_triangles_on_one_edge_vs_oracle(ctx):
    """A fan of 20 000 triangles on one edge (0, 1): one run of 20 000 sides
    after the sort, numbered in linear time."""
    k = np.arange(2, 20002, dtype=np.int32)
    conn3 = np.stack([np.zeros_like(k), np.ones_like(k), k], axis=1).astype(np.int32)
    xy = np.random.default_rng(2).normal(size=(20002, 2))
    c6, mid = ag.add_midpoints_gpu(ctx, xy, conn3)
    o6, omid = orc.add_midpoints(xy, conn3)
    np.testing.assert_array_equal(c6, o6)
    np.testing.assert_array_equal(mid, omid)
    assert (c6[:, 1] == xy.shape[0]).all()
