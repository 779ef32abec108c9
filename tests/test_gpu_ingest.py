"""GPU parity of the mesh input kernels (amie_gpu_renumber,
amie_gpu_add_midpoints) through the C ABI: against meshes the compiled
reference's FeatureTree made and numbered itself (tests/golden/ref/ingest_*,
written by oracle/ref_harness.cpp), and against the oracle's serial
restatement on larger and irregular inputs.  Integer results and the new
nodes' coordinates are compared bit for bit."""
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
