"""CPU-side checks of the product library: it loads without a GPU, exports
every symbol include/amie_gpu.h declares, and its host helpers (meshes,
orientation, DOF maps, inclusions, random field, C matrices) agree with the
oracle and with the configurations' definitions.  No GPU compute calls."""
import os
import re
import subprocess

import numpy as np
import pytest

import amie_gpu as ag
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "amie_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(amie_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ag.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(ag.SIGNATURES)


@pytest.mark.parametrize("etype,n", [(ag.P1, 7), (ag.P2, 4), (ag.TET4, 3)])
def test_structured_mesh_orientation_matches_reference_rule(etype, n):
    coords, conn = ag.structured_mesh(etype, n)
    for e in range(conn.shape[0]):
        perm = orc.orient(etype, coords[conn[e]])
        np.testing.assert_array_equal(perm, np.arange(conn.shape[1]))
    # the orientation helper reverses a flipped mesh back
    flipped = conn.copy()
    if etype == ag.P1:
        flipped[:, [1, 2]] = flipped[:, [2, 1]]
    elif etype == ag.P2:
        flipped[:, [2, 4]] = flipped[:, [4, 2]]
        flipped[:, [1, 5]] = flipped[:, [5, 1]]
    else:
        flipped[:, [0, 1]] = flipped[:, [1, 0]]
    np.testing.assert_array_equal(ag.orient(etype, flipped, coords), conn)


def test_structured_sizes():
    c, e = ag.structured_mesh(ag.P1, 128)
    assert c.shape[0] == 16641 and 2 * c.shape[0] == 33282 and e.shape[0] == 32768
    rs, ci = orc.pattern(orc.P1, e, c.shape[0])
    assert len(ci) == 115457
    c, e = ag.structured_mesh(ag.TET4, 8)
    assert c.shape[0] == 729 and e.shape[0] == 6 * 512
    rs, ci = orc.pattern(orc.TET4, e, c.shape[0])
    assert rs.max() == 15  # interior nodes of the 6-tet split


def test_p2_mesh_midpoints():
    coords, conn = ag.structured_mesh(ag.P2, 3)
    for e in conn:
        v = coords[e]
        np.testing.assert_allclose(v[1], 0.5 * (v[0] + v[2]))
        np.testing.assert_allclose(v[3], 0.5 * (v[2] + v[4]))
        np.testing.assert_allclose(v[5], 0.5 * (v[4] + v[0]))


def test_periodic_dofmap():
    n = 5
    m = ag.periodic_dofmap(n)
    coords, conn = ag.structured_mesh(ag.P1, n)
    assert m.max() == n * n - 1 and len(np.unique(m)) == n * n
    # opposite edges share DOF nodes
    left = np.nonzero(np.isclose(coords[:, 0], 0))[0]
    right = np.nonzero(np.isclose(coords[:, 0], 1))[0]
    np.testing.assert_array_equal(m[left], m[right])


def test_inclusions_rejection_sampling():
    xyr, frac = ag.inclusions_2d(2, 0.35, 0.01, 0.04, 0.002)
    assert frac >= 0.35
    assert np.all(xyr[:, 2] >= 0.01) and np.all(xyr[:, 2] <= 0.04)
    assert np.all(xyr[:, 0] - xyr[:, 2] >= 0) and np.all(xyr[:, 0] + xyr[:, 2] <= 1)
    d = np.hypot(xyr[:, None, 0] - xyr[None, :, 0], xyr[:, None, 1] - xyr[None, :, 1])
    gap = d - xyr[:, None, 2] - xyr[None, :, 2]
    np.fill_diagonal(gap, 1.0)
    assert gap.min() >= 0.002 - 1e-12
    again, _ = ag.inclusions_2d(2, 0.35, 0.01, 0.04, 0.002)
    np.testing.assert_array_equal(xyr, again)


def test_phase_by_centroid():
    coords, conn = ag.structured_mesh(ag.P1, 40)
    xyr = np.array([[0.5, 0.5, 0.2]])
    ph = ag.phase_by_centroid(ag.P1, conn, coords, xyr)
    cen = coords[conn].mean(axis=1)
    inside = np.hypot(cen[:, 0] - 0.5, cen[:, 1] - 0.5) <= 0.2
    np.testing.assert_array_equal(ph, inside.astype(np.int32))
    area = 0.5 / 40 ** 2 * ph.sum()
    assert abs(area - np.pi * 0.04) < 0.01


def test_damage_field_range_and_seed():
    d = ag.damage_field_3d(3, 16, 0.05, 0.0, 0.3)
    assert d.shape == (6 * 16 ** 3,)
    assert d.min() == pytest.approx(0.0) and d.max() == pytest.approx(0.3)
    np.testing.assert_array_equal(d, ag.damage_field_3d(3, 16, 0.05, 0.0, 0.3))
    # correlation length 0.05 = 3.2 voxels at n = 64: neighbouring voxels
    # differ by much less than the spread of the field
    v = ag.damage_field_3d(3, 64, 0.05, 0.0, 0.3)[::6].reshape(64, 64, 64)
    for ax in range(3):
        assert np.abs(np.diff(v, axis=ax)).mean() < 0.25 * v.std()


def test_cauchy_green_matches_oracle():
    np.testing.assert_array_equal(ag.cauchy_green(1e10, 0.2), orc.cauchy_green(1e10, 0.2))
    np.testing.assert_array_equal(ag.cauchy_green(30e9, 0.2, 3), orc.cauchy_green(30e9, 0.2, dim=3))


def test_context_without_gpu_is_an_error():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import amie_gpu as ag\n"
            "try:\n    ag.Context(0)\nexcept ag.AmieError as e:\n    print('ERR', e.code)\n") % os.path.join(ROOT, "xfem-amie_amd")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"})
    assert "ERR" in r.stdout


GOLD = os.path.join(ROOT, "tests", "golden", "ref")


@pytest.mark.parametrize("tag,etype,E,nu", [("p1_a", ag.P1, 1e10, 0.2), ("p1_unit", ag.P1, 1e10, 0.2),
                                            ("p2_a", ag.P2, 60e9, 0.2), ("p2_unit", ag.P2, 12e9, 0.3),
                                            ("tet_a", ag.TET4, 30e9, 0.2), ("tet_b", ag.TET4, 30e9, 0.2)])
def test_device_element_formulation_matches_reference(tag, etype, E, nu):
    """The element routine the assembly kernel runs (host instantiation of the
    same __host__ __device__ code) against the compiled reference's element
    matrices (tests/golden/ref)."""
    xy = np.load(os.path.join(GOLD, "Ke_%s_%s.npy" % (tag, "xyz" if etype == ag.TET4 else "xy")))
    if etype != ag.P2:
        xy = xy[np.load(os.path.join(GOLD, "Ke_%s_order.npy" % tag))]
    K = ag.element_matrix(etype, xy, ag.cauchy_green(E, nu, dim=ag.DIM[etype]))
    Kref = np.load(os.path.join(GOLD, "Ke_%s.npy" % tag))
    assert np.abs(K - Kref).max() <= 1e-15 * np.abs(Kref).max()


DROPIN = os.path.join(os.path.dirname(__file__), "..", "xfem-amie_amd", "host", "dropin_check")


def test_cpp_dropin_built_and_fails_loudly_without_gpu(tmp_path):
    """The C++ LinearSolver drop-in links against the library and, with no GPU,
    reports the device error and exits 2 (no CPU fallback)."""
    assert os.path.exists(DROPIN), "run __graft_entry__.build()"
    g = lambda n: os.path.join(os.path.dirname(__file__), "golden", "ref", n + ".npy")
    files = []
    for name, dt in (("row_size", np.uint32), ("column_index", np.uint32), ("array", np.float64), ("forces", np.float64)):
        f = tmp_path / (name + ".bin")
        np.load(g("sys_p1_8x8_" + name)).astype(dt).tofile(f)
        files.append(str(f))
    r = subprocess.run([DROPIN, "2", *files, "1e-10", "4", str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 2
    assert "GPU CG" in r.stderr


def test_triangle_writer_matches_reference_file(tmp_path):
    """The file the compiled reference's TriangleWriter wrote after a
    FeatureTree solve (coordinates, displacements, TOTAL_STRAIN_FIELD,
    REAL_STRESS_FIELD; tests/golden/ref/writer_p1_*), rewritten from the
    inputs that writer read (the element states' values at the vertices):
    byte for byte the same."""
    ref = os.path.join(ROOT, "tests", "golden", "ref")
    g = lambda n: np.load(os.path.join(ref, "writer_p1_" + n + ".npy"))
    out = tmp_path / "trg.txt"
    ag.write_triangles(out, g("conn"), g("coords"), g("uvertex"), [g("strain"), g("stress")])
    assert out.read_bytes() == open(os.path.join(ref, "writer_p1_trg.txt"), "rb").read()


def test_binary_triangle_writer_matches_reference_file(tmp_path):
    """The BinaryTriangleWriter file of the same reference solve, rewritten
    from the same inputs: byte for byte the same."""
    ref = os.path.join(ROOT, "tests", "golden", "ref")
    g = lambda n: np.load(os.path.join(ref, "writer_p1_" + n + ".npy"))
    out = tmp_path / "btrg.txt"
    ag.write_triangles(out, g("conn"), g("coords"), g("uvertex"), [g("strain"), g("stress")], binary=True)
    assert out.read_bytes() == open(os.path.join(ref, "writer_p1_btrg.txt"), "rb").read()


def test_multi_triangle_writer_matches_reference(tmp_path):
    """MultiTriangleWriter of the same reference solve, appended twice: the
    index file (the reference's, with its directory replaced by ours) and
    two step files equal to the reference's TriangleWriter file."""
    ref = os.path.join(ROOT, "tests", "golden", "ref")
    g = lambda n: np.load(os.path.join(ref, "writer_p1_" + n + ".npy"))
    want_head = open(os.path.join(ref, "writer_p1_multi_head.txt")).read().splitlines()
    ref_base = want_head[1][: -len("_1")]
    base = str(tmp_path / "writer_p1_multi")
    w = ag.MultiTriangleWriter(tmp_path / "head.txt", base)
    for _ in range(2):
        w.append(g("conn"), g("coords"), g("uvertex"), [g("strain"), g("stress")])
    got = (tmp_path / "head.txt").read_text().splitlines()
    assert got == [line.replace(ref_base, base) for line in want_head]
    trg = open(os.path.join(ref, "writer_p1_trg.txt"), "rb").read()
    for step in (1, 2):
        assert open(f"{base}_{step}", "rb").read() == trg


def test_triangle_writer_argument_errors(tmp_path):
    conn = np.array([[0, 1, 2]], dtype=np.int32)
    xy = np.zeros((3, 2))
    with pytest.raises(ag.AmieError):
        ag.write_triangles(tmp_path / "no_such_dir" / "t.txt", conn, xy, np.zeros(6))
    with pytest.raises(ag.AmieError):
        ag.write_triangles(tmp_path / "t.txt", np.zeros((1, 4), dtype=np.int32), xy, np.zeros((1, 3, 2)))
    with pytest.raises(ag.AmieError, match="outside"):
        ag.write_triangles(tmp_path / "t.txt", np.array([[0, 1, 3]], dtype=np.int32), xy, np.zeros((1, 3, 2)))


@pytest.mark.parametrize("tag", ["place_gap", "place_nogap"])
def test_inclusion_placement_matches_reference(tag):
    """placement2D of the compiled reference (utilities/placement.cpp:87-162)
    on a 1 x 0.8 rectangle: 60 radii with a 0.005 gap (all placed), and 200
    larger radii without a gap (the 3000-try budget runs out after 21).  The
    same inclusions are placed, at the same centres with the same radii, bit
    for bit."""
    ref = os.path.join(ROOT, "tests", "golden", "ref")
    g = lambda n: np.load(os.path.join(ref, tag + "_" + n + ".npy"))
    gap, tries, seed = g("params")
    idx, xyr = ag.place_inclusions_2d(g("radii"), 1.0, 0.8, (0.5, 0.4), min_dist=gap, tries=int(tries), seed=int(seed))
    np.testing.assert_array_equal(idx, g("placed"))
    np.testing.assert_array_equal(xyr, g("out"))


def test_inclusion_placement_properties():
    """Placed circles lie in the rectangle and keep the gap between them."""
    radii = np.sort(np.random.default_rng(3).uniform(0.01, 0.04, 80))[::-1]
    idx, xyr = ag.place_inclusions_2d(radii, 2.0, 1.0, (1.0, 0.5), min_dist=0.01, tries=20000, seed=9)
    assert len(idx) > 40
    x, y, r = xyr.T
    assert (x - r >= -1e-12).all() and (x + r <= 2 + 1e-12).all() and (y - r >= -1e-12).all() and (y + r <= 1 + 1e-12).all()
    d = np.hypot(x[:, None] - x[None], y[:, None] - y[None]) + np.eye(len(x)) * 10
    assert (d >= np.maximum(r[:, None], r[None]) + np.minimum(r[:, None], r[None]) + 0.01 - 1e-12).all()


def test_integration_shows_the_compiled_dropin_file_verbatim():
    """INTEGRATION.md's reference-side LinearSolver is the file that
    oracle/ref_dropin.cpp compiles against the reference headers."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "INTEGRATION.md")).read()
    hdr = open(os.path.join(root, "xfem-amie_amd", "host", "gpuconjugategradient.h")).read()
    a = text.index("This is that file, verbatim")
    b = text.index("```cpp\n", a) + len("```cpp\n")
    c = text.index("```\n", b)
    assert text[b:c] == hdr
