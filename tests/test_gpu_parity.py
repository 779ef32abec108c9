"""GPU parity: the HIP path through the C ABI against the CPU oracle (which is
itself pinned to the compiled reference, tests/test_oracle_golden.py) and
against the reference's golden vectors directly.

Tolerances: pattern, element->block maps and DOF maps are compared bit for
bit; FP64 matrices of the default (gather) assembly bit for bit where the
element matrices are (P1, Tet4), and matrices and right-hand sides otherwise
to 1e-14 of their largest entry (the coloured scatter sums in colour order,
not the reference's serial Kahan order); the FP64 SpMV to 1e-13; the scaled FP32 SpMV to 2e-6; solutions through the
relative residual they reach plus the displacement error against the FP64
oracle solution.
"""
import os
import time
import subprocess

import numpy as np
import pytest

import amie_gpu as ag
import configs
import homogenisation as hm
import oracle as orc

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "golden", "ref")


def g(name):
    return np.load(os.path.join(REF, name + ".npy"))


@pytest.fixture(scope="module")
def ctx():
    c = ag.Context(0)
    yield c
    # batches still referenced by tracebacks keep the context alive until
    # they are collected
    c.close()


def golden_problem(tag):
    if tag.startswith("sys_tet"):
        coords = g(tag + "_xyz")
        E, nu = g(tag + "_mat")
        cm = ag.cauchy_green(E, nu, dim=3)[None]
        return ag.Problem(ag.TET4, coords, g(tag + "_conn"), cm, bc_dofs=g(tag + "_bc_dof"),
                          bc_vals=g(tag + "_bc_val"))
    E0, nu0, E1, nu1 = g(tag + "_mat")
    cm = np.stack([ag.cauchy_green(E0, nu0), ag.cauchy_green(E1, nu1)])
    return ag.Problem(ag.P1, g(tag + "_xy"), g(tag + "_conn"), cm, phase=g(tag + "_phase"),
                      bc_dofs=g(tag + "_bc_dof"), bc_vals=g(tag + "_bc_val"))


SYSTEMS = ["sys_p1_4x3", "sys_p1_8x8", "sys_p1_12x10_mix", "sys_tet_2", "sys_tet_3"]


@pytest.mark.parametrize("method", [ag.ASSEMBLE_GATHER, ag.ASSEMBLE_COLOURED])
@pytest.mark.parametrize("tag", SYSTEMS)
def test_assembly_matches_reference_golden(ctx, tag, method):
    """The gather assembly adds each block's contributions in the reference's
    order with its Kahan compensation, so its matrix is the reference's own,
    bit for bit; the coloured scatter sums in colour order (to 1e-14)."""
    print("case", tag, method, flush=True)
    B = ag.Batch(ctx, [golden_problem(tag)])
    if method != ag.ASSEMBLE_GATHER:
        B.set_assembly(method)
    rs, ci, ar, rhs = B.matrix()
    np.testing.assert_array_equal(rs, g(tag + "_row_size"))
    np.testing.assert_array_equal(ci, g(tag + "_column_index"))
    ref = g(tag + "_array")
    fref = g(tag + "_forces")
    if method == ag.ASSEMBLE_GATHER:
        np.testing.assert_array_equal(ar, ref)
        np.testing.assert_array_equal(rhs, fref)
    assert np.abs(ar - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(rhs - fref).max() <= 1e-13 * np.abs(fref).max()


def element_blocks_oracle(etype, conn, dofmap=None):
    dconn = conn if dofmap is None else dofmap[conn]
    nrows = int(dconn.max()) + 1
    rs, ci = orc.pattern(etype, dconn, nrows)
    start = np.concatenate([[0], np.cumsum(rs.astype(np.int64))])
    n = conn.shape[1]
    out = np.zeros((conn.shape[0], n, n), dtype=np.int64)
    for i in range(n):
        for j in range(n):
            r = dconn[:, i]
            c = dconn[:, j]
            pos = np.array([start[a] + np.searchsorted(ci[start[a]:start[a + 1]], b) for a, b in zip(r, c)])
            out[:, i, j] = pos
    return out


@pytest.mark.parametrize("etype,n", [(ag.P1, 9), (ag.P2, 5), (ag.TET4, 3)])
def test_element_block_map_bit_exact(ctx, etype, n):
    print("case", etype, n, flush=True)
    coords, conn = ag.structured_mesh(etype, n)
    C = ag.cauchy_green(1e10, 0.2, dim=ag.DIM[etype])[None]
    B = ag.Batch(ctx, [ag.Problem(etype, coords, conn, C)])
    np.testing.assert_array_equal(B.element_blocks(), element_blocks_oracle(etype, conn))
    rs, ci, _, _ = B.matrix()
    ors, oci = orc.pattern(etype, conn, coords.shape[0])
    np.testing.assert_array_equal(rs, ors)
    np.testing.assert_array_equal(ci, oci)


def test_periodic_block_map_bit_exact(ctx):
    p = configs.rve(10, 7, 0.3)
    B = ag.Batch(ctx, [p])
    np.testing.assert_array_equal(B.element_blocks(), element_blocks_oracle(ag.P1, p.conn, p.dofmap))


@pytest.mark.parametrize("method", [ag.ASSEMBLE_GATHER, ag.ASSEMBLE_COLOURED])
@pytest.mark.parametrize("make", ["p1", "p2", "tet"])
def test_assembly_vs_oracle(ctx, make, method):
    """Damaged Tet4, two-phase P1 and P2 against the oracle's serial Kahan
    assembly: bit for bit with the gather for P1 and Tet4 (whose element
    matrices are bit-exact); P2 element matrices differ by an ulp, so P2 and
    the coloured scatter are held to 1e-14."""
    print("case", make, method, flush=True)
    p = {"p1": lambda: configs.config1(16), "p2": lambda: configs.config2(10),
         "tet": lambda: configs.config3(5)}[make]()
    B = ag.Batch(ctx, [p])
    if method != ag.ASSEMBLE_GATHER:
        B.set_assembly(method)
    rs, ci, ar, rhs = B.matrix()
    A = orc.assemble(p.etype, p.conn, p.coords, p.cmats, p.phase, p.damage)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p.bc_dofs, p.bc_vals)
    np.testing.assert_array_equal(rs, A.row_size)
    np.testing.assert_array_equal(ci, A.column_index)
    if method == ag.ASSEMBLE_GATHER and make != "p2":
        np.testing.assert_array_equal(ar, A.array)
        np.testing.assert_array_equal(rhs, b)
    err = np.abs(ar - A.array)
    if err.max() > 1e-14 * np.abs(A.array).max():
        b2 = p.coords.shape[1] * (p.coords.shape[1] + p.coords.shape[1] % 2)
        q = int(np.argmax(err)) // b2
        row = int(np.searchsorted(np.cumsum(A.row_size.astype(np.int64)), q, side="right"))
        print("worst block", q, "row", row, "col", int(A.column_index[q]), "gpu", ar[q * b2:(q + 1) * b2],
              "oracle", A.array[q * b2:(q + 1) * b2], "nbad blocks", int((err.reshape(-1, b2).max(1) > 1).sum()))
    assert err.max() <= 1e-14 * np.abs(A.array).max()
    assert np.abs(rhs - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("make", ["p1", "p2", "tet"])
def test_spmv_fp64_fp32(ctx, make):
    print("case", make, flush=True)
    p = {"p1": lambda: configs.config1(20), "p2": lambda: configs.config2(12),
         "tet": lambda: configs.config3(6)}[make]()
    B = ag.Batch(ctx, [p])
    rs, ci, ar, _ = B.matrix()
    A = orc.BlockCSR(ag.DIM[p.etype], rs, ci, ar)
    x = np.random.default_rng(1).uniform(-1, 1, A.n)
    yref = orc.spmv(A, x)
    scale = np.abs(yref).max()
    y64 = B.spmv(x, precision=ag.FP64)
    assert np.abs(y64 - yref).max() <= 1e-13 * scale
    y32 = B.spmv(x, precision=ag.MIXED)
    assert np.abs(y32 - yref).max() <= 2e-6 * scale
    # FP32 + bf16 split values (relative error < 2^-32), FP64 accumulation
    ysp = B.spmv(x, precision=ag.MIXED_SPLIT)
    err = np.abs(ysp - yref).max()
    assert err <= 1e-9 * scale
    assert err < 1e-2 * np.abs(y32 - yref).max()
    # 48-bit fixed point of the scaled matrix (absolute error < 2^-47 per
    # scaled entry), FP64 accumulation
    yfx = B.spmv(x, precision=ag.MIXED_FIX48)
    assert np.abs(yfx - yref).max() <= 1e-12 * scale


def test_fp64_cg_patch_test(ctx):
    # uniaxial patch test: exact P1 solution u_y = 1e-3 y, u_x = -nu 1e-3 x
    p = configs.config1(32)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=ag.FP64, rtol=1e-10)
    assert st["status"] == 0 and st["relres_max"] <= 1e-10
    u = B.solution().reshape(-1, 2)
    x, y = p.coords[:, 0], p.coords[:, 1]
    np.testing.assert_allclose(u[:, 1], 1e-3 * y, rtol=0, atol=1e-10)
    np.testing.assert_allclose(u[:, 0], -0.2 * 1e-3 * x, rtol=0, atol=1e-10)


def oracle_solution(p, rtol=1e-13, precond=0):
    A = orc.assemble(p.etype, p.conn, p.coords, p.cmats, p.phase, p.damage)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p.bc_dofs, p.bc_vals)
    x, it, rr, ok = orc.pcg(A, b, rtol=rtol, maxit=200000, precond=precond)
    assert ok
    return A, b, x


@pytest.mark.parametrize("precision,precond", [(ag.FP64, ag.JACOBI), (ag.MIXED, ag.JACOBI),
                                               (ag.MIXED, ag.BLOCK_JACOBI), (ag.FP64, ag.BLOCK_JACOBI),
                                               (ag.MIXED64, ag.JACOBI), (ag.MIXED_K64, ag.JACOBI),
                                               (ag.MIXED_SPLIT, ag.JACOBI), (ag.MIXED_SPLIT, ag.BLOCK_JACOBI),
                                               (ag.MIXED_FIX48, ag.JACOBI), (ag.MIXED_FIX48, ag.BLOCK_JACOBI)])
def test_cg_composite_p2_vs_oracle(ctx, precision, precond):
    """Relative residual 1e-8 measured with the oracle's FP64 SpMV on the
    oracle's own matrix; the displacement error against the oracle solution
    (rtol 1e-13) is bounded by cond(K) * 1e-8, checked loosely here and tightly
    in test_tight_solve_error."""
    print("case", precision, precond, flush=True)
    p = configs.config2(24)
    A, b, xref = oracle_solution(p)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=precision, precond=precond, rtol=1e-8)
    assert st["status"] == 0, st
    x = B.solution()
    rel = np.linalg.norm(orc.spmv(A, x, b)) / np.linalg.norm(b)
    assert rel <= 1.01e-8
    assert abs(rel - st["relres_max"]) <= 1e-3 * rel + 1e-15
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-4


@pytest.mark.parametrize("precision", [ag.FP64, ag.MIXED, ag.MIXED_SPLIT, ag.MIXED_FIX48])
def test_tight_solve_error(ctx, precision):
    """Solved to relative residual 1e-12, FP64 and mixed precision (FP64
    refinement) reproduce the FP64 oracle displacement to 1e-7."""
    p = configs.config2(24)
    A, b, xref = oracle_solution(p)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=precision, rtol=1e-12)
    assert st["status"] == 0, st
    x = B.solution()
    assert np.linalg.norm(orc.spmv(A, x, b)) / np.linalg.norm(b) <= 1.01e-12
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-7


@pytest.mark.parametrize("precision,precond", [(ag.MIXED, ag.BLOCK_JACOBI), (ag.FP64, ag.JACOBI),
                                               (ag.MIXED_SPLIT, ag.BLOCK_JACOBI), (ag.MIXED_FIX48, ag.BLOCK_JACOBI)])
def test_cg_damaged_tet_vs_oracle(ctx, precision, precond):
    p = configs.config3(8)
    A, b, xref = oracle_solution(p)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=precision, precond=precond, rtol=1e-8)
    assert st["status"] == 0, st
    x = B.solution()
    assert np.linalg.norm(orc.spmv(A, x, b)) / np.linalg.norm(b) <= 1.01e-8
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-5


@pytest.mark.parametrize("tag", ["sys_p1_8x8", "sys_p1_12x10_mix", "sys_tet_3"])
def test_bcsr_dropin_vs_reference_cg(ctx, tag):
    """Drop-in path: the reference-assembled matrix and forces in, the
    reference ConjugateGradient's converged solution out."""
    stride = 3 if "tet" in tag else 2
    B = ag.Batch(ctx, bcsr=(stride, g(tag + "_row_size"), g(tag + "_column_index"), g(tag + "_array"),
                            g(tag + "_forces")))
    st = B.solve(precision=ag.FP64, rtol=1e-12)
    assert st["status"] == 0
    x = B.solution()
    xref = g(tag + "_cg_x")
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-8
    for prec in (ag.MIXED, ag.MIXED_SPLIT, ag.MIXED_FIX48):
        st = B.solve(precision=prec, rtol=1e-10)
        assert st["status"] == 0
        assert np.linalg.norm(B.solution() - xref) / np.linalg.norm(xref) <= 1e-7


@pytest.mark.parametrize("precision", [ag.FP64, ag.MIXED_SPLIT, ag.MIXED_FIX48])
def test_reference_absolute_stop(ctx, precision):
    """ref_eps: the reference's stop sqrt(r . D^-1 r) <= eps
    (solvers/conjugategradient.cpp:101,218,305), measured here with the
    oracle's FP64 SpMV and inverse diagonal on the reference matrix."""
    tag = "sys_p1_12x10_mix"
    A = orc.BlockCSR(2, g(tag + "_row_size"), g(tag + "_column_index"), g(tag + "_array"))
    b = g(tag + "_forces")
    B = ag.Batch(ctx, bcsr=(2, A.row_size, A.column_index, A.array, b))
    eps = 1e-10
    st = B.solve(precision=precision, ref_eps=eps, rtol=1e-30)
    assert st["status"] == 0, st
    x = B.solution()
    r = orc.spmv(A, x, b)
    meas = np.sqrt(np.dot(r, orc.inverse_diagonal(A) * r))
    assert meas <= eps
    # and not far past it: the stop is the first refinement below eps
    assert meas >= 1e-4 * eps
    xref = g(tag + "_cg_x")
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-6


@pytest.mark.parametrize("method", [ag.ASSEMBLE_GATHER, ag.ASSEMBLE_COLOURED])
def test_rve_macro_load_vs_oracle(ctx, method):
    """Periodic RVE batch: matrices and the three macro-strain loads of each
    problem against the oracle, bit for bit with the gather (loads summed per
    row in element order, each element row over its DOFs in order), to 1e-14
    with the coloured scatter."""
    probs = [configs.rve(n, 100 + n, f) for n, f in [(8, 0.25), (12, 0.3), (16, 0.4)]]
    B = ag.Batch(ctx, probs)
    if method != ag.ASSEMBLE_GATHER:
        B.set_assembly(method)
    for k, p in enumerate(probs):
        print("case", method, k, flush=True)
        rs, ci, ar, rhs = B.matrix(k)
        A = orc.assemble_periodic(ag.P1, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
        loads = [orc.macro_rhs(ag.P1, p.conn, p.dofmap, p.coords, p.cmats, p.phase, p.macro[lc], A.nrows)
                 for lc in range(3)]
        orc.apply_dirichlet(A, np.zeros(A.n), p.bc_dofs, p.bc_vals)
        if method == ag.ASSEMBLE_GATHER:
            np.testing.assert_array_equal(ar, A.array)
        assert np.abs(ar - A.array).max() <= 1e-14 * np.abs(A.array).max()
        for lc in range(3):
            f = loads[lc]
            f[p.bc_dofs] = p.bc_vals
            if method == ag.ASSEMBLE_GATHER:
                np.testing.assert_array_equal(rhs[lc], f)
            assert np.abs(rhs[lc] - f).max() <= 1e-13 * np.abs(f).max()


@pytest.mark.parametrize("precision", [ag.MIXED, ag.MIXED_SPLIT, ag.MIXED_FIX48])
def test_rve_batch_effective_stiffness_vs_oracle(ctx, precision):
    probs = [configs.rve(n, 100 + n, f) for n, f in [(8, 0.25), (12, 0.3), (16, 0.4)]]
    B = ag.Batch(ctx, probs)
    st = B.solve(precision=precision, rtol=1e-10)
    assert st["status"] == 0 and st["relres_max"] <= 1e-10
    ceff = B.average_stress()  # [prob][load][stress]
    for k, p in enumerate(probs):
        A = orc.assemble_periodic(ag.P1, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
        loads = [orc.macro_rhs(ag.P1, p.conn, p.dofmap, p.coords, p.cmats, p.phase, p.macro[lc], A.nrows)
                 for lc in range(3)]
        orc.apply_dirichlet(A, np.zeros(A.n), p.bc_dofs, p.bc_vals)  # pinned node: unit rows, value 0
        ref = np.zeros((3, 3))
        for lc in range(3):
            f = loads[lc]
            f[p.bc_dofs] = p.bc_vals
            x, it, rr, ok = orc.pcg(A, f, rtol=1e-13, maxit=100000)
            assert ok
            xg = B.solution(k)[lc]
            assert np.linalg.norm(xg - x) <= 1e-7 * np.linalg.norm(x)
            eps = orc.element_strain_dofmap(ag.P1, p.conn, p.dofmap, p.coords, x, p.macro[lc])
            s, vol = orc.average_stress(ag.P1, p.conn, p.coords, p.cmats, p.phase, eps)
            ref[lc] = s
        np.testing.assert_allclose(ceff[k], ref, rtol=1e-7, atol=1e-7 * np.abs(ref).max())
        C = ceff[k] / 1e-3
        assert np.abs(C - C.T).max() <= 1e-5 * np.abs(C).max()
        # strictly inside the Voigt / Reuss bracket (a sign error in the
        # load would give 2 <C> - C_eff, outside it)
        voigt, reuss = hm.voigt_reuss(ag.P1, p)
        sym = 0.5 * (C + C.T)
        assert np.linalg.eigvalsh(voigt - sym).min() > 1e-4 * np.abs(voigt).max()
        assert np.linalg.eigvalsh(sym - reuss).min() > 1e-4 * np.abs(voigt).max()


@pytest.mark.parametrize("precision", [ag.FP64, ag.MIXED_FIX48])
def test_rve3d_effective_stiffness_vs_oracle(ctx, precision):
    """Periodic Tet4 RVE under all six unit macro strains (two problems of
    three load cases each): matrices and loads bit for bit against the
    oracle, solutions and the 6x6 C_eff against the oracle's (scipy direct
    solve), and C_eff strictly inside the Voigt / Reuss bracket."""
    E = np.eye(6) * 1e-3
    probs = [configs.rve3d(6, 41, 0.35, macro=E[:3]), configs.rve3d(6, 41, 0.35, macro=E[3:])]
    B = ag.Batch(ctx, probs)
    ref, _, xs = hm.effective_stiffness(ag.TET4, probs[0], list(E), solutions=True)
    for k, p in enumerate(probs):
        print("case", precision, k, flush=True)
        rs, ci, ar, rhs = B.matrix(k)
        A = orc.assemble_periodic(ag.TET4, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
        loads = [orc.macro_rhs(ag.TET4, p.conn, p.dofmap, p.coords, p.cmats, p.phase, p.macro[lc], A.nrows)
                 for lc in range(3)]
        orc.apply_dirichlet(A, np.zeros(A.n), p.bc_dofs, p.bc_vals)
        np.testing.assert_array_equal(ar, A.array)
        for lc in range(3):
            f = loads[lc]
            f[p.bc_dofs] = p.bc_vals
            np.testing.assert_array_equal(rhs[lc], f)
    st = B.solve(precision=precision, rtol=1e-10)
    assert st["status"] == 0 and st["relres_max"] <= 1e-10, st
    got = B.average_stress()             # [problem][load][stress]
    C = np.concatenate([got[0], got[1]]) / 1e-3
    for k in range(2):
        for lc in range(3):
            x = xs[3 * k + lc]
            assert np.linalg.norm(B.solution(k)[lc] - x) <= 1e-7 * np.linalg.norm(x)
    np.testing.assert_allclose(C, ref / 1e-3, rtol=0, atol=1e-7 * np.abs(ref / 1e-3).max())
    voigt, reuss = hm.voigt_reuss(ag.TET4, probs[0])
    sym = 0.5 * (C + C.T)
    assert np.abs(C - C.T).max() <= 1e-6 * np.abs(C).max()
    assert np.linalg.eigvalsh(voigt - sym).min() > 1e-4 * np.abs(voigt).max()
    assert np.linalg.eigvalsh(sym - reuss).min() > 1e-4 * np.abs(voigt).max()


@pytest.mark.parametrize("etype", [ag.P1, ag.TET4])
def test_laminate_effective_stiffness_on_device(ctx, etype):
    """A laminate RVE's C_eff has a closed form that P1 / Tet4 reproduce
    exactly (piecewise-linear fluctuation): the device solve, mixed
    precision to rel. res 1e-10, matches it to 1e-7."""
    dim = ag.DIM[etype]
    cm = np.stack([ag.cauchy_green(12e9, 0.3, dim=dim), ag.cauchy_green(60e9, 0.2, dim=dim)])
    if etype == ag.P1:
        n, pattern = 16, [0, 1, 1, 0, 0, 0, 1, 0, 1, 1, 1, 0, 0, 1, 0, 0]
        ph = hm.layer_phases(etype, n, pattern)
        probs = [hm.periodic_problem(etype, n, ph, cm, macro=np.eye(3) * 1e-3)]
        t, nn = [0], [1, 2]
    else:
        n, pattern = 6, [1, 0, 0, 1, 1, 0]
        ph = hm.layer_phases(etype, n, pattern)
        probs = [hm.periodic_problem(etype, n, ph, cm, macro=np.eye(6)[:3] * 1e-3),
                 hm.periodic_problem(etype, n, ph, cm, macro=np.eye(6)[3:] * 1e-3)]
        t, nn = [0, 1, 5], [2, 3, 4]
    B = ag.Batch(ctx, probs)
    st = B.solve(precision=ag.MIXED_FIX48, rtol=1e-10)
    assert st["status"] == 0, st
    C = np.concatenate(list(B.average_stress())).T / 1e-3     # column k: stress of unit strain k
    f1 = float(np.mean(pattern))
    exact = hm.laminate_stiffness([cm[0], cm[1]], [1. - f1, f1], t, nn)
    np.testing.assert_allclose(C, exact, rtol=0, atol=1e-7 * np.abs(exact).max())


@pytest.mark.parametrize("precision", [ag.MIXED, ag.MIXED_SPLIT, ag.MIXED_FIX48])
def test_batch_result_independent_of_mix(ctx, precision):
    probs = [configs.rve(n, 300 + n, 0.3) for n in (10, 14, 20)]
    alone = ag.Batch(ctx, [configs.rve(14, 314, 0.3)])
    alone.solve(precision=precision, rtol=1e-8)
    mixed = ag.Batch(ctx, probs)
    mixed.solve(precision=precision, rtol=1e-8)
    np.testing.assert_array_equal(alone.solution(0), mixed.solution(1))


@pytest.mark.parametrize("precision", [ag.MIXED, ag.MIXED_SPLIT, ag.MIXED_FIX48])
def test_graph_replay_matches_launches(ctx, precision):
    p = configs.config2(16)
    B = ag.Batch(ctx, [p])
    B.solve(precision=precision, rtol=1e-8, use_graph=1)
    x1 = B.solution().copy()
    B.solve(precision=precision, rtol=1e-8, use_graph=0)
    x2 = B.solution().copy()
    B.solve(precision=precision, rtol=1e-8, use_graph=1)
    x3 = B.solution().copy()
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(x1, x3)


@pytest.mark.parametrize("n", [1, 3, 16])
@pytest.mark.parametrize("precision", [ag.FP64, ag.MIXED_FIX48])
def test_stiffness_behaviour_fixture(ctx, n, precision):
    """The reference's elastic behaviour test (examples/test/test_stiffness.ini,
    check_behaviour_test_stiffness_base, examples/main_check_behaviour.cpp:60-118)
    through the device path: assembly, Dirichlet elimination, CG with the
    reference's stop rule and default eps, element strains, volume-averaged
    stress.  n = 1 is the reference's 2-element sample."""
    import stiffness_fixture as sf
    E, nu, smax, expected = sf.load()
    coords, conn = ag.structured_mesh(ag.P1, n)
    xy, dofs, vals, f = sf.sample(coords, traction=smax)
    p = ag.Problem(ag.P1, xy, conn, ag.cauchy_green(E, nu)[None], bc_dofs=dofs, bc_vals=vals, forces=f)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=precision, ref_eps=1e-10)
    assert st["status"] == 0, st
    eps = B.element_strain(0, 0)
    sig = B.average_stress()[0, 0]
    assert eps[:, 1].mean() * 1e3 == pytest.approx(expected[1], rel=5e-6)
    assert sig[1] / 1e6 == pytest.approx(expected[2], rel=5e-6)
    np.testing.assert_allclose(eps, np.tile([-nu * smax / E, smax / E, 0.], (len(eps), 1)), atol=1e-6 * smax / E)


@pytest.mark.parametrize("case,kernel", [("rve", 0), ("rve", 1), ("tet", 0)])
def test_smoothed_strain_vs_oracle(ctx, case, kernel):
    """Non-local strains on the device against the oracle's brute force, from
    the device's own local strains: a P1 RVE batch with two phases, and a
    Tet4 block whose damage field leaves the elements above the threshold
    out."""
    print("case", case, kernel, flush=True)
    if case == "rve":
        probs = [configs.rve(n, 600 + n, 0.3) for n in (12, 20)]
        radius, threshold, etype = 0.0377, 1.0, ag.P1
    else:
        probs = [configs.config3(6)]
        radius, threshold, etype = 0.09, 0.25, ag.TET4
    B = ag.Batch(ctx, probs)
    assert B.solve(precision=ag.FP64, rtol=1e-10)["status"] == 0
    B.set_nonlocal(radius, kernel, threshold)
    for k, p in enumerate(probs):
        d = B.damage(k)[0] if p.damage is not None else None
        for rhs in range(p.nrhs):
            eps = B.element_strain(k, rhs)
            dev = B.smoothed_strain(k, rhs)
            ora = orc.smoothed_strain(etype, p.conn, p.coords, eps, radius, kernel=kernel, phase=p.phase, damage=d,
                                      threshold=threshold)
            np.testing.assert_allclose(dev, ora, rtol=1e-12, atol=1e-12 * np.abs(ora).max())
            if p.damage is not None:
                assert (d >= threshold).any() and (d < threshold).any()


def test_nonlocal_damage_step_vs_oracle(ctx):
    """A damage step driven by the non-local strains: the oracle's law on the
    oracle's smoothed strains of the device's local strains."""
    probs = [configs.rve(n, 700 + n, 0.3) for n in (10, 16)]
    B = ag.Batch(ctx, probs)
    assert B.solve(precision=ag.FP64, rtol=1e-10)["status"] == 0
    B.set_nonlocal(0.05, 0, 1.0)
    before = [(B.element_strain(k, 0), B.damage(k)) for k in range(len(probs))]
    B.damage_step(kappa0=1e-4, kappaf=1e-3)
    for k, p in enumerate(probs):
        eps, (d0, kap0) = before[k]
        sm = orc.smoothed_strain(ag.P1, p.conn, p.coords, eps, 0.05, phase=p.phase, damage=d0, threshold=1.0)
        d, kap = d0.copy(), kap0.copy()
        orc.damage_update(sm, kap, d, kappa0=1e-4, kappaf=1e-3)
        dg, kg = B.damage(k)
        np.testing.assert_allclose(kg, kap, rtol=1e-12, atol=0)
        np.testing.assert_allclose(dg, d, rtol=1e-12, atol=1e-15)
        assert (dg > 0).any()


def test_damage_step_vs_oracle(ctx):
    p = configs.config3(6, step=10)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=ag.FP64, rtol=1e-12)
    assert st["status"] == 0
    x = B.solution()
    d0 = p.damage.copy()
    grown = B.damage_step(kappa0=1e-5, kappaf=1e-3)
    d, kap = B.damage()
    eps = orc.element_strain(ag.TET4, p.conn, p.coords, x)
    kref = np.zeros(len(d0))
    dref = d0.copy()
    gref = orc.damage_update(eps, kref, dref, d0=d0, kappa0=1e-5, kappaf=1e-3)
    assert grown == gref
    np.testing.assert_allclose(kap, kref, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(d, dref, rtol=1e-12, atol=1e-15)
    # re-assembled with C (1 - d)
    rs, ci, ar, _ = B.matrix()
    A = orc.assemble(ag.TET4, p.conn, p.coords, p.cmats, None, d)
    bb = np.zeros(A.n)
    orc.apply_dirichlet(A, bb, p.bc_dofs, p.bc_vals)
    assert np.abs(ar - A.array).max() <= 1e-13 * np.abs(A.array).max()


def test_homogeneous_rhs_returns_zero(ctx):
    coords, conn = ag.structured_mesh(ag.P1, 6)
    p = ag.Problem(ag.P1, coords, conn, ag.cauchy_green(1e10, 0.2)[None],
                   bc_dofs=np.array([0, 1, 3], dtype=np.int64), bc_vals=np.zeros(3))
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=ag.MIXED)
    assert st["status"] == 0
    assert np.all(B.solution() == 0.)


def test_bad_input_is_an_error(ctx):
    coords, conn = ag.structured_mesh(ag.P1, 4)
    bad = conn.copy()
    bad[0, 0] = 10 ** 6
    with pytest.raises(ag.AmieError):
        ag.Batch(ctx, [ag.Problem(ag.P1, coords, bad, ag.cauchy_green(1e10, 0.2)[None])])


def test_fix48_refuses_a_matrix_outside_its_range(ctx):
    """The 48-bit format holds a scaled SPD matrix (|Ks_ij| <= 1); a matrix
    whose diagonal is negative (a negative modulus) is refused with an error,
    not solved with clipped values; the value-indexed candidate too."""
    coords, conn = ag.structured_mesh(ag.P1, 4)
    p = ag.Problem(ag.P1, coords, conn, ag.cauchy_green(-1e10, 0.2)[None],
                   bc_dofs=np.array([0, 1, 3], dtype=np.int64), bc_vals=np.array([0., 0., 1e-3]))
    B = ag.Batch(ctx, [p])
    with pytest.raises(ag.AmieError):
        B.solve(precision=ag.MIXED_FIX48)
    with pytest.raises(ag.AmieError):
        B.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, 8, 0, 1)


def test_iteration_cap_reports_not_converged(ctx):
    p = configs.config2(10)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=ag.MIXED_FIX48, rtol=1e-10, maxit=5)
    assert st["status"] == 1 and st["nproblems_converged"] == 0
    st = B.solve(precision=ag.MIXED_FIX48, rtol=1e-10)
    assert st["status"] == 0


@pytest.mark.parametrize("case", ["rve_fix48", "rve_fix48_dict", "rve_fp64", "tet_fix48_bj", "tet_fp64_bj"])
@pytest.mark.parametrize("grid", [16, 256])
def test_persistent_solve_matches_launches(ctx, case, grid):
    """One persistent launch per solve pass (device-wide barriers between the
    phases) gives the launch path's results bit for bit: solutions,
    iterations and refinements, for 2D RVE batches (value planes and the
    value-indexed matrix) and a damaged Tet4 block with block-Jacobi."""
    print("case", case, grid, flush=True)
    if case.startswith("rve"):
        probs = [configs.rve(n, 1100 + n, 0.3) for n in (8, 20, 33)]
        precond = ag.JACOBI
    else:
        probs = [configs.config3(7, seed=31), configs.config3(5, seed=32)]
        precond = ag.BLOCK_JACOBI
    prec = ag.FP64 if "fp64" in case else ag.MIXED_FIX48
    B = ag.Batch(ctx, probs)
    if case == "rve_fix48_dict":
        B.set_spmv_kernel(prec, precond, 8, 8, 1)
    ref = B.solve(precision=prec, precond=precond, rtol=1e-10, use_graph=0)
    xr = [B.solution(k).copy() for k in range(len(probs))]
    ir = [B.problem_info(k) for k in range(len(probs))]
    got = B.solve(precision=prec, precond=precond, rtol=1e-10, persistent=grid)
    assert ref["status"] == 0 and got["status"] == 0, (ref, got)
    for k in range(len(probs)):
        np.testing.assert_array_equal(B.solution(k), xr[k])
        assert B.problem_info(k) == ir[k]
    assert got["iterations_max"] == ref["iterations_max"] and got["refinements"] == ref["refinements"]
    B.close()


def test_persistent_solve_escalates_like_launches(ctx):
    """The escalation to FP64 of a stalled problem runs as a second
    persistent pass, with the launch path's results and counts."""
    probs = [configs.rve(3, 601, 0.3), configs.rve(48, 602, 0.4)]
    kw = dict(precision=ag.MIXED_FIX48, rtol=1e-10, stall_iters=16, ckpt_interval=8)
    B = ag.Batch(ctx, probs)
    ref = B.solve(use_graph=0, **kw)
    xr = [B.solution(k).copy() for k in range(2)]
    got = B.solve(persistent=64, **kw)
    assert ref["nproblems_escalated"] == 1 and got["nproblems_escalated"] == 1
    for k in range(2):
        np.testing.assert_array_equal(B.solution(k), xr[k])
    for key in ("iterations_max", "iterations_sum", "refinements", "status"):
        assert got[key] == ref[key], key
    B.close()


def test_escalation_keeps_per_problem_counts(ctx):
    """A tiny stall window (16 iterations without halving the FP64
    residual, refinements every 8) stalls the larger RVE, which continues in
    FP64; the small one converges first.  Every problem reports its own
    iterations and refinements: the same as when solved alone, and the
    batch totals are their sum and maximum."""
    small = configs.rve(3, 601, 0.3)
    large = configs.rve(48, 602, 0.4)
    kw = dict(precision=ag.MIXED_FIX48, rtol=1e-10, stall_iters=16, ckpt_interval=8)
    B = ag.Batch(ctx, [small, large])
    st = B.solve(**kw)
    info = [B.problem_info(k) for k in range(2)]
    alone = []
    for p in (small, large):
        A = ag.Batch(ctx, [p])
        sa = A.solve(**kw)
        alone.append((sa, A.problem_info(0)))
        A.close()
    print("batch", st, info, "alone", alone, flush=True)
    assert st["status"] == 0 and st["relres_max"] <= 1e-10, st
    assert alone[0][0]["nproblems_escalated"] == 0 and alone[1][0]["nproblems_escalated"] == 1
    assert st["nproblems_escalated"] == 1
    for k in range(2):
        assert info[k]["iterations"] == alone[k][1]["iterations"], k
        assert info[k]["refinements"] == alone[k][1]["refinements"], k
        assert info[k]["status"] == 0
    assert st["iterations_sum"] == sum(i["iterations"] for i in info)
    assert st["iterations_max"] == max(i["iterations"] for i in info)
    assert st["refinements"] == sum(i["refinements"] for i in info)


def test_nan_in_the_input_is_reported(ctx):
    coords, conn = ag.structured_mesh(ag.P1, 6)
    f = np.zeros(2 * len(coords))
    f[-1] = np.nan
    p = ag.Problem(ag.P1, coords, conn, ag.cauchy_green(1e10, 0.2)[None],
                   bc_dofs=np.array([0, 1, 3], dtype=np.int64), bc_vals=np.zeros(3), forces=f)
    B = ag.Batch(ctx, [p])
    st = B.solve(precision=ag.FP64)
    assert st["status"] == -4, st   # AMIE_ERR_NAN: the reference exits the process instead


@pytest.mark.parametrize("precision,tune", [(ag.MIXED, 0), (ag.MIXED_FIX48, 1)])
def test_server_result_same_alone_and_in_mix(ctx, precision, tune):
    probs = [configs.rve(n, 500 + n, 0.3) for n in (8, 12, 16, 10)]
    S = ag.Server(ctx, 8 << 30, 8, precision=precision, rtol=1e-8, autotune=tune)
    t = S.submit(probs[1], 0.0)
    res = []
    t_end = time.monotonic() + 60.0   # batches are built on worker threads: steps return at once
    while not res and time.monotonic() < t_end:
        S.step(0.0)
        res = S.poll()
    assert res
    x_alone = S.solution(t, probs[1].ndof * 3)
    tickets = [S.submit(p, 1.0) for p in probs]
    got = {}
    while time.monotonic() < t_end:
        S.step(1.0)
        for r in S.poll():
            got[r["ticket"]] = S.solution(r["ticket"], probs[tickets.index(r["ticket"])].ndof * 3)
        if len(got) == len(probs):
            break
    assert len(got) == len(probs)
    np.testing.assert_array_equal(got[tickets[1]], x_alone)
    S.close()


def serve(S, reqs, limit_s=120.0):
    """Submits (problem, key, damage) requests at t = 0 and steps the server
    until every one has a result: {ticket: (result, solution)}."""
    tickets = [S.submit(p, 0.0, key=key, damage=dmg) for p, key, dmg in reqs]
    ndof = {t: p.ndof * p.nrhs for t, (p, _, _) in zip(tickets, reqs)}
    got = {}
    t_end = time.monotonic() + limit_s
    while len(got) < len(tickets) and time.monotonic() < t_end:
        S.step(1.0)
        for r in S.poll():
            x = S.solution(r["ticket"], ndof[r["ticket"]]) if r["status"] == 0 else None
            got[r["ticket"]] = (r, x)
    assert len(got) == len(tickets), (len(got), len(tickets))
    return [got[t] for t in tickets]


def oracle_load_steps(p_of_step, steps, prm):
    """The incremental-loading loop in the reference's step order
    (features/features.cpp:6150-6373): loads of step s, assembly with the
    damage so far, solve (scipy direct), element strains, damage update."""
    import scipy.sparse.linalg as spla
    p = p_of_step(1)
    d0 = p.damage.copy()
    d, kap = d0.copy(), np.zeros(len(d0))
    xs = []
    for s in range(1, steps + 1):
        ps = p_of_step(s)
        A = orc.assemble(ag.TET4, ps.conn, ps.coords, ps.cmats, None, d)
        b = np.zeros(A.n)
        orc.apply_dirichlet(A, b, ps.bc_dofs, ps.bc_vals)
        x = spla.spsolve(A.to_scipy().tocsc(), b)
        eps = orc.element_strain(ag.TET4, ps.conn, ps.coords, x)
        orc.damage_update(eps, kap, d, d0=d0, kappa0=prm[0], kappaf=prm[1], dmax=prm[2])
        xs.append(x)
    return xs, d, kap


def test_served_mix_vs_oracle(ctx):
    """The served mix: 2D P1 RVEs (three load cases), 3D Tet4 RVEs (three
    load cases) and four load steps on a persistent damage key, submitted
    together.  Every RVE solution and C_eff matches the oracle's (scipy
    direct solve); each load step's solution and the damage after the four
    steps match the oracle's sequential loop; and the 3D RVEs and the load
    steps give bitwise the same results served alone.  Solved to rel. res
    1e-10: displacements to 1e-6, C_eff to 1e-7, damage to 1e-6."""
    E6 = np.eye(6)[:3] * 1e-3
    rve2 = [configs.rve(n, 900 + n, 0.3) for n in (12, 20)]
    rve3 = [configs.rve3d(n, 950 + n, 0.3, macro=E6) for n in (4, 6)]
    prm = (1e-4, 1e-3, 1 - 1e-6)

    def dmg_step(s):
        return configs.config3(8, seed=301, step=s, du=1e-4)

    steps = [(dmg_step(s), 7, prm) for s in range(1, 5)]
    S = ag.Server(ctx, 8 << 30, 8, precision=ag.MIXED_FIX48, rtol=1e-10)
    mix = serve(S, [(p, 0, None) for p in rve2 + rve3] + steps)
    assert all(r["status"] == 0 for r, _ in mix), [r for r, _ in mix]
    d_mix, k_mix = S.damage(7, dmg_step(1).conn.shape[0])
    S.close()
    # RVEs against the oracle
    for (r, x), p in zip(mix[:4], rve2 + rve3):
        et = p.etype
        rows, _, xs = hm.effective_stiffness(et, p, list(p.macro), solutions=True)
        xg = x.reshape(p.nrhs, -1)
        for lc in range(p.nrhs):
            assert np.linalg.norm(xg[lc] - xs[lc]) <= 1e-6 * np.linalg.norm(xs[lc]), (et, lc)
            eps = orc.element_strain_dofmap(et, p.conn, p.dofmap, p.coords, xg[lc], p.macro[lc])
            sg, _ = orc.average_stress(et, p.conn, p.coords, p.cmats, p.phase, eps)
            np.testing.assert_allclose(sg, rows[lc], rtol=0, atol=1e-7 * np.abs(rows[lc]).max())
    # load steps against the oracle's loop
    xs, d_ref, k_ref = oracle_load_steps(dmg_step, 4, prm)
    for (r, x), xr in zip(mix[4:], xs):
        assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)
    d_init = dmg_step(1).damage
    assert (d_ref > d_init + 1e-3).any()          # the damage grew
    np.testing.assert_allclose(k_mix, k_ref, rtol=1e-6, atol=0)
    np.testing.assert_allclose(d_mix, d_ref, rtol=0, atol=1e-6)
    # alone = in the mix
    S = ag.Server(ctx, 8 << 30, 8, precision=ag.MIXED_FIX48, rtol=1e-10)
    alone = [serve(S, [(p, 0, None)])[0] for p in rve3] + serve(S, steps)
    d_alone, _ = S.damage(7, dmg_step(1).conn.shape[0])
    S.close()
    for (ra, xa), (rm, xm) in zip(alone, mix[2:]):
        np.testing.assert_array_equal(xa, xm)
        assert ra["iterations"] == rm["iterations"]
    np.testing.assert_array_equal(d_alone, d_mix)


@pytest.mark.parametrize("du,precision,precond", [(2e-5, ag.MIXED_FIX48, ag.BLOCK_JACOBI),
                                                   (1e-4, ag.MIXED_FIX48, ag.BLOCK_JACOBI),
                                                   (1e-4, ag.FP64, ag.JACOBI)])
def test_incremental_damage_trajectory_vs_oracle(ctx, du, precision, precond):
    """The incremental-loading loop of the Tet4 damage case as the benchmark
    runs it (load scaled step by step, re-assembly, a warm-started solve,
    then the damage step), five steps on 8^3, against the oracle's loop in
    the reference's step order (scipy direct solves): every step's
    displacements to 1e-6 (solved to rel. res 1e-10), kappa to 1e-6
    relative and the damage to 1e-6 after every step.  du = 2e-5 is the
    configured increment; 1e-4 makes the damage grow from the first step."""
    print("case", du, precision, precond, flush=True)
    prm = (1e-4, 1e-3, 1 - 1e-6)

    def at(s):
        return configs.config3(8, seed=3, step=s, du=du)

    xs, _, _ = oracle_load_steps(at, 5, prm)
    B = ag.Batch(ctx, [at(1)])
    d = at(1).damage.copy()
    kap = np.zeros(len(d))
    for s in range(1, 6):
        if s > 1:
            B.scale_bc(s / (s - 1))
        B.assemble()
        st = B.solve(precision=precision, precond=precond, rtol=1e-10, use_x0=1 if s > 1 else 0)
        assert st["status"] == 0, (s, st)
        x = B.solution()
        assert np.linalg.norm(x - xs[s - 1]) <= 1e-6 * np.linalg.norm(xs[s - 1]), s
        B.damage_step(*prm)
        # the oracle's update from the device's own solution of this step
        eps = orc.element_strain(ag.TET4, at(s).conn, at(s).coords, x)
        orc.damage_update(eps, kap, d, d0=at(1).damage, kappa0=prm[0], kappaf=prm[1], dmax=prm[2])
        dg, kg = B.damage()
        np.testing.assert_allclose(kg, kap, rtol=1e-6, atol=0)
        np.testing.assert_allclose(dg, d, rtol=0, atol=1e-6)
    if du > 5e-5:
        assert (d > at(1).damage + 1e-3).any()
    B.close()


def test_server_keys_release_and_errors(ctx):
    """A request that no longer fits next to idle resident problems
    completes with AMIE_ERR_NOMEM instead of waiting forever; releasing the
    key frees its memory; a load step whose Dirichlet DOFs differ from the
    resident problem's completes with AMIE_ERR_ARG; the queue drains."""
    p = configs.config3(6, seed=302, step=1)
    big = configs.config3(6, seed=303, step=1)
    B = ag.Batch(ctx, [p])
    resident = B.nbytes            # what the key holds once built
    B.close()
    need = ag.Server.request_bytes(big)
    assert need <= 2 * resident
    # a budget that holds the resident key or the other problem, not both
    S = ag.Server(ctx, resident + need // 2, 4, precision=ag.MIXED_FIX48, rtol=1e-8)
    (r1, _), = serve(S, [(p, 11, None)])
    assert r1["status"] == 0
    (r2, _), = serve(S, [(big, 0, None)])
    assert r2["status"] == -3, r2             # AMIE_ERR_NOMEM: the key still holds its memory
    bad = configs.config3(6, seed=302, step=2)
    bad = ag.Problem(ag.TET4, bad.coords, bad.conn, bad.cmats, damage=bad.damage, bc_dofs=bad.bc_dofs[:-1],
                     bc_vals=bad.bc_vals[:-1])
    (r3, _), = serve(S, [(bad, 11, None)])
    assert r3["status"] == -1, r3             # AMIE_ERR_ARG: other Dirichlet DOFs
    S.release(11)
    with pytest.raises(ag.AmieError):
        S.release(11)
    (r4, _), = serve(S, [(big, 0, None)])
    assert r4["status"] == 0, r4
    S.close()


@pytest.mark.parametrize("etype,n", [(ag.P1, 6), (ag.P2, 6), (ag.TET4, 3)])
def test_device_element_matrices_match_host_and_oracle(ctx, etype, n):
    print("case", etype, n, flush=True)
    coords, conn = ag.structured_mesh(etype, n)
    dim = ag.DIM[etype]
    X = coords[conn]
    rng = np.random.default_rng(5)
    X = X + rng.uniform(-0.1, 0.1, X.shape) / n if etype != ag.P2 else X
    C = np.stack([ag.cauchy_green(1e10 * (1 + e % 3), 0.2, dim=dim) for e in range(conn.shape[0])])
    K = ag.element_matrices_gpu(ctx, etype, X, C)
    for e in range(conn.shape[0]):
        Kh = ag.element_matrix(etype, X[e], C[e])
        Ko = orc.element_matrix(etype, X[e], C[e])
        assert np.abs(K[e] - Kh).max() <= 1e-14 * np.abs(Kh).max(), e
        assert np.abs(K[e] - Ko).max() <= 1e-14 * np.abs(Ko).max(), e


def test_result_independent_of_grid(ctx):
    """Many tiles per workgroup (tiny grid) and one tile per workgroup give the
    same bits: reductions are summed in tile order whatever the schedule."""
    probs = [configs.rve(n, 700 + n, 0.3) for n in (16, 24, 40, 33, 20, 48)]
    B = ag.Batch(ctx, probs)
    B.solve(precision=ag.MIXED, rtol=1e-8)
    ref = [B.solution(k).copy() for k in range(len(probs))]
    for g in (8, 24, 100000):
        st = B.solve(precision=ag.MIXED, rtol=1e-8, max_grid=g)
        assert st["status"] == 0, (g, st)
        for k in range(len(probs)):
            np.testing.assert_array_equal(B.solution(k), ref[k])


@pytest.mark.parametrize("kind", ["rve2d", "rve3d"])
@pytest.mark.parametrize("precision", [ag.FP64, ag.MIXED, ag.MIXED_FIX48])
def test_spmv_candidates_give_same_bits(ctx, kind, precision):
    """Every SpMV candidate of the autotuner (blocks per step x tile order,
    and with FIX48 the value-indexed matrix) gives the same solution bits:
    the same values, the row sums in the same order and the partial sums per
    tile.  autotune keeps one of the candidates, and the autotune solve
    option runs it."""
    if kind == "rve2d":
        probs = [configs.rve(n, 900 + n, 0.3) for n in (12, 20, 31)]
    else:
        probs = [configs.rve3d(n, 950 + n) for n in (4, 6)]
    B = ag.Batch(ctx, probs)
    assert B.solve(precision=precision, rtol=1e-8)["status"] == 0
    ref = [B.solution(k).copy() for k in range(len(probs))]
    tuned = B.autotune(precision, ag.JACOBI, iters=3)
    assert tuned["tile_order"] in (0, 1, 8, 64) and tuned["us"] > 0
    assert tuned["format"] in ((0, 1) if precision == ag.MIXED_FIX48 else (0,))
    found = set()
    for fmt in (0, 1):
        for steps in (1, 2, 4, 8, 16):
            for order in (0, 1, 8, 64):
                try:
                    B.set_spmv_kernel(precision, ag.JACOBI, steps, order, fmt)
                except ag.AmieError:
                    continue
                found.add((fmt, steps))
                st = B.solve(precision=precision, rtol=1e-8)
                assert st["status"] == 0, (fmt, steps, order, st)
                for k in range(len(probs)):
                    np.testing.assert_array_equal(B.solution(k), ref[k])
    # three step counts per format; the value-indexed matrix with FIX48 only
    assert len(found) == (6 if precision == ag.MIXED_FIX48 else 3)
    assert (tuned["format"], tuned["steps"]) in found
    with pytest.raises(ag.AmieError):
        B.set_spmv_kernel(precision, ag.JACOBI, 3, 0)
    with pytest.raises(ag.AmieError):
        B.set_spmv_kernel(precision, ag.JACOBI, tuned["steps"], -1)
    B2 = ag.Batch(ctx, probs)
    assert B2.solve(precision=precision, rtol=1e-8, autotune=1)["status"] == 0
    for k in range(len(probs)):
        np.testing.assert_array_equal(B2.solution(k), ref[k])


def test_value_indexed_matrix_follows_assembly(ctx):
    """The value-indexed matrix is rebuilt at every assembly: after a damage
    step changes the values, solves with it still match the FIX48 planes bit
    for bit; a matrix with more distinct blocks than 16-bit indices reach
    (a damage field over every element) has none, and the candidate is
    refused."""
    probs = [configs.rve(n, 990 + n, 0.3) for n in (10, 17)]
    B = ag.Batch(ctx, probs)
    B.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, 8, 0, 1)
    for step in range(2):
        if step:
            B.damage_step()   # new element stiffness, new values, re-assembly
        st = B.solve(precision=ag.MIXED_FIX48, rtol=1e-8)
        assert st["status"] == 0
        xd = [B.solution(k).copy() for k in range(len(probs))]
        B.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, 8, 0, 0)
        st = B.solve(precision=ag.MIXED_FIX48, rtol=1e-8)
        for k in range(len(probs)):
            np.testing.assert_array_equal(B.solution(k), xd[k])
        B.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, 8, 0, 1)
    D = ag.Batch(ctx, [configs.config3(40)])   # 384 000 elements, each with its own damage
    with pytest.raises(ag.AmieError):
        D.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, 8, 0, 1)
    assert D.autotune(ag.MIXED_FIX48, ag.JACOBI, iters=2)["format"] == 0


DROPIN = os.path.join(os.path.dirname(__file__), "..", "xfem-amie_amd", "host", "dropin_check")


@pytest.mark.parametrize("tag,precision", [("sys_p1_12x10_mix", ag.MIXED_FIX48), ("sys_tet_3", ag.MIXED_FIX48),
                                           ("sys_p1_8x8", ag.FP64), ("sys_p1_8x8", ag.MIXED_SPLIT)])
def test_cpp_linear_solver_dropin(tmp_path, tag, precision):
    """The C++ LinearSolver drop-in (host/gpu_conjugate_gradient.hpp), driven as
    Assembly::mgsolve drives a solver, on reference-assembled systems: the
    reference's stop rule holds and the solution is the reference CG's."""
    print("case", tag, precision, flush=True)
    stride = 3 if "tet" in tag else 2
    files = []
    for name, dt in (("row_size", np.uint32), ("column_index", np.uint32), ("array", np.float64),
                     ("forces", np.float64)):
        f = tmp_path / (name + ".bin")
        g(tag + "_" + name).astype(dt).tofile(f)
        files.append(str(f))
    out = tmp_path / "x.bin"
    r = subprocess.run([DROPIN, str(stride), *files, "1e-10", str(precision), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x = np.fromfile(out, dtype=np.float64)
    A = orc.BlockCSR(stride, g(tag + "_row_size"), g(tag + "_column_index"), g(tag + "_array"))
    res = orc.spmv(A, x, g(tag + "_forces"))
    assert np.sqrt(np.dot(res, orc.inverse_diagonal(A) * res)) <= 1e-10
    xref = g(tag + "_cg_x")
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-6


REF_DROPIN = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "ref_dropin")


@pytest.mark.skipif(not os.path.exists(REF_DROPIN), reason="oracle/_ref/ref_dropin is built only where the "
                                                           "reference sources are (oracle/Makefile.ref)")
@pytest.mark.parametrize("precision", [ag.MIXED_FIX48, ag.FP64])
def test_reference_assembly_mgsolve_dropin(precision):
    """The reference's own Assembly drives the drop-in: the reference-side
    Amie::GpuConjugateGradient : LinearSolver (xfem-amie_amd/host/
    gpuconjugategradient.h, compiled against the reference headers) through
    Assembly::mgprepare + mgsolve over three load steps of a P1 and a Tet4
    case; every step's displacements match Assembly::cgsolve's to 1e-6, one
    device system serves all three steps (oracle/ref_dropin.cpp)."""
    print("case", precision, flush=True)
    r = subprocess.run([REF_DROPIN, str(precision)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert "ALL OK" in r.stdout


def test_triangle_writer_from_device_solution(ctx, tmp_path):
    """The reference's own solve written by its TriangleWriter
    (tests/golden/ref/writer_p1_trg.txt: FeatureTree mesh of a unit box,
    u_x = 0 left, u_y = 0 bottom, u_x = 1e-3 right) against the same problem
    solved on the device and written by amie_write_triangles: same header,
    same coordinates text, displacements, strains and stresses equal to the
    printed 6 digits (within 1e-5 of each field's largest value)."""
    ref = lambda n: g("writer_p1_" + n)
    conn, xy = ref("conn").astype(np.int32), ref("coords")
    x, y = xy[:, 0], xy[:, 1]
    left, bottom, right = np.where(x == x.min())[0], np.where(y == y.min())[0], np.where(x == x.max())[0]
    bc_dofs = np.concatenate([2 * left, 2 * bottom + 1, 2 * right])
    bc_vals = np.concatenate([np.zeros(len(left)), np.zeros(len(bottom)), np.full(len(right), 1e-3)])
    C = ag.cauchy_green(1e10, 0.2)
    p = ag.Problem(ag.P1, xy, conn, C[None], bc_dofs=bc_dofs, bc_vals=bc_vals)
    B = ag.Batch(ctx, [p])
    assert B.solve(precision=ag.FP64, rtol=1e-12)["status"] == 0
    u = B.solution(0)
    eps = B.element_strain(0)
    out = tmp_path / "trg.txt"
    ag.write_triangles(out, conn, xy, u, [eps, eps @ C.T])
    got = out.read_text().splitlines()
    want = open(os.path.join(REF, "writer_p1_trg.txt")).read().splitlines()
    assert got[:4] == want[:4] and len(got) == len(want)
    a = np.array([[float(v) for v in line.split()] for line in got[4:]])
    b = np.array([[float(v) for v in line.split()] for line in want[4:]])
    assert [line.split()[:6] for line in got[4:]] == [line.split()[:6] for line in want[4:]]
    for lo, hi in ((6, 12), (12, 21), (21, 30)):
        scale = np.abs(b[:, lo:hi]).max()
        assert np.abs(a[:, lo:hi] - b[:, lo:hi]).max() <= 1e-5 * scale, (lo, hi)



@pytest.mark.parametrize("case", ["p2", "tet4_damaged"])
def test_gauss_fields_vs_oracle(ctx, case):
    """Strain and stress at every Gauss point (4 for P2, 1 for Tet4) of the
    device's solution against the oracle's restatement of
    ElementState::getField at the Gauss points, to 1e-12 of the largest
    value (the stress with each element's C (1 - d))."""
    print("case", case, flush=True)
    p = configs.config2(20) if case == "p2" else configs.config3(6, seed=3, step=1)
    B = ag.Batch(ctx, [p])
    assert B.solve(precision=ag.FP64, rtol=1e-10)["status"] == 0
    x = B.solution()
    eps, sig = B.gauss_fields()
    oe, os_ = orc.gauss_fields(p.etype, p.conn, p.coords, p.cmats, x, phase=p.phase, damage=p.damage)
    assert eps.shape == oe.shape and eps.shape[1] == (4 if case == "p2" else 1)
    np.testing.assert_allclose(eps, oe, rtol=0, atol=1e-12 * np.abs(oe).max())
    np.testing.assert_allclose(sig, os_, rtol=0, atol=1e-12 * np.abs(os_).max())
    if case != "p2":
        np.testing.assert_allclose(eps[:, 0, :], B.element_strain(), rtol=0, atol=1e-14 * np.abs(oe).max())
    B.close()


def test_gauss_fields_p1_vs_reference_writer(ctx):
    """The P1 element strain and stress of the reference's own FeatureTree
    solve (REAL_STRESS_FIELD / TOTAL_STRAIN_FIELD at the vertices, constant
    over a P1 element; tests/golden/ref/writer_p1_*) from its displacements
    handed to the device."""
    conn, coords, u = g("writer_p1_conn").astype(np.int32), g("writer_p1_coords"), g("writer_p1_u")
    strain, stress = g("writer_p1_strain"), g("writer_p1_stress")
    p = ag.Problem(ag.P1, coords, conn, ag.cauchy_green(1e10, 0.2)[None], bc_dofs=np.array([0], dtype=np.int64),
                   bc_vals=np.zeros(1))
    B = ag.Batch(ctx, [p])
    B.set_solution(u)
    eps, sig = B.gauss_fields()
    np.testing.assert_allclose(eps[:, 0, :], strain[:, 0, :], rtol=0, atol=1e-12 * np.abs(strain).max())
    np.testing.assert_allclose(sig[:, 0, :], stress[:, 0, :], rtol=0, atol=1e-9 * np.abs(stress).max())
    B.close()
