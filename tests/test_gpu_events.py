"""The device's event-driven damage (amie_gpu_batch_event_step) against the
compiled reference's FeatureTree runs (tests/events_case.py): the damage of
every element after every load step, the number of passes (solves) of each
step, the fracture scores of the last pass and the displacements."""
import numpy as np
import pytest

import amie_gpu as ag
import events_case as ev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = ag.Context(0)
    yield c
    c.close()


def problem(case):
    return ag.Problem(ag.P1, case["coords"], case["conn"], case["cmats"], phase=case["phase"],
                      bc_dofs=case["bc_dofs"], bc_vals=ev.bc_values(case, 0.))


def run(B, case, nprob=1, **kw):
    out = []
    for v in case["loads"]:
        for p in range(nprob):
            B.set_bc_values(case["bc_dofs"], ev.bc_values(case, v), p=p)
        st = B.event_step(case["crit"], case["prm"], case["radius"], **kw)
        out.append((st, [B.damage(p)[0] for p in range(nprob)], [B.scores(p) for p in range(nprob)],
                    [B.solution(p) for p in range(nprob)]))
    return out


SOLVES = {"fp64": dict(precision=ag.FP64, rtol=1e-12), "fix48": dict(precision=ag.MIXED_FIX48, rtol=1e-12)}


@pytest.mark.parametrize("solver", sorted(SOLVES))
@pytest.mark.parametrize("tag", ev.TAGS)
def test_event_step_trajectory_vs_reference(ctx, tag, solver):
    """Every load step equal to the reference's: damage of every element
    exactly (multiples of 0.1, fractured at 0.6), passes equal, scores of
    the last pass to 1e-8, displacements to 1e-9 (FP64) / 1e-8 (FIX48
    mixed precision, same stopping residual)."""
    print("case", tag, solver, flush=True)
    case = ev.load(tag)
    B = ag.Batch(ctx, [problem(case)])
    m = case["phase"] == 0
    utol = 1e-9 if solver == "fp64" else 1e-8
    for k, (st, d, sc, x) in enumerate(run(B, case, **SOLVES[solver])):
        assert st["status"] == 0 and st["converged"] == 1, (k, st)
        np.testing.assert_allclose(d[0], case["damage"][k], rtol=0, atol=1e-12, err_msg=f"step {k}")
        assert st["passes"] == case["passes"][k], (k, st["passes"], case["passes"][k])
        np.testing.assert_allclose(sc[0][m], case["score"][k][m], rtol=0, atol=1e-8)
        assert np.linalg.norm(x[0] - case["u"][k]) <= utol * np.linalg.norm(case["u"][k]), k
        assert st["fractured"] == int((case["damage"][k] >= 0.6).sum())
    B.close()


def test_event_step_same_in_a_batch(ctx):
    """Two copies of a case in one batch (the visiting order and the
    neighbourhoods per problem): each copy's damage, scores and passes equal
    the reference's, and the two copies are bitwise equal."""
    case = ev.load("events_r01")
    B = ag.Batch(ctx, [problem(case), problem(case)])
    for k, (st, d, sc, x) in enumerate(run(B, case, nprob=2, precision=ag.FP64, rtol=1e-12)):
        assert st["passes"] == case["passes"][k], (k, st)
        for p in range(2):
            np.testing.assert_allclose(d[p], case["damage"][k], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(d[0], d[1])
        np.testing.assert_array_equal(sc[0], sc[1])
        np.testing.assert_array_equal(x[0], x[1])
    B.close()


def test_event_step_visiting_order_matters(ctx):
    """Mesh order (first_deferred 0) instead of the reference's does not
    reproduce the small-radius run: the order is honoured."""
    case = ev.load("events_r01")
    B = ag.Batch(ctx, [problem(case)])
    B.set_bc_values(case["bc_dofs"], ev.bc_values(case, case["loads"][0]))
    B.event_step(case["crit"], case["prm"], case["radius"], first_deferred=0, precision=ag.FP64, rtol=1e-12)
    assert np.abs(B.damage()[0] - case["damage"][0]).max() > 0.05
    B.close()


def test_event_step_rejects_bad_input(ctx):
    """Quadratic elements, material tables shorter than the problem's, and
    bad parameters are refused with an error, nothing is run."""
    case = ev.load("events_p1")
    B = ag.Batch(ctx, [problem(case)])
    with pytest.raises(ag.AmieError):
        B.event_step(case["crit"][:1], case["prm"][:1], case["radius"])
    with pytest.raises(ag.AmieError):
        B.event_step(case["crit"], case["prm"], -1.)
    with pytest.raises(ag.AmieError):
        B.event_step(case["crit"], case["prm"], case["radius"], fibre_fraction=0.)
    B.close()
    coords, conn = ag.structured_mesh(ag.P2, 4)
    B2 = ag.Batch(ctx, [ag.Problem(ag.P2, coords, conn, case["cmats"][:1], bc_dofs=[0, 1], bc_vals=[0., 0.])])
    with pytest.raises(ag.AmieError):
        B2.event_step([1], case["prm"][:1], 0.05)
    B2.close()


# --- IsotropicLinearDamage (the Newton branch) ------------------------------

NEWTON = dict(damage_model=ag.DAMAGE_ISOTROPIC_LINEAR)


@pytest.mark.parametrize("solver", sorted(SOLVES))
def test_newton_first_step_vs_reference(ctx, solver):
    """The first load step of the reference's IsotropicLinearDamage run: 201
    passes, the damage of every element to 1e-8 (1e-7 with the FIX48 mixed
    solve), scores of the last pass to 1e-8, displacements to 1e-9 / 1e-8."""
    print("case events_newton step 0", solver, flush=True)
    case = ev.load(ev.NEWTON)
    m = case["phase"] == 0
    B = ag.Batch(ctx, [problem(case)])
    B.set_bc_values(case["bc_dofs"], ev.bc_values(case, case["loads"][0]))
    st = B.event_step(case["crit"], case["prm"], case["radius"], **NEWTON, **SOLVES[solver])
    fp64 = solver == "fp64"
    assert st["status"] == 0 and st["converged"] == 1 and st["fractured"] == 0, st
    assert st["passes"] == case["passes"][0] == 201
    np.testing.assert_allclose(B.damage()[0], case["damage"][0], rtol=0, atol=1e-8 if fp64 else 1e-7)
    np.testing.assert_allclose(B.scores()[m], case["score"][0][m], rtol=0, atol=1e-8)
    x = B.solution()
    assert np.linalg.norm(x - case["u"][0]) <= (1e-9 if fp64 else 1e-8) * np.linalg.norm(case["u"][0])
    B.close()


def test_newton_passes_vs_reference(ctx):
    """Twenty passes of the reference's run, each from the reference's
    damage before it (one solve and one pass on the device): the scores
    equal the reference's to 1e-8 and the damage after the pass equals the
    reference's to 1e-8 (every window and floor decision has a margin far
    above the 1e-10 by which the two solvers' scores differ)."""
    case = ev.load(ev.NEWTON)
    m = case["phase"] == 0
    tol = 2e-3
    B = ag.Batch(ctx, [problem(case)])
    for j, k in enumerate(case["pass_index"]):
        sc_ref = case["pass_score"][j]
        T = sc_ref[m].max()
        margin = min(np.abs(T - sc_ref[m] - 4 * tol).min(), np.abs(sc_ref[m] - 0.05 * tol).min())
        assert margin > 1e-9, (k, margin)
        print("pass", k, flush=True)
        B.set_damage(case["pass_state"][j])
        B.set_bc_values(case["bc_dofs"], ev.bc_values(case, case["loads"][case["pass_load"][j]]))
        st = B.event_step(case["crit"], case["prm"], case["radius"], max_passes=1, precision=ag.FP64, rtol=1e-12,
                          **NEWTON)
        assert st["passes"] == 1 and st["status"] == 0, st
        np.testing.assert_allclose(B.scores()[m], sc_ref[m], rtol=0, atol=1e-8, err_msg=f"pass {k}")
        np.testing.assert_allclose(B.damage()[0], case["pass_next"][j], rtol=0, atol=1e-8, err_msg=f"pass {k}")
    B.close()


def device_pass_vs_oracle(ctx, prob, crit, prm, radius, state, dofs, vals, etype):
    """One device pass from state; the oracle's pass on the device's own
    scores must give the same damage bit for bit."""
    import oracle as orc
    B = ag.Batch(ctx, [prob])
    B.set_damage(state)
    B.set_bc_values(dofs, vals)
    st = B.event_step(crit, prm, radius, max_passes=1, precision=ag.FP64, rtol=1e-12, **NEWTON)
    assert st["status"] == 0, st
    sc = B.scores()
    d_dev = B.damage()[0]
    d_or = np.array(state, dtype=np.float64)
    with orc.one_thread():
        ch = orc.newton_pass(etype, prob.conn, prob.coords, prob.phase, crit, sc, d_or, radius)
    B.close()
    np.testing.assert_array_equal(d_dev, d_or)
    assert st["events"] == ch.sum() > 0, (st, ch.sum())


@pytest.mark.parametrize("j", [2, 9, 10, 17])
def test_newton_pass_equals_oracle_pass_wave(ctx, j):
    """The reference's mesh (neighbourhoods of several hundred elements: a
    wavefront per element): the device's pass equals or_newton_pass on the
    device's own scores, bit for bit."""
    case = ev.load(ev.NEWTON)
    device_pass_vs_oracle(ctx, problem(case), case["crit"], case["prm"], case["radius"], case["pass_state"][j],
                          case["bc_dofs"], ev.bc_values(case, case["loads"][case["pass_load"][j]]), ag.P1)


def test_newton_pass_equals_oracle_pass_small(ctx):
    """A structured 8 x 8 P1 square (128 elements: a thread per element)
    with random damage, stretched beyond the threshold: the device's pass
    equals or_newton_pass on the device's own scores, bit for bit."""
    import configs
    prob, dofs, nfix, crit, maz, radius = configs.event_square(8)
    rng = np.random.default_rng(11)
    state = np.where(prob.phase == 0, rng.uniform(0., 0.5, prob.conn.shape[0]), 0.)
    vals = np.concatenate([np.zeros(nfix), np.full(len(dofs) - nfix, 3e-4)])
    device_pass_vs_oracle(ctx, prob, crit, maz, radius, state, dofs, vals, ag.P1)


def test_newton_pass_equals_oracle_pass_tet4(ctx):
    """A structured Tet4 cube (3D scores with the compressive branch), one
    material, random damage, stretched along x: the device's pass equals
    or_newton_pass on the device's own scores, bit for bit."""
    coords, conn = ag.structured_mesh(ag.TET4, 4)
    x = coords[:, 0]
    left, right = np.nonzero(np.isclose(x, 0.))[0], np.nonzero(np.isclose(x, 1.))[0]
    y0 = np.nonzero(np.isclose(coords[:, 1], 0.))[0]
    z0 = np.nonzero(np.isclose(coords[:, 2], 0.))[0]
    dofs = np.concatenate([3 * left, 3 * y0 + 1, 3 * z0 + 2, 3 * right]).astype(np.int64)
    nfix = len(left) + len(y0) + len(z0)
    vals = np.concatenate([np.zeros(nfix), np.full(len(right), 3e-4)])
    C = ag.cauchy_green(1e10, 0.2, dim=3)
    prob = ag.Problem(ag.TET4, coords, conn, C[None], phase=np.zeros(conn.shape[0], dtype=np.int32),
                      bc_dofs=dofs, bc_vals=vals)
    rng = np.random.default_rng(12)
    state = rng.uniform(0., 0.5, conn.shape[0])
    device_pass_vs_oracle(ctx, prob, np.array([1], dtype=np.int32), np.array([[1e-4, 1e10, 0.2, 100.]]), 0.1,
                          state, dofs, vals, ag.TET4)


def test_newton_later_steps_and_batch(ctx):
    """All three load steps, two copies of the case in one batch: the copies
    bitwise equal; the first step as the reference's; later steps, whose
    element-by-element trajectory amplifies 1e-10 score differences to 1e-2,
    within the reference's totals (damage sum to 0.5 %, passes to 5 %, as
    far as the oracle's own runs at solver rtol 1e-12 and 1e-10 differ) and
    within 1.5 times the distance between the device's own runs at solver
    rtol 1e-12 and 1e-10."""
    case = ev.load(ev.NEWTON)
    B2 = ag.Batch(ctx, [problem(case), problem(case)])
    fine = run(B2, case, nprob=2, precision=ag.FP64, rtol=1e-12, **NEWTON)
    B2.close()
    B1 = ag.Batch(ctx, [problem(case)])
    coarse = run(B1, case, precision=ag.FP64, rtol=1e-10, **NEWTON)
    B1.close()
    np.testing.assert_allclose(fine[0][1][0], case["damage"][0], rtol=0, atol=1e-8)
    for k, (st, d, sc, x) in enumerate(fine):
        np.testing.assert_array_equal(d[0], d[1])
        np.testing.assert_array_equal(x[0], x[1])
        if k == 0:
            continue
        ref = case["damage"][k]
        gap_ref = np.abs(d[0] - ref).mean()
        gap_self = np.abs(d[0] - coarse[k][1][0]).mean()
        assert gap_ref < 1.5 * gap_self, (k, gap_ref, gap_self)
        assert abs(d[0].sum() - ref.sum()) <= 5e-3 * ref.sum(), k
        assert abs(st["passes"] - case["passes"][k]) <= 0.05 * case["passes"][k], (k, st["passes"])


def test_newton_rejects_bad_model(ctx):
    case = ev.load(ev.NEWTON)
    B = ag.Batch(ctx, [problem(case)])
    with pytest.raises(ag.AmieError):
        B.event_step(case["crit"], case["prm"], case["radius"], damage_model=7)
    with pytest.raises(ag.AmieError):
        B.event_step(case["crit"], case["prm"], case["radius"], max_passes_per_step=-1, **NEWTON)
    B.close()


@pytest.mark.parametrize("tag", ev.TAGS)
def test_event_step_small_and_batch_solvers_agree(ctx, tag):
    """The reference cases are small enough for the one-workgroup FP64 CG
    (opts.persistent 0); with persistent launches requested the batch solver
    runs instead.  Both give the reference's damage and pass counts, and
    displacements within 1e-9 of each other."""
    case = ev.load(tag)
    out = {}
    for name, kw in (("small", {}), ("batch", dict(persistent=64))):
        B = ag.Batch(ctx, [problem(case)])
        out[name] = run(B, case, precision=ag.FP64, rtol=1e-12, **kw)
        B.close()
    for k in range(len(case["loads"])):
        for name in out:
            st, d, sc, x = out[name][k]
            assert st["passes"] == case["passes"][k], (name, k, st)
            np.testing.assert_allclose(d[0], case["damage"][k], rtol=0, atol=1e-12, err_msg=f"{name} step {k}")
        xs, xb = out["small"][k][3][0], out["batch"][k][3][0]
        assert np.linalg.norm(xs - xb) <= 1e-9 * np.linalg.norm(xb), k


def test_event_step_tet4_vs_oracle(ctx):
    """A Tet4 cube, left face clamped, right face stretched and sheared (the
    3D scores with the compressive branch; the one-workgroup solver with
    3 x 3 blocks): two load steps of the device's fibre-model event step
    equal the oracle's loop on the same mesh, damage exactly and passes
    equal.  (The 3D neighbourhoods are the oracle's centroid balls, not the
    reference's, so this pins device against oracle only.)"""
    import oracle as orc
    coords, conn = ag.structured_mesh(ag.TET4, 4)
    x = coords[:, 0]
    left, right = np.nonzero(np.isclose(x, 0.))[0], np.nonzero(np.isclose(x, 1.))[0]
    dofs = np.concatenate([3 * left, 3 * left + 1, 3 * left + 2, 3 * right, 3 * right + 1]).astype(np.int64)
    C = ag.cauchy_green(1e10, 0.2, dim=3)
    phase = np.zeros(conn.shape[0], dtype=np.int32)
    crit, prm, radius = np.array([1], dtype=np.int32), np.array([[1e-4, 1e10, 0.2, 100.]]), 0.15
    prob = ag.Problem(ag.TET4, coords, conn, C[None], phase=phase, bc_dofs=dofs, bc_vals=np.zeros(len(dofs)))
    B = ag.Batch(ctx, [prob])
    d_or = np.zeros(conn.shape[0])
    for load in (1.4e-4, 2e-4):
        vals = np.concatenate([np.zeros(3 * len(left)), np.full(len(right), load), np.full(len(right), 0.5 * load)])
        print("tet4", load, flush=True)
        B.set_bc_values(dofs, vals)
        st = B.event_step(crit, prm, radius, precision=ag.FP64, rtol=1e-12)
        d_or, _, sc, passes = orc.event_damage(orc.TET4, conn, coords, C[None], phase, crit, prm, dofs, vals, d_or,
                                               radius)
        assert not np.isnan(sc).any()
        assert st["status"] == 0 and st["passes"] == passes, (st, passes)
        np.testing.assert_allclose(B.damage()[0], d_or, rtol=0, atol=1e-12)
    assert d_or.max() > 0
    B.close()


@pytest.mark.parametrize("etype", ["p1", "tet4"])
def test_event_step_cooperative_grid_agrees(ctx, etype):
    """One problem too large for one workgroup (P1 56 x 56: 51 slices; a
    Tet4 12^3 cube: 35 slices) runs its solves over a cooperative grid of
    workgroups (device-wide barriers between the phases); with persistent
    launches requested the batch solver runs instead.  One load step each:
    the same damage, the same passes, displacements within 1e-9."""
    if etype == "p1":
        import configs
        prob, dofs, nfix, crit, maz, radius = configs.event_square(56)
        vals = np.concatenate([np.zeros(nfix), np.full(len(dofs) - nfix, 0.8e-4)])
    else:
        coords, conn = ag.structured_mesh(ag.TET4, 12)
        x = coords[:, 0]
        left, right = np.nonzero(np.isclose(x, 0.))[0], np.nonzero(np.isclose(x, 1.))[0]
        dofs = np.concatenate([3 * left, 3 * left + 1, 3 * left + 2, 3 * right, 3 * right + 1]).astype(np.int64)
        vals = np.concatenate([np.zeros(3 * len(left)), np.full(len(right), 1.4e-4), np.full(len(right), 0.7e-4)])
        C = ag.cauchy_green(1e10, 0.2, dim=3)
        prob = ag.Problem(ag.TET4, coords, conn, C[None], phase=np.zeros(conn.shape[0], dtype=np.int32),
                          bc_dofs=dofs, bc_vals=np.zeros(len(dofs)))
        crit, maz, radius = np.array([1], dtype=np.int32), np.array([[1e-4, 1e10, 0.2, 100.]]), 0.15
    out = {}
    for name, kw in (("grid", {}), ("batch", dict(persistent=64))):
        print("cooperative", etype, name, flush=True)
        B = ag.Batch(ctx, [prob])
        B.set_bc_values(dofs, vals)
        st = B.event_step(crit, maz, radius, precision=ag.FP64, rtol=1e-12, **kw)
        assert st["status"] == 0 and st["converged"] == 1, (name, st)
        out[name] = (st, B.damage()[0], B.solution())
        B.close()
    (sg, dg, xg), (sb, db, xb) = out["grid"], out["batch"]
    assert sg["passes"] == sb["passes"] and sg["events"] == sb["events"], (sg, sb)
    np.testing.assert_array_equal(dg, db)
    assert np.linalg.norm(xg - xb) <= 1e-9 * np.linalg.norm(xb)
    assert dg.max() > 0
