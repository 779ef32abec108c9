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
