"""The event-driven damage cases the compiled reference ran
(oracle/ref_harness.cpp, --events): a unit square of StiffnessAndFracture
(NonLocalMazars with FiberBasedIsotropicLinearDamage, or with
IsotropicLinearDamage for events_newton) around a stiffer elastic
inclusion, meshed by the reference's FeatureTree, u_x = 0 on the left,
u_y = 0 at the bottom, u_x = load on the right, one FeatureTree::step per
load.  Shared by the CPU pins of the oracle and the GPU tests."""
import os

import numpy as np

import oracle as orc

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref")
TAGS = ("events_p1", "events_r01")
NEWTON = "events_newton"


def load(tag):
    g = lambda n: np.load(os.path.join(REF, f"{tag}_{n}.npy"))
    conn, coords, phase = g("conn").astype(np.int32), g("coords"), g("phase").astype(np.int32)
    E, nu, thr, Gf, radius, E_inc = g("params")
    x, y = coords[:, 0], coords[:, 1]
    left = np.nonzero(np.abs(x - x.min()) < 1e-9)[0]
    bottom = np.nonzero(np.abs(y - y.min()) < 1e-9)[0]
    right = np.nonzero(np.abs(x - x.max()) < 1e-9)[0]
    ncrit = int((phase == 0).sum())
    if tag == NEWTON:
        # selected passes of the run: damage before pass k and k + 1, scores
        # of pass k, the load step of pass k
        extra = dict(pass_index=g("pass"), pass_load=g("pass_load"), pass_state=g("pass_state"),
                     pass_score=g("pass_score"), pass_next=g("pass_next"))
    else:
        extra = dict(order=g("order"), first=g("first"))
    return dict(
        conn=conn, coords=coords, phase=phase, radius=float(radius),
        cmats=np.stack([orc.cauchy_green(E, nu), orc.cauchy_green(E_inc, nu)]),
        crit=np.array([1, 0], dtype=np.int32),
        prm=np.array([[thr, E, nu, Gf], [0., 0., 0., 1.]]),
        loads=g("loads"), left=left, bottom=bottom, right=right,
        bc_dofs=np.concatenate([2 * left, 2 * bottom + 1, 2 * right]).astype(np.int64),
        damage=g("damage"), score=g("score"), u=g("u"), passes=g("ncalls") // ncrit,
        ncrit=ncrit, **extra)


def bc_values(case, load_value):
    return np.concatenate([np.zeros(len(case["left"]) + len(case["bottom"])),
                           np.full(len(case["right"]), load_value)])
