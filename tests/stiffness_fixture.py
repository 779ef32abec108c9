"""The reference's elastic behaviour fixture, read as data.

examples/test/test_stiffness.ini holds the behaviour (Stiffness, E = 10 GPa,
nu = 0.2) and the arguments of the run (--steps 1 --maximum-value 1e6
--stress --constant); examples/test/check_behaviour_test_stiffness_base holds
the line the reference wrote: time, eps_yy * 1e3, sigma_yy / 1e6,
damage * 100.  examples/main_check_behaviour.cpp:60-118 sets up the run: a
0.01 x 0.01 RectangularFeature (plane stress), bottom FIX_ALONG_ETA,
bottom-left corner FIX_ALONG_XI, the top loaded by SET_STRESS_ETA with the
maximum value (a uniform traction).  Copies of both files are in
golden/ref_fixtures.
"""
import os
import re

import numpy as np

FIX = os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures")


def load():
    ini = open(os.path.join(FIX, "test_stiffness.ini")).read()
    E = float(re.search(r"young_modulus\s*=\s*(\S+)", ini).group(1))
    nu = float(re.search(r"poisson_ratio\s*=\s*(\S+)", ini).group(1))
    smax = float(re.search(r"--maximum-value\s+(\S+)", ini).group(1))
    with open(os.path.join(FIX, "check_behaviour_test_stiffness_base")) as f:
        expected = [float(v) for v in f.read().split()]
    return E, nu, smax, expected


def sample(coords_unit, side=0.01, traction=1e6):
    """Unit-square mesh scaled to the sample; Dirichlet DOFs/values and the
    consistent nodal forces of a uniform traction on the top edge (t h / 2 at
    both ends of every linear edge)."""
    xy = np.asarray(coords_unit, dtype=np.float64) * side
    x, y = xy[:, 0], xy[:, 1]
    bottom = np.nonzero(np.isclose(y, 0.))[0]
    corner = np.nonzero(np.isclose(x, 0.) & np.isclose(y, 0.))[0]
    top = np.nonzero(np.isclose(y, side))[0]
    dofs = np.concatenate([2 * bottom + 1, 2 * corner]).astype(np.int64)
    vals = np.zeros(len(dofs))
    f = np.zeros(2 * len(xy))
    ts = top[np.argsort(x[top])]
    for a, b in zip(ts[:-1], ts[1:]):
        h = x[b] - x[a]
        f[2 * a + 1] += traction * h / 2
        f[2 * b + 1] += traction * h / 2
    return xy, dofs, vals, f
