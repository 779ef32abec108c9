"""Pins of the oracle's periodic homogenisation path (the macro-strain load
or_macro_rhs, the strain recovery or_element_strain_dofmap and, with them,
the effective stiffness of an RVE) by facts that do not come from the
oracle's own formulas:

* the load of an imposed macro strain E is -P^T K_full u_E, with K_full the
  non-periodic stiffness (or_assemble, pinned bit for bit to the compiled
  reference), P the 0/1 periodic DOF map and u_E any linear field whose
  symmetric gradient is E: the test uses u = G x with a non-symmetric G
  (simple shear plus stretch), which differs from the oracle's symmetric
  split by a rigid rotation that K_full annihilates;
* the strain of a periodic fluctuation plus E equals the reference-pinned
  element strain (or_element_strain) of the full field P x + G x;
* a laminate RVE (phases in layers aligned with element rows / planes) has
  a closed-form effective stiffness (continuous in-plane strains and
  continuous tractions across the layers); its exact fluctuation is
  piecewise linear in the layer coordinate, so P1 / Tet4 reproduce it to
  rounding;
* Voigt / Reuss bounds (physics/homogenization/composite.cpp:649-680: the
  Voigt scheme's strain concentration is the identity, the Reuss scheme's
  uniform stress): <C^-1>^-1 <= C_eff <= <C> as quadratic forms.  The
  discrete solution is stiffer than the exact one, so the Reuss bound holds
  for it too, and the Voigt bound holds because zero fluctuation is in the
  discrete space.  A sign error in the load gives 2<C> - C_eff, outside it.

The linear systems are solved by scipy's sparse direct solver (a library
routine), not by the oracle's CG.
"""
import numpy as np
import pytest

import amie_gpu as ag
import configs
import oracle as orc
from homogenisation import (P1, TET4, effective_stiffness, full_projection, laminate_stiffness, layer_phases,
                            linear_field, periodic_problem, voigt_reuss)


def macros_of(etype):
    s = 3 if etype == P1 else 6
    r = np.random.default_rng(17 + s)
    return [r.uniform(-2e-3, 2e-3, s) for _ in range(3)]


# ------------------------------------------------ load: -P^T K_full u_E
@pytest.mark.parametrize("etype,n,seed", [(P1, 6, 3), (P1, 9, 4), (TET4, 3, 5), (TET4, 4, 6)])
def test_macro_load_is_minus_pt_k_u(etype, n, seed):
    p = configs.rve(n, seed, 0.35) if etype == P1 else configs.rve3d(n, seed)
    b = ag.DIM[etype]
    full = orc.assemble(etype, p.conn, p.coords, p.cmats, phase=p.phase).to_dense()
    P = full_projection(p.dofmap, p.ndofnodes, b)
    for E in macros_of(etype):
        f = orc.macro_rhs(etype, p.conn, p.dofmap, p.coords, p.cmats, p.phase, E, p.ndofnodes)
        exp = -P.T @ (full @ linear_field(p.coords, E))
        assert np.abs(exp).max() > 0.
        np.testing.assert_allclose(f, exp, rtol=0, atol=1e-12 * np.abs(exp).max())


@pytest.mark.parametrize("etype,n,seed", [(P1, 7, 8), (TET4, 4, 9)])
def test_linear_field_has_the_macro_strain(etype, n, seed):
    """The non-symmetric G of linear_field has strain E under the
    reference-pinned element strain (engineering shear)."""
    p = configs.rve(n, seed, 0.3) if etype == P1 else configs.rve3d(n, seed)
    for E in macros_of(etype):
        eps = orc.element_strain(etype, p.conn, p.coords, linear_field(p.coords, E))
        np.testing.assert_allclose(eps, np.tile(E, (p.conn.shape[0], 1)), rtol=0, atol=1e-15)


@pytest.mark.parametrize("etype,n,seed", [(P1, 8, 10), (TET4, 4, 11)])
def test_strain_of_fluctuation_plus_macro(etype, n, seed):
    p = configs.rve(n, seed, 0.3) if etype == P1 else configs.rve3d(n, seed)
    b = ag.DIM[etype]
    P = full_projection(p.dofmap, p.ndofnodes, b)
    x = np.random.default_rng(seed).normal(0., 1e-3, p.ndofnodes * b)
    for E in macros_of(etype):
        got = orc.element_strain_dofmap(etype, p.conn, p.dofmap, p.coords, x, E)
        exp = orc.element_strain(etype, p.conn, p.coords, P @ x + linear_field(p.coords, E))
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13 * np.abs(exp).max())


# ------------------------------------------------------- laminates
@pytest.mark.parametrize("etype,n,pattern", [
    (P1, 8, [0, 0, 0, 1, 1, 0, 1, 0]),
    (P1, 12, [1, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 0]),
    (TET4, 4, [0, 1, 1, 0]),
    (TET4, 6, [1, 0, 0, 0, 1, 0]),
])
def test_laminate_effective_stiffness_closed_form(etype, n, pattern):
    dim = 2 if etype == P1 else 3
    cm = np.stack([ag.cauchy_green(12e9, 0.3, dim=dim), ag.cauchy_green(60e9, 0.2, dim=dim)])
    p = periodic_problem(etype, n, layer_phases(etype, n, pattern), cm)
    s = 3 if dim == 2 else 6
    unit = [np.eye(s)[k] * 1e-3 for k in range(s)]
    ceff, _ = effective_stiffness(etype, p, unit)
    ceff = ceff.T / 1e-3                                       # column k: stress of unit strain k
    f1 = float(np.mean(pattern))
    t, nn = ([0], [1, 2]) if dim == 2 else ([0, 1, 5], [2, 3, 4])
    exact = laminate_stiffness([cm[0], cm[1]], [1. - f1, f1], t, nn)
    np.testing.assert_allclose(ceff, exact, rtol=0, atol=1e-10 * np.abs(exact).max())
    # the laminate is neither the Voigt nor the Reuss average
    voigt = (1 - f1) * cm[0] + f1 * cm[1]
    assert np.abs(voigt - exact).max() > 1e-2 * np.abs(exact).max()


# ------------------------------------------------------- bounds
@pytest.mark.parametrize("etype,n,seed,frac", [(P1, 16, 21, 0.3), (P1, 20, 22, 0.45), (TET4, 6, 23, 0.3),
                                               (TET4, 6, 24, 0.5)])
def test_voigt_reuss_bracket(etype, n, seed, frac):
    p = configs.rve(n, seed, frac) if etype == P1 else configs.rve3d(n, seed, frac)
    s = 3 if etype == P1 else 6
    assert 0 < p.phase.mean() < 1
    ceff, _ = effective_stiffness(etype, p, [np.eye(s)[k] * 1e-3 for k in range(s)])
    ceff = ceff.T / 1e-3
    np.testing.assert_allclose(ceff, ceff.T, rtol=0, atol=1e-9 * np.abs(ceff).max())
    voigt, reuss = voigt_reuss(etype, p)
    sym = 0.5 * (ceff + ceff.T)
    scale = np.abs(voigt).max()
    lo = np.linalg.eigvalsh(sym - reuss)
    hi = np.linalg.eigvalsh(voigt - sym)
    # strictly inside: two phases of different stiffness, not a laminate
    assert lo.min() > 1e-4 * scale, lo
    assert hi.min() > 1e-4 * scale, hi
    # a sign error in the load would give 2 <C> - C_eff: outside the bracket
    flipped = 2 * voigt - sym
    assert np.linalg.eigvalsh(voigt - flipped).min() < -1e-4 * scale


def test_homogeneous_rve3d_reproduces_c():
    p = configs.rve3d(4, 31)
    C = p.cmats[0]
    p = ag.Problem(ag.TET4, p.coords, p.conn, np.stack([C, C]), phase=p.phase, dofmap=p.dofmap,
                   ndofnodes=p.ndofnodes, bc_dofs=p.bc_dofs, bc_vals=p.bc_vals, macro=p.macro, nrhs=1)
    ceff, _ = effective_stiffness(TET4, p, [np.eye(6)[k] * 1e-3 for k in range(6)])
    np.testing.assert_allclose(ceff.T / 1e-3, C, rtol=0, atol=1e-10 * np.abs(C).max())
