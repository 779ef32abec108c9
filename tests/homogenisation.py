"""Helpers of the homogenisation pins: the periodic projection, linear
fields of a macro strain, the oracle's effective stiffness (scipy direct
solve), element measures, the closed-form laminate stiffness and layered
phase fields.  Shared by the CPU pins of the oracle and the GPU tests."""
import numpy as np
import scipy.sparse.linalg as spla

import amie_gpu as ag
import oracle as orc

P1, TET4 = 1, 3


def full_projection(dofmap, ndofnodes, b):
    """0/1 matrix P: full DOF (node*b + axis) <- periodic DOF."""
    nn = len(dofmap)
    P = np.zeros((nn * b, ndofnodes * b))
    for i in range(nn):
        for a in range(b):
            P[i * b + a, dofmap[i] * b + a] = 1.
    return P


def linear_field(coords, E):
    """u = G x with a NON-symmetric G whose symmetric part (engineering
    shear) is E: all shear carried by one off-diagonal entry."""
    if coords.shape[1] == 2:
        exx, eyy, gxy = E
        G = np.array([[exx, gxy], [0., eyy]])
    else:
        exx, eyy, ezz, gyz, gxz, gxy = E
        G = np.array([[exx, gxy, gxz], [0., eyy, gyz], [0., 0., ezz]])
    return (coords @ G.T).ravel()


def effective_stiffness(etype, p, macros, solutions=False):
    """Average stress per macro strain (one row per entry of macros), by the
    oracle's assembly, load and strain recovery and scipy's direct solve;
    with solutions=True also the periodic fluctuations."""
    A = orc.assemble_periodic(etype, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
    bc = np.arange(ag.DIM[etype], dtype=np.int64)      # fluctuation pinned at DOF node 0
    loads = [orc.macro_rhs(etype, p.conn, p.dofmap, p.coords, p.cmats, p.phase, E, A.nrows) for E in macros]
    orc.apply_dirichlet(A, np.zeros(A.n), bc, np.zeros(len(bc)))
    K = A.to_scipy().tocsc()
    rows, xs = [], []
    for E, f in zip(macros, loads):
        f[bc] = 0.
        x = spla.spsolve(K, f)
        eps = orc.element_strain_dofmap(etype, p.conn, p.dofmap, p.coords, x, E)
        s, _ = orc.average_stress(etype, p.conn, p.coords, p.cmats, p.phase, eps)
        rows.append(s)
        xs.append(x)
    return (np.array(rows), K, xs) if solutions else (np.array(rows), K)


def measures(etype, conn, coords):
    X = coords[conn]
    if etype == P1:
        a, b, c = X[:, 0], X[:, 1], X[:, 2]
        return 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    u, v, w = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]
    return np.abs(np.einsum("ij,ij->i", u, np.cross(v, w))) / 6.


def laminate_stiffness(Cs, fractions, t, n):
    """Closed-form effective stiffness of a laminate with layer normal along
    the Voigt components n (t: the in-plane ones): in-plane strains are the
    macro ones in every layer, the normal tractions are one constant S,
    and the layer-averaged normal strains are the macro ones."""
    s = Cs[0].shape[0]
    Ninv = [np.linalg.inv(C[np.ix_(n, n)]) for C in Cs]
    Am = sum(f * Ni for f, Ni in zip(fractions, Ninv))
    Bm = sum(f * Ni @ C[np.ix_(n, t)] for f, Ni, C in zip(fractions, Ninv, Cs))
    Ceff = np.zeros((s, s))
    for k in range(s):
        E = np.zeros(s)
        E[k] = 1.
        S = np.linalg.solve(Am, E[n] + Bm @ E[t])
        sig = np.zeros(s)
        sig[n] = S
        for f, Ni, C in zip(fractions, Ninv, Cs):
            en = Ni @ (S - C[np.ix_(n, t)] @ E[t])
            sig[t] += f * (C[np.ix_(t, t)] @ E[t] + C[np.ix_(t, n)] @ en)
        Ceff[:, k] = sig
    return Ceff


def periodic_problem(etype, n, phase, cmats, macro=None):
    coords, conn = ag.structured_mesh(etype, n)
    if etype == P1:
        dofmap = ag.periodic_dofmap(n)
        nd = n * n
    else:
        m = n + 1
        i, j, k = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
        dofmap = (((i % n) * n + (j % n)) * n + (k % n)).ravel().astype(np.int32)
        nd = n ** 3
    b = ag.DIM[etype]
    macro = np.eye(3 if b == 2 else 6)[:1] * 1e-3 if macro is None else np.atleast_2d(macro)
    return ag.Problem(etype, coords, conn, cmats, phase=phase, dofmap=dofmap, ndofnodes=nd,
                      bc_dofs=np.arange(b, dtype=np.int64), bc_vals=np.zeros(b), macro=macro,
                      nrhs=macro.shape[0])


def layer_phases(etype, n, pattern):
    """Phase of every element from its layer (element row j in 2D, plane k
    in 3D), the layer's phase pattern[layer]."""
    coords, conn = ag.structured_mesh(etype, n)
    c = coords[conn].mean(axis=1)[:, 1 if etype == P1 else 2]
    layer = np.floor(c * n).astype(int)
    return np.asarray(pattern)[layer].astype(np.int32)



def voigt_reuss(etype, p):
    """Volume-weighted <C> and <C^-1>^-1 of the RVE's elements."""
    w = measures(etype, p.conn, p.coords)
    w = w / w.sum()
    Ce = p.cmats[p.phase]
    voigt = np.einsum("e,eij->ij", w, Ce)
    reuss = np.linalg.inv(np.einsum("e,eij->ij", w, np.linalg.inv(Ce)))
    return voigt, reuss
