"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product (xfem-amie_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P1, P2, TET4 = 1, 2, 3
NPER = {P1: 3, P2: 6, TET4: 4}
DIM = {P1: 2, P2: 2, TET4: 3}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        i64, i32, dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        vp = ctypes.c_void_p
        L.or_set_threads.argtypes = [i32]
        L.or_get_threads.restype = i32
        L.or_cauchy_green.argtypes = [dbl, dbl, i32, i32, vp]
        L.or_orient.argtypes = [i32, vp, vp]
        L.or_element_matrix.argtypes = [i32, vp, vp, vp]
        L.or_pattern.argtypes = [i32, i64, vp, i64, vp, vp]
        L.or_pattern.restype = i64
        L.or_assemble.argtypes = [i32, i64, vp, vp, vp, vp, vp, i64, vp, vp, vp]
        L.or_apply_dirichlet.argtypes = [i32, i64, vp, vp, vp, vp, i64, vp, vp]
        L.or_spmv.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp]
        L.or_dot.argtypes = [vp, vp, i64]
        L.or_dot.restype = dbl
        L.or_inverse_diagonal.argtypes = [i32, i64, vp, vp, vp, vp]
        L.or_inverse_2x2_diagonal.argtypes = [i32, i64, vp, vp, vp, vp]
        L.or_inverse_2x2_diagonal.restype = i32
        L.or_cg_reference.argtypes = [i32, i64, vp, vp, vp, vp, vp, dbl, i64, i32, vp, vp]
        L.or_cg_reference.restype = i32
        L.or_pcg.argtypes = [i32, i64, vp, vp, vp, vp, vp, dbl, i64, i32, vp, vp, vp]
        L.or_pcg.restype = i32
        L.or_element_strain.argtypes = [i32, i64, vp, vp, vp, vp]
        L.or_assemble_dofmap.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp]
        L.or_macro_rhs.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, vp, i64, vp]
        L.or_element_strain_dofmap.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp]
        L.or_damage_update.argtypes = [i64, i32, vp, vp, dbl, dbl, dbl, vp, vp]
        L.or_damage_update.restype = i64
        L.or_average_stress.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, vp]
        L.or_smoothed_strain.argtypes = [i32, i64, vp, vp, vp, vp, dbl, dbl, i32, vp, vp]
        L.or_assemble_fractured.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, dbl, dbl, i64, vp, vp, vp]
        L.or_principal.argtypes = [i32, i32, vp, vp]
        L.or_gauss_fields.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, vp, vp]
        L.or_mazars_score.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, dbl, vp]
        L.or_fiber_pass.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, dbl, dbl, dbl, dbl, i64, vp, vp]
        L.or_fiber_pass.restype = i64
        L.or_newton_pass.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, dbl, dbl, vp]
        L.or_newton_pass.restype = i64
        L.or_renumber.argtypes = [i64, i64, i32, vp, vp]
        L.or_renumber.restype = i64
        L.or_add_midpoints.argtypes = [i64, vp, i64, vp, vp, vp]
        L.or_add_midpoints.restype = i64
        L.or_voxel_owner.argtypes = [i64, vp, i64, vp, vp, i32, i32, i32, vp]
        _LIB = L
    return _LIB


def set_threads(n):
    lib().or_set_threads(n)


def get_threads():
    return lib().or_get_threads()


class one_thread:
    """Context manager: run the oracle with one OpenMP thread."""

    def __enter__(self):
        self.n = get_threads()
        set_threads(1)

    def __exit__(self, *a):
        set_threads(self.n)


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def cauchy_green(E, nu, dim=2, plane_strain=False):
    s = 3 if dim == 2 else 6
    C = np.zeros((s, s))
    lib().or_cauchy_green(E, nu, dim, int(plane_strain), _p(C))
    return C


def orient(etype, coords):
    coords = _c(coords, np.float64)
    perm = np.zeros(NPER[etype], dtype=np.int32)
    lib().or_orient(etype, _p(coords), _p(perm))
    return perm


def element_matrix(etype, coords, C):
    coords = _c(coords, np.float64)
    C = _c(C, np.float64)
    b = DIM[etype]
    n = NPER[etype] * b
    K = np.zeros((n, n))
    lib().or_element_matrix(etype, _p(coords), _p(C), _p(K))
    return K


class BlockCSR:
    """Node-block CSR in the reference layout (padded column-major blocks)."""

    def __init__(self, stride, row_size, column_index, array=None):
        self.stride = stride
        self.row_size = np.ascontiguousarray(row_size, dtype=np.uint32)
        self.column_index = np.ascontiguousarray(column_index, dtype=np.uint32)
        self.nrows = len(self.row_size)
        self.col_len = stride + stride % 2
        nnzb = len(self.column_index)
        self.array = (np.zeros(nnzb * stride * self.col_len) if array is None
                      else np.ascontiguousarray(array, dtype=np.float64))

    @property
    def n(self):
        return self.nrows * self.stride

    def args(self):
        return (self.stride, self.nrows, _p(self.row_size), _p(self.column_index), _p(self.array))

    def to_dense(self):
        b, cl = self.stride, self.col_len
        A = np.zeros((self.n, self.n))
        start = np.concatenate([[0], np.cumsum(self.row_size, dtype=np.int64)])
        blocks = self.array.reshape(-1, b, cl)  # [blk][col][row]
        for r in range(self.nrows):
            for q in range(start[r], start[r + 1]):
                c = int(self.column_index[q])
                A[r * b:(r + 1) * b, c * b:(c + 1) * b] = blocks[q, :, :b].T
        return A

    def to_scipy(self):
        import scipy.sparse as sp
        b, cl = self.stride, self.col_len
        start = np.concatenate([[0], np.cumsum(self.row_size, dtype=np.int64)])
        blocks = self.array.reshape(-1, b, cl)[:, :, :b].transpose(0, 2, 1)
        return sp.bsr_matrix((np.ascontiguousarray(blocks), self.column_index.astype(np.int64), start),
                             shape=(self.n, self.n)).tocsr()


def pattern(etype, conn, nnodes):
    conn = _c(conn, np.int32)
    nelem = conn.shape[0]
    nnzb = lib().or_pattern(etype, nelem, _p(conn), nnodes, None, None)
    rs = np.zeros(nnodes, dtype=np.uint32)
    ci = np.zeros(nnzb, dtype=np.uint32)
    lib().or_pattern(etype, nelem, _p(conn), nnodes, _p(rs), _p(ci))
    return rs, ci


def assemble(etype, conn, coords, cmats, phase=None, damage=None):
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    nnodes = coords.shape[0]
    rs, ci = pattern(etype, conn, nnodes)
    A = BlockCSR(DIM[etype], rs, ci)
    lib().or_assemble(etype, conn.shape[0], _p(conn), _p(coords), _p(cmats), _p(phase), _p(damage),
                      A.nrows, _p(A.row_size), _p(A.column_index), _p(A.array))
    return A


def assemble_periodic(etype, conn, dofmap, coords, cmats, phase=None, damage=None):
    """Stiffness through a DOF map (periodic condensation)."""
    conn = _c(conn, np.int32)
    dofmap = _c(dofmap, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    nrows = int(dofmap.max()) + 1
    rs, ci = pattern(etype, dofmap[conn], nrows)
    A = BlockCSR(DIM[etype], rs, ci)
    lib().or_assemble_dofmap(etype, conn.shape[0], _p(conn), _p(dofmap), _p(coords), _p(cmats), _p(phase),
                             _p(damage), A.nrows, _p(A.row_size), _p(A.column_index), _p(A.array))
    return A


def macro_rhs(etype, conn, dofmap, coords, cmats, phase, macro, nrows, damage=None):
    conn = _c(conn, np.int32)
    dofmap = _c(dofmap, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    macro = _c(macro, np.float64)
    f = np.zeros(nrows * DIM[etype])
    lib().or_macro_rhs(etype, conn.shape[0], _p(conn), _p(dofmap), _p(coords), _p(cmats), _p(phase), _p(damage),
                       _p(macro), nrows, _p(f))
    return f


def element_strain_dofmap(etype, conn, dofmap, coords, u, macro=None):
    conn = _c(conn, np.int32)
    dofmap = _c(dofmap, np.int32)
    coords = _c(coords, np.float64)
    u = _c(u, np.float64)
    macro = _c(macro, np.float64)
    s = 3 if DIM[etype] == 2 else 6
    out = np.zeros((conn.shape[0], s))
    lib().or_element_strain_dofmap(etype, conn.shape[0], _p(conn), _p(dofmap), _p(coords), _p(u), _p(macro),
                                   _p(out))
    return out


def apply_dirichlet(A, rhs, dofs, vals):
    dofs = _c(dofs, np.int64)
    vals = _c(vals, np.float64)
    lib().or_apply_dirichlet(*A.args(), _p(rhs), len(dofs), _p(dofs), _p(vals))


def spmv(A, x, b=None):
    x = _c(x, np.float64)
    b = _c(b, np.float64)
    y = np.zeros(A.n)
    lib().or_spmv(*A.args(), _p(x), _p(b), _p(y))
    return y


def dot(a, b):
    a = _c(a, np.float64)
    b = _c(b, np.float64)
    return lib().or_dot(_p(a), _p(b), len(a))


def inverse_diagonal(A):
    d = np.zeros(A.n)
    lib().or_inverse_diagonal(*A.args(), _p(d))
    return d


def inverse_2x2_diagonal(A):
    blk = np.zeros((A.n // 2, 4))
    if lib().or_inverse_2x2_diagonal(*A.args(), _p(blk)) != 0:
        raise ValueError("odd number of dofs")
    return blk


def cg_reference(A, b, x0=None, eps=1e-10, maxit=-1, nssor=32):
    b = _c(b, np.float64)
    x0 = _c(x0, np.float64)
    x = np.zeros(A.n)
    nit = np.zeros(1, dtype=np.int64)
    ok = lib().or_cg_reference(*A.args(), _p(b), _p(x0), eps, maxit, nssor, _p(x), _p(nit))
    return x, int(nit[0]), ok


def pcg(A, b, x0=None, rtol=1e-10, maxit=100000, precond=0):
    b = _c(b, np.float64)
    x0 = _c(x0, np.float64)
    x = np.zeros(A.n)
    nit = np.zeros(1, dtype=np.int64)
    rr = np.zeros(1)
    ok = lib().or_pcg(*A.args(), _p(b), _p(x0), rtol, maxit, precond, _p(x), _p(nit), _p(rr))
    return x, int(nit[0]), float(rr[0]), ok


NGP = {P1: 1, P2: 4, TET4: 1}


def gauss_fields(etype, conn, coords, cmats, u, phase=None, damage=None):
    """Strain and stress at every Gauss point: (nelem, ngp, 3|6) each."""
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    u = _c(u, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    s = 3 if DIM[etype] == 2 else 6
    eps = np.zeros((conn.shape[0], NGP[etype], s))
    sig = np.zeros_like(eps)
    lib().or_gauss_fields(etype, conn.shape[0], _p(conn), _p(coords), _p(cmats), _p(phase), _p(damage), _p(u),
                          _p(eps), _p(sig))
    return eps, sig


def element_strain(etype, conn, coords, u):
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    u = _c(u, np.float64)
    s = 3 if DIM[etype] == 2 else 6
    out = np.zeros((conn.shape[0], s))
    lib().or_element_strain(etype, conn.shape[0], _p(conn), _p(coords), _p(u), _p(out))
    return out


def damage_update(strain, kappa, d, d0=None, kappa0=1e-4, kappaf=1e-3, dmax=1 - 1e-6):
    strain = _c(strain, np.float64)
    d0 = _c(d0, np.float64)
    dim = 2 if strain.shape[1] == 3 else 3
    grown = lib().or_damage_update(strain.shape[0], dim, _p(strain), _p(d0), kappa0, kappaf, dmax,
                                   _p(kappa), _p(d))
    return int(grown)


def average_stress(etype, conn, coords, cmats, phase, strain):
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    phase = _c(phase, np.int32)
    strain = _c(strain, np.float64)
    s = strain.shape[1]
    out = np.zeros(s)
    vol = np.zeros(1)
    lib().or_average_stress(etype, conn.shape[0], _p(conn), _p(coords), _p(cmats), _p(phase), _p(strain),
                            _p(out), _p(vol))
    return out, float(vol[0])


QUARTIC_COMPACT, GAUSSIAN_NONCOMPACT = 0, 1


def smoothed_strain(etype, conn, coords, strain, radius, kernel=QUARTIC_COMPACT, phase=None, damage=None,
                    threshold=1.0):
    """Non-local element strains (linear elements): the reference's
    getSmoothedFields over its generateCache neighbourhoods."""
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    strain = _c(strain, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    out = np.zeros_like(strain)
    lib().or_smoothed_strain(etype, conn.shape[0], _p(conn), _p(coords), _p(phase), _p(damage), threshold, radius,
                             kernel, _p(strain), _p(out))
    return out


def renumber(conn, nnodes=None):
    """FeatureTree::renumber: first-touch ids over the elements in order
    (-1 for nodes no element names) and their count."""
    conn = _c(conn, np.int32)
    n = int(conn.max()) + 1 if nnodes is None else int(nnodes)
    ids = np.empty(n, dtype=np.int32)
    count = lib().or_renumber(n, conn.shape[0], conn.shape[1], _p(conn), _p(ids))
    return ids, int(count)


def add_midpoints(xy, conn3):
    """DelaunayTree::addSharedNodes(1, 1): P2 triangles [v0, m01, v1, m12,
    v2, m20] and the new mid-edge nodes' coordinates (numbered from
    len(xy) in the order they are made)."""
    xy = _c(xy, np.float64)
    conn3 = _c(conn3, np.int32)
    conn6 = np.empty((conn3.shape[0], 6), dtype=np.int32)
    mid = np.empty((3 * conn3.shape[0], 2))
    n = lib().or_add_midpoints(xy.shape[0], _p(xy), conn3.shape[0], _p(conn3), _p(conn6), _p(mid))
    return conn6, mid[:n].copy()


def voxel_owner(xyz, conn, box, n):
    """VoxelWriter sampling: owning tetrahedron of every voxel of an n^3 grid
    (-1 where none), index k + n*j + n*n*i."""
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int32)
    box = _c(box, np.float64)
    owner = np.empty(n ** 3, dtype=np.int32)
    lib().or_voxel_owner(xyz.shape[0], _p(xyz), conn.shape[0], _p(conn), _p(box), n, n, n, _p(owner))
    return owner



# ------------------------------------------------------- event-driven damage
# FeatureTree's damage iteration with StiffnessAndFracture + NonLocalMazars +
# FiberBasedIsotropicLinearDamage (its default damage model): see
# or_mazars_score and or_fiber_pass in amie_oracle.cpp.
FIBRE_FRACTION, FRACTURED_AT, RESIDUAL, SCORE_TOLERANCE, FIRST_DEFERRED = 0.1, 0.6, 1e-4, 2e-3, 65


def assemble_fractured(etype, conn, coords, cmats, phase, damage, fractured_at=FRACTURED_AT, residual=RESIDUAL):
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    phase = _c(phase, np.int32)
    damage = _c(damage, np.float64)
    rs, ci = pattern(etype, conn, coords.shape[0])
    A = BlockCSR(DIM[etype], rs, ci)
    lib().or_assemble_fractured(etype, conn.shape[0], _p(conn), None, _p(coords), _p(cmats), _p(phase), _p(damage),
                                fractured_at, residual, A.nrows, _p(A.row_size), _p(A.column_index), _p(A.array))
    return A


def principal(v, doubled):
    """toPrincipal of one stress (doubled False) or engineering strain."""
    v = _c(np.asarray(v, dtype=np.float64), np.float64)
    dim = 2 if v.shape[0] == 3 else 3
    out = np.zeros(dim)
    lib().or_principal(dim, int(bool(doubled)), _p(v), _p(out))
    return out


def mazars_score(dim, phase, crit, prm, cmats, sbar, damage, fractured_at=FRACTURED_AT):
    """NonLocalMazars fracture score of every element from its smoothed
    strain; prm[phase] = (threshold, E, nu, Gf)."""
    phase = _c(phase, np.int32)
    crit = _c(np.asarray(crit), np.int32)
    prm = _c(np.asarray(prm, dtype=np.float64).reshape(-1, 4), np.float64)
    cmats = _c(cmats, np.float64)
    sbar = _c(sbar, np.float64)
    damage = _c(damage, np.float64)
    out = np.zeros(sbar.shape[0])
    lib().or_mazars_score(dim, sbar.shape[0], _p(phase), _p(crit), _p(prm), _p(cmats), _p(sbar), _p(damage),
                          fractured_at, _p(out))
    return out


def fiber_pass(etype, conn, coords, phase, crit, score, damage, radius, fraction=FIBRE_FRACTION,
               fractured_at=FRACTURED_AT, tol=SCORE_TOLERANCE, first_deferred=FIRST_DEFERRED, nbr_max=None):
    """One FiberBasedIsotropicLinearDamage pass; updates damage in place and
    returns the mask of the elements it incremented (nbr_max, if given,
    receives each criterion element's neighbourhood maximum)."""
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    phase = _c(phase, np.int32)
    crit = _c(np.asarray(crit), np.int32)
    score = _c(score, np.float64)
    assert damage.dtype == np.float64 and damage.flags.c_contiguous
    changed = np.zeros(conn.shape[0], dtype=np.int32)
    lib().or_fiber_pass(etype, conn.shape[0], _p(conn), _p(coords), _p(phase), _p(crit), _p(score), _p(damage),
                        radius, fraction, fractured_at, tol, first_deferred, _p(changed), _p(nbr_max))
    return changed.astype(bool)


def newton_pass(etype, conn, coords, phase, crit, score, damage, radius, tol=SCORE_TOLERANCE):
    """One IsotropicLinearDamage (Newton) pass; updates damage in place and
    returns the mask of the elements it changed."""
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    phase = _c(phase, np.int32)
    crit = _c(np.asarray(crit), np.int32)
    score = _c(score, np.float64)
    assert damage.dtype == np.float64 and damage.flags.c_contiguous
    changed = np.zeros(conn.shape[0], dtype=np.int32)
    lib().or_newton_pass(etype, conn.shape[0], _p(conn), _p(coords), _p(phase), _p(crit), _p(score), _p(damage),
                         radius, tol, _p(changed))
    return changed.astype(bool)


MAX_ITERATIONS_PER_STEP = 200   # FeatureTree::maxitPerStep (features/features.cpp:202)


def event_damage(etype, conn, coords, cmats, phase, crit, prm, bc_dofs, bc_vals, damage, radius,
                 kernel=QUARTIC_COMPACT, rtol=1e-12, max_passes=100000, first_deferred=FIRST_DEFERRED,
                 model="fibre"):
    """One FeatureTree::step of the event-driven damage (features/features.cpp:
    6150-6373): solve with the current damage (fractured elements at the
    residual stiffness), score every element on its smoothed strain, run one
    damage pass; model "fibre" (FiberBasedIsotropicLinearDamage, the default
    of StiffnessAndFracture) or "newton" (IsotropicLinearDamage).  The step
    ends at a pass that changes nothing and leaves no element scoring above
    0, or at such a pass once more than maxitPerStep passes have run.
    Returns (damage, displacements, scores of the last pass, passes)."""
    d = np.array(damage, dtype=np.float64)
    dim = DIM[etype]
    frac_at = FRACTURED_AT if model == "fibre" else np.inf
    passes = 0
    while True:
        A = assemble_fractured(etype, conn, coords, cmats, phase, d, fractured_at=frac_at)
        b = np.zeros(A.n)
        apply_dirichlet(A, b, bc_dofs, bc_vals)
        x = pcg(A, b, rtol=rtol, maxit=100000)[0]
        eps = element_strain(etype, conn, coords, x)
        sbar = smoothed_strain(etype, conn, coords, eps, radius, kernel, phase, d, frac_at)
        score = mazars_score(dim, phase, crit, prm, cmats, sbar, d, fractured_at=frac_at)
        if model == "fibre":
            changed = fiber_pass(etype, conn, coords, phase, crit, score, d, radius, first_deferred=first_deferred)
        else:
            changed = newton_pass(etype, conn, coords, phase, crit, score, d, radius)
        passes += 1
        crit_e = np.asarray(crit)[np.asarray(phase)]
        met = bool((score[crit_e != 0] > 0).any())
        if passes >= max_passes:
            return d, x, score, passes
        if not changed.any() and (not met or passes > MAX_ITERATIONS_PER_STEP):
            return d, x, score, passes
