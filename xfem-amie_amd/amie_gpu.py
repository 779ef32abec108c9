"""ctypes binding of libamie_gpu.so (include/amie_gpu.h).

The binding is thin: numpy arrays in, numpy arrays out, every call checked.
There is no fallback: when the HIP library is missing or the GPU is not a
gfx950, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamie_gpu.so")

P1, P2, TET4 = 1, 2, 3
FP64, MIXED, MIXED64, MIXED_K64, MIXED_SPLIT, MIXED_FIX48 = 0, 1, 2, 3, 4, 5
DAMAGE_FIBRE, DAMAGE_ISOTROPIC_LINEAR = 0, 1   # event-damage models (amie_event_params.damage_model)
JACOBI, BLOCK_JACOBI = 0, 1
REFINE_REPLACE, REFINE_RESTART = 0, 1
ASSEMBLE_GATHER, ASSEMBLE_COLOURED = 0, 1
NPER = {P1: 3, P2: 6, TET4: 4}
DIM = {P1: 2, P2: 2, TET4: 3}
VOIGT = {P1: 3, P2: 3, TET4: 6}

c_i32, c_i64, c_dbl, c_u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
vp = ctypes.c_void_p


class ProblemDesc(ctypes.Structure):
    _fields_ = [("elem_type", c_i32), ("nnodes", c_i64), ("nelem", c_i64), ("coords", vp), ("conn", vp),
                ("dofmap", vp), ("ndofnodes", c_i64), ("nphase", c_i32), ("cmat", vp), ("phase", vp),
                ("damage", vp), ("nbc", c_i64), ("bc_dofs", vp), ("bc_vals", vp), ("nrhs", c_i32),
                ("forces", vp), ("macro_strain", vp)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("precision", c_i32), ("precond", c_i32), ("refine", c_i32), ("rtol", c_dbl),
                ("refine_drop", c_dbl), ("ckpt_interval", c_i64), ("maxit", c_i64), ("use_x0", c_i32), ("use_graph", c_i32),
                ("chunk", c_i32), ("time_kernels", c_i32), ("escalate", c_i32),
                ("max_grid", c_i32), ("ref_eps", c_dbl), ("autotune", c_i32), ("stall_iters", c_i64),
                ("persistent", c_i32)]


class SolveStats(ctypes.Structure):
    _fields_ = [("status", c_i32), ("nproblems_converged", c_i32), ("iterations_max", c_i64),
                ("iterations_sum", c_i64), ("refinements", c_i64), ("relres_max", c_dbl), ("solve_ms", c_dbl),
                ("spmv_kernel_ms", c_dbl), ("spmv_launches", c_i64), ("spmv_bytes", c_dbl),
                ("nproblems_escalated", c_i32), ("nproblems_handover", c_i32), ("iterations_handover", c_i64),
                ("update_kernel_ms", c_dbl), ("update_launches", c_i64), ("update_bytes", c_dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DamageParams(ctypes.Structure):
    _fields_ = [("kappa0", c_dbl), ("kappaf", c_dbl), ("dmax", c_dbl)]


class EventParams(ctypes.Structure):
    _fields_ = [("nphase", c_i32), ("criterion", vp), ("mazars", vp), ("radius", c_dbl), ("kernel", c_i32),
                ("fibre_fraction", c_dbl), ("fractured_at", c_dbl), ("residual", c_dbl), ("score_tolerance", c_dbl),
                ("first_deferred", c_i64), ("max_passes", c_i64), ("damage_model", c_i32),
                ("max_passes_per_step", c_i64)]


class EventStats(ctypes.Structure):
    _fields_ = [("passes", c_i64), ("cg_iterations", c_i64), ("events", c_i64), ("fractured", c_i64),
                ("status", c_i32), ("converged", c_i32), ("solve_ms", c_dbl), ("damage_ms", c_dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RequestResult(ctypes.Structure):
    _fields_ = [("ticket", c_u64), ("status", c_i32), ("iterations", c_i32), ("relres", c_dbl),
                ("submit_ms", c_dbl), ("start_ms", c_dbl), ("done_ms", c_dbl)]


# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "amie_gpu_create": (c_i32, [vp, c_i32]),
    "amie_gpu_destroy": (None, [vp]),
    "amie_gpu_last_error": (ctypes.c_char_p, []),
    "amie_gpu_device_info": (c_i32, [vp, vp, vp, vp]),
    "amie_gpu_default_opts": (None, [vp]),
    "amie_mesh_orient": (c_i32, [c_i32, c_i64, vp, vp]),
    "amie_mesh_structured_size": (c_i32, [c_i32, c_i32, vp, vp]),
    "amie_mesh_structured": (c_i32, [c_i32, c_i32, vp, vp]),
    "amie_mesh_periodic_dofmap": (c_i32, [c_i32, vp]),
    "amie_inclusions_2d": (c_i32, [c_u64, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, vp, vp, vp]),
    "amie_phase_by_centroid": (c_i32, [c_i32, c_i64, vp, vp, c_i64, vp, vp]),
    "amie_damage_field_3d": (c_i32, [c_u64, c_i32, c_dbl, c_dbl, c_dbl, vp]),
    "amie_cauchy_green": (None, [c_dbl, c_dbl, c_i32, vp]),
    "amie_element_matrix": (c_i32, [c_i32, vp, vp, vp]),
    "amie_gpu_element_matrices": (c_i32, [vp, c_i32, c_i64, vp, vp, vp]),
    "amie_gpu_renumber": (c_i32, [vp, c_i64, c_i64, c_i32, vp, vp, vp, vp]),
    "amie_write_triangles": (c_i32, [ctypes.c_char_p, c_i64, c_i32, vp, c_i64, vp, vp, c_i32, vp, vp]),
    "amie_write_triangles_bin": (c_i32, [ctypes.c_char_p, c_i64, c_i32, vp, c_i64, vp, vp, c_i32, vp, vp]),
    "amie_multi_triangles_begin": (c_i32, [ctypes.c_char_p]),
    "amie_multi_triangles_append": (c_i32, [ctypes.c_char_p, ctypes.c_char_p, c_i32, c_i64, c_i32, vp, c_i64, vp, vp,
                                            c_i32, vp, vp]),
    "amie_gpu_voxelize": (c_i32, [vp, c_i64, vp, c_i64, vp, c_i32, vp, vp, c_i32, c_i32, c_i32, vp, vp]),
    "amie_gpu_voxelize_noalloc": (c_i32, [vp, vp, c_i64, vp, c_i64, vp, c_i32, vp, vp, c_i32, c_i32, c_i32, vp]),
    "amie_write_voxels": (c_i32, [ctypes.c_char_p, c_i32, c_i32, c_i32, c_i32, vp]),
    "amie_place_inclusions_2d": (c_i32, [c_dbl, c_dbl, c_dbl, c_dbl, c_i64, vp, c_dbl, c_i64, c_dbl, c_u64,
                                         vp, vp, vp]),
    "amie_gpu_add_midpoints": (c_i32, [vp, c_i64, vp, c_i64, vp, vp, vp, vp, vp]),
    "amie_gpu_stream_copy": (c_i32, [vp, ctypes.c_size_t, c_i32, vp]),
    "amie_gpu_batch_create": (c_i32, [vp, c_i32, vp, vp]),
    "amie_gpu_batch_from_bcsr": (c_i32, [vp, c_i32, c_i64, vp, vp, vp, vp, vp]),
    "amie_gpu_batch_destroy": (None, [vp]),
    "amie_gpu_batch_update_bcsr": (c_i32, [vp, vp, vp]),
    "amie_gpu_batch_bytes": (ctypes.c_size_t, [vp]),
    "amie_gpu_batch_assemble": (c_i32, [vp]),
    "amie_gpu_batch_set_assembly": (c_i32, [vp, c_i32]),
    "amie_gpu_batch_solve": (c_i32, [vp, vp, vp]),
    "amie_gpu_batch_problem_info": (c_i32, [vp, c_i32, vp, vp, vp, vp]),
    "amie_gpu_batch_get_solution": (c_i32, [vp, c_i32, vp]),
    "amie_gpu_batch_set_solution": (c_i32, [vp, c_i32, vp]),
    "amie_gpu_batch_get_nnzb": (c_i32, [vp, c_i32, vp, vp]),
    "amie_gpu_batch_get_matrix": (c_i32, [vp, c_i32, vp, vp, vp, vp]),
    "amie_gpu_batch_get_element_blocks": (c_i32, [vp, c_i32, vp]),
    "amie_gpu_batch_spmv": (c_i32, [vp, c_i32, c_i32, vp, vp]),
    "amie_gpu_batch_time_spmv": (c_i32, [vp, c_i32, c_i32, c_i32, vp, vp]),
    "amie_gpu_batch_autotune": (c_i32, [vp, c_i32, c_i32, c_i32, vp, vp, vp, vp]),
    "amie_gpu_batch_set_nonlocal": (c_i32, [vp, c_dbl, c_i32, c_dbl]),
    "amie_gpu_batch_smoothed_strain": (c_i32, [vp, c_i32, c_i32, vp]),
    "amie_gpu_batch_set_spmv_kernel": (c_i32, [vp, c_i32, c_i32, c_i32, c_i32, c_i32]),
    "amie_gpu_batch_element_strain": (c_i32, [vp, c_i32, c_i32, vp]),
    "amie_gpu_batch_average_stress": (c_i32, [vp, vp]),
    "amie_gpu_batch_damage_step": (c_i32, [vp, vp, vp]),
    "amie_gpu_default_event_params": (None, [vp]),
    "amie_gpu_batch_gauss_fields": (c_i32, [vp, c_i32, c_i32, vp, vp]),
    "amie_gpu_batch_event_step": (c_i32, [vp, vp, vp, vp]),
    "amie_gpu_batch_get_scores": (c_i32, [vp, c_i32, vp]),
    "amie_gpu_batch_get_damage": (c_i32, [vp, c_i32, vp, vp]),
    "amie_gpu_batch_set_damage": (c_i32, [vp, c_i32, vp]),
    "amie_gpu_batch_scale_bc": (c_i32, [vp, c_dbl]),
    "amie_gpu_server_create": (c_i32, [vp, ctypes.c_size_t, c_i32, vp, vp]),
    "amie_gpu_server_destroy": (None, [vp]),
    "amie_gpu_server_submit": (c_i32, [vp, vp, c_u64, c_dbl, vp]),
    "amie_gpu_server_step": (c_i32, [vp, c_dbl, vp, vp]),
    "amie_gpu_server_poll": (c_i32, [vp, vp, c_i32, vp]),
    "amie_gpu_server_get_solution": (c_i32, [vp, c_u64, vp, c_i64]),
    "amie_gpu_server_submit_step": (c_i32, [vp, vp, c_u64, vp, c_dbl, vp]),
    "amie_gpu_server_release": (c_i32, [vp, c_u64]),
    "amie_gpu_server_request_bytes": (c_i32, [vp, vp]),
    "amie_gpu_server_get_damage": (c_i32, [vp, c_u64, vp, vp, c_i64]),
    "amie_gpu_batch_set_bc_values": (c_i32, [vp, c_i32, c_i64, vp, vp]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libamie_gpu.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class AmieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amie error {code}: {msg}")
        self.code = code


def check(rc, allow=(0,)):
    if rc not in allow:
        raise AmieError(rc, lib().amie_gpu_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------ mesh helpers
def cauchy_green(E, nu, dim=2):
    s = 3 if dim == 2 else 6
    C = np.zeros((s, s))
    lib().amie_cauchy_green(E, nu, dim, _p(C))
    return C


def element_matrix(etype, coords, C):
    coords = _c(coords, np.float64)
    C = _c(C, np.float64)
    n = NPER[etype] * DIM[etype]
    K = np.zeros((n, n))
    check(lib().amie_element_matrix(etype, _p(coords), _p(C), _p(K)))
    return K


def structured_mesh(etype, n):
    nn, ne = c_i64(), c_i64()
    check(lib().amie_mesh_structured_size(etype, n, ctypes.byref(nn), ctypes.byref(ne)))
    coords = np.zeros((nn.value, DIM[etype]))
    conn = np.zeros((ne.value, NPER[etype]), dtype=np.int32)
    check(lib().amie_mesh_structured(etype, n, _p(coords), _p(conn)))
    return coords, conn


def orient(etype, conn, coords):
    conn = np.array(conn, dtype=np.int32, order="C")
    coords = _c(coords, np.float64)
    check(lib().amie_mesh_orient(etype, conn.shape[0], _p(conn), _p(coords)))
    return conn


def periodic_dofmap(n):
    m = np.zeros((n + 1) * (n + 1), dtype=np.int32)
    check(lib().amie_mesh_periodic_dofmap(n, _p(m)))
    return m


def inclusions_2d(seed, target, rmin, rmax, gap, max_circles=200000):
    xyr = np.zeros((max_circles, 3))
    nc, frac = c_i64(), c_dbl()
    check(lib().amie_inclusions_2d(seed, target, rmin, rmax, gap, max_circles, _p(xyr), ctypes.byref(nc),
                                   ctypes.byref(frac)))
    return xyr[:nc.value].copy(), frac.value


def phase_by_centroid(etype, conn, coords, xyr):
    conn = _c(conn, np.int32)
    coords = _c(coords, np.float64)
    xyr = _c(xyr, np.float64)
    ph = np.zeros(conn.shape[0], dtype=np.int32)
    check(lib().amie_phase_by_centroid(etype, conn.shape[0], _p(conn), _p(coords), xyr.shape[0], _p(xyr), _p(ph)))
    return ph


def damage_field_3d(seed, n, ell, lo, hi):
    d = np.zeros(6 * n ** 3)
    check(lib().amie_damage_field_3d(seed, n, ell, lo, hi, _p(d)))
    return d


# --------------------------------------------------------------- problems
class Problem:
    """Owns the numpy arrays behind one amie_problem_desc."""

    def __init__(self, etype, coords, conn, cmats, phase=None, damage=None, bc_dofs=None, bc_vals=None,
                 forces=None, macro=None, dofmap=None, ndofnodes=None, nrhs=1):
        self.etype = etype
        self.coords = _c(coords, np.float64)
        self.conn = _c(conn, np.int32)
        self.cmats = _c(cmats, np.float64)
        self.phase = _c(phase, np.int32)
        self.damage = _c(damage, np.float64)
        self.bc_dofs = _c(bc_dofs if bc_dofs is not None else np.zeros(0), np.int64)
        self.bc_vals = _c(bc_vals if bc_vals is not None else np.zeros(0), np.float64)
        self.forces = _c(forces, np.float64)
        self.macro = _c(macro, np.float64)
        self.dofmap = _c(dofmap, np.int32)
        self.nrhs = nrhs
        self.ndofnodes = ndofnodes if ndofnodes is not None else (
            int(self.dofmap.max()) + 1 if self.dofmap is not None else self.coords.shape[0])
        d = ProblemDesc()
        d.elem_type = etype
        d.nnodes = self.coords.shape[0]
        d.nelem = self.conn.shape[0]
        d.coords = _p(self.coords)
        d.conn = _p(self.conn)
        d.dofmap = _p(self.dofmap)
        d.ndofnodes = self.ndofnodes
        d.nphase = self.cmats.shape[0] if self.cmats.ndim == 3 else 1
        d.cmat = _p(self.cmats)
        d.phase = _p(self.phase)
        d.damage = _p(self.damage)
        d.nbc = len(self.bc_dofs)
        d.bc_dofs = _p(self.bc_dofs)
        d.bc_vals = _p(self.bc_vals)
        d.nrhs = nrhs
        d.forces = _p(self.forces)
        d.macro_strain = _p(self.macro)
        self.desc = d

    @property
    def ndof(self):
        return self.ndofnodes * DIM[self.etype]


class Context:
    def __init__(self, device=0):
        self.h = vp()
        check(lib().amie_gpu_create(ctypes.byref(self.h), device))

    def info(self):
        f, t, cu = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        check(lib().amie_gpu_device_info(self.h, ctypes.byref(f), ctypes.byref(t), ctypes.byref(cu)))
        return f.value, t.value, cu.value

    def stream_copy_gbs(self, nbytes=1 << 30, iters=20):
        """Achievable HBM bandwidth of a plain copy (read + write bytes / s)."""
        ms = c_dbl()
        check(lib().amie_gpu_stream_copy(self.h, nbytes, iters, ctypes.byref(ms)))
        return 2 * nbytes / (ms.value * 1e-3) / 1e9

    def close(self):
        if self.h:
            lib().amie_gpu_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def element_matrices_gpu(ctx, etype, coords, cmats):
    """coords: nelem x nper x dim, cmats: nelem x s x s -> nelem x N x N."""
    coords = _c(coords, np.float64)
    cmats = _c(cmats, np.float64)
    ne = coords.shape[0]
    n = NPER[etype] * DIM[etype]
    K = np.zeros((ne, n, n))
    check(lib().amie_gpu_element_matrices(ctx.h, etype, ne, _p(coords), _p(cmats), _p(K)))
    return K


def renumber_gpu(ctx, conn, nnodes, timed=False):
    """FeatureTree::renumber on the device: first-touch ids over the elements
    in order (-1 for nodes no element names), their count, and (timed) the
    device milliseconds."""
    conn = _c(conn, np.int32)
    ids = np.empty(int(nnodes), dtype=np.int32)
    cnt, ms = c_i64(), c_dbl()
    check(lib().amie_gpu_renumber(ctx.h, int(nnodes), conn.shape[0], conn.shape[1], _p(conn), _p(ids),
                                  ctypes.byref(cnt), ctypes.byref(ms) if timed else None))
    return (ids, cnt.value, ms.value) if timed else (ids, cnt.value)


def add_midpoints_gpu(ctx, xy, conn3, timed=False):
    """DelaunayTree::addSharedNodes(1, 1) on the device: P2 triangles [v0,
    m01, v1, m12, v2, m20], the new mid-edge nodes' coordinates (numbered
    from len(xy)), and (timed) the device milliseconds."""
    xy = _c(xy, np.float64)
    conn3 = _c(conn3, np.int32)
    conn6 = np.empty((conn3.shape[0], 6), dtype=np.int32)
    mid = np.empty((3 * conn3.shape[0], 2))
    n, ms = c_i64(), c_dbl()
    check(lib().amie_gpu_add_midpoints(ctx.h, xy.shape[0], _p(xy), conn3.shape[0], _p(conn3), _p(conn6), _p(mid),
                                       ctypes.byref(n), ctypes.byref(ms) if timed else None))
    out = (conn6, mid[:n.value].copy())
    return out + (ms.value,) if timed else out


def place_inclusions_2d(radii, width=1.0, height=1.0, center=(0.5, 0.5), min_dist=0.0, tries=6400,
                        orientation=np.pi, seed=1):
    """The reference's placement2D for circles in a rectangle: indices of
    the placed radii, and their centres and final radii (k x 3)."""
    radii = _c(radii, np.float64)
    n = radii.shape[0]
    idx = np.zeros(max(1, n), dtype=np.int64)
    xyr = np.zeros((max(1, n), 3))
    k = c_i64()
    check(lib().amie_place_inclusions_2d(width, height, center[0], center[1], n, _p(radii), min_dist, tries,
                                         orientation, seed, _p(idx), _p(xyr), ctypes.byref(k)))
    return idx[:k.value].copy(), xyr[:k.value].copy()


def voxelize_gpu(ctx, xyz, conn, field, box, n, timed=False):
    """VoxelWriter sampling on the device: field (ntet x ncomp, constant per
    Tet4) on an n^3 grid over box (cx cy cz sx sy sz) -> ncomp x n^3 values
    (-1e9 where no tetrahedron owns the voxel), and (timed) device ms."""
    xyz = _c(xyz, np.float64)
    conn = _c(conn, np.int32)
    field = _c(np.atleast_2d(np.asarray(field, dtype=np.float64).T).T, np.float64)
    box = _c(box, np.float64)
    ncomp = field.shape[1]
    values = np.empty((ncomp, n ** 3))
    ms = c_dbl()
    check(lib().amie_gpu_voxelize(ctx.h, xyz.shape[0], _p(xyz), conn.shape[0], _p(conn), ncomp, _p(field), _p(box),
                                  n, n, n, _p(values), ctypes.byref(ms) if timed else None))
    return (values, ms.value) if timed else values


def write_voxels(base, values, n):
    """VoxelWriter files <base>_<n>_<c>.vox from ncomp x n^3 values."""
    values = _c(np.atleast_2d(values), np.float64)
    check(lib().amie_write_voxels(str(base).encode(), n, n, n, values.shape[0], _p(values)))


class MultiTriangleWriter:
    """MultiTriangleWriter: an index file `head` and one TriangleWriter file
    <base>_<step> per append."""

    def __init__(self, head, base):
        self.head, self.base, self.step = str(head), str(base), 0
        check(lib().amie_multi_triangles_begin(self.head.encode()))

    def append(self, conn, xy, u, fields=()):
        self.step += 1
        conn, xy, u, sizes, ptrs, keep = _triangle_args(conn, xy, u, fields)
        check(lib().amie_multi_triangles_append(self.head.encode(), self.base.encode(), self.step, conn.shape[0],
                                                conn.shape[1], _p(conn), xy.shape[0], _p(xy), _p(u), len(keep),
                                                _p(sizes), ptrs))


def _triangle_args(conn, xy, u, fields):
    conn = _c(conn, np.int32)
    xy = _c(xy, np.float64)
    u = np.asarray(u, dtype=np.float64)
    if u.ndim != 3:
        n = conn.shape[1] // 3
        u = u.reshape(-1, 2)[conn[:, [0, n, 2 * n]]]
    u = np.ascontiguousarray(u)
    fs = []
    for f in fields:
        f = np.asarray(f, dtype=np.float64)
        if f.ndim == 2:
            f = np.repeat(f[:, None, :], 3, axis=1)
        fs.append(np.ascontiguousarray(f))
    sizes = np.array([f.shape[2] for f in fs], dtype=np.int32)
    ptrs = (ctypes.c_void_p * max(1, len(fs)))(*[f.ctypes.data for f in fs])
    return conn, xy, u, sizes, ptrs, fs


def write_triangles(path, conn, xy, u, fields=(), binary=False):
    """TriangleWriter (binary: BinaryTriangleWriter) file: conn ntri x nper,
    xy nnodes x 2; u either ntri x 3 x 2 (values at the vertices) or nodal
    (nnodes x 2, or flat by DOF); fields a sequence of ntri x 3 x size arrays
    (values at the vertices) or ntri x size arrays (constant over the
    element)."""
    conn, xy, u, sizes, ptrs, keep = _triangle_args(conn, xy, u, fields)
    fn = lib().amie_write_triangles_bin if binary else lib().amie_write_triangles
    check(fn(str(path).encode(), conn.shape[0], conn.shape[1], _p(conn), xy.shape[0], _p(xy), _p(u), len(keep),
             _p(sizes), ptrs))


def default_opts(**kw):
    o = SolveOpts()
    lib().amie_gpu_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Batch:
    def __init__(self, ctx, problems=None, bcsr=None):
        self.ctx = ctx
        self.h = vp()
        if problems is not None:
            self.problems = list(problems)
            arr = (ProblemDesc * len(self.problems))(*[p.desc for p in self.problems])
            check(lib().amie_gpu_batch_create(ctx.h, len(self.problems), arr, ctypes.byref(self.h)))
            self.etype = self.problems[0].etype
            self.nrhs = self.problems[0].nrhs
            self.b = DIM[self.etype]
        else:
            stride, row_size, column_index, array, rhs = bcsr
            self.problems = None
            rs = _c(row_size, np.uint32)
            ci = _c(column_index, np.uint32)
            ar = _c(array, np.float64)
            b = _c(rhs, np.float64)
            check(lib().amie_gpu_batch_from_bcsr(ctx.h, stride, len(rs), _p(rs), _p(ci), _p(ar), _p(b),
                                                 ctypes.byref(self.h)))
            self.etype = P1 if stride == 2 else TET4
            self.nrhs = 1
            self.b = stride

    def update_bcsr(self, array, rhs):
        """New values and right-hand side, same pattern (drop-in batches)."""
        ar = _c(array, np.float64)
        b = _c(rhs, np.float64)
        check(lib().amie_gpu_batch_update_bcsr(self.h, _p(ar), _p(b)))

    def close(self):
        if self.h:
            lib().amie_gpu_batch_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbytes(self):
        return lib().amie_gpu_batch_bytes(self.h)

    def shape(self, p=0):
        nr, nz = c_i64(), c_i64()
        check(lib().amie_gpu_batch_get_nnzb(self.h, p, ctypes.byref(nr), ctypes.byref(nz)))
        return nr.value, nz.value

    def assemble(self):
        check(lib().amie_gpu_batch_assemble(self.h))

    def set_assembly(self, method):
        """ASSEMBLE_GATHER (default) or ASSEMBLE_COLOURED; re-assembles."""
        check(lib().amie_gpu_batch_set_assembly(self.h, method))

    def solve(self, **kw):
        o = default_opts(**kw)
        st = SolveStats()
        rc = lib().amie_gpu_batch_solve(self.h, ctypes.byref(o), ctypes.byref(st))
        # a NaN or a non-converged problem is reported in the stats (status)
        check(rc, allow=(0, 1, -4))
        return st.as_dict()

    def problem_info(self, p=0):
        st, it, ck, rr = c_i32(), c_i64(), c_i64(), c_dbl()
        check(lib().amie_gpu_batch_problem_info(self.h, p, ctypes.byref(st), ctypes.byref(it), ctypes.byref(ck),
                                                ctypes.byref(rr)))
        return {"status": st.value, "iterations": it.value, "refinements": ck.value, "relres": rr.value}

    def solution(self, p=0):
        nr, _ = self.shape(p)
        x = np.zeros(self.nrhs * nr * self.b)
        check(lib().amie_gpu_batch_get_solution(self.h, p, _p(x)))
        return x.reshape(self.nrhs, -1) if self.nrhs > 1 else x

    def set_solution(self, x, p=0):
        x = _c(x, np.float64)
        check(lib().amie_gpu_batch_set_solution(self.h, p, _p(x)))

    def matrix(self, p=0):
        nr, nz = self.shape(p)
        rs = np.zeros(nr, dtype=np.uint32)
        ci = np.zeros(nz, dtype=np.uint32)
        cl = self.b + self.b % 2
        ar = np.zeros(nz * self.b * cl)
        rhs = np.zeros(self.nrhs * nr * self.b)
        check(lib().amie_gpu_batch_get_matrix(self.h, p, _p(rs), _p(ci), _p(ar), _p(rhs)))
        return rs, ci, ar, (rhs.reshape(self.nrhs, -1) if self.nrhs > 1 else rhs)

    def element_blocks(self, p=0):
        pr = self.problems[p]
        n = NPER[self.etype]
        s = np.zeros(pr.conn.shape[0] * n * n, dtype=np.int64)
        check(lib().amie_gpu_batch_get_element_blocks(self.h, p, _p(s)))
        return s.reshape(-1, n, n)

    def spmv(self, x, p=0, precision=FP64):
        x = _c(x, np.float64)
        y = np.zeros_like(x)
        check(lib().amie_gpu_batch_spmv(self.h, p, precision, _p(x), _p(y)))
        return y

    def time_spmv(self, precision=MIXED_FIX48, precond=JACOBI, iters=20):
        """Mean SpMV-kernel launch time (us) and algorithmic bytes per launch
        with every problem of the batch active (the autotuned kernel when
        there is one)."""
        us, by = c_dbl(), c_dbl()
        check(lib().amie_gpu_batch_time_spmv(self.h, precision, precond, iters, ctypes.byref(us), ctypes.byref(by)))
        return us.value, by.value

    def autotune(self, precision=MIXED_FIX48, precond=JACOBI, iters=5):
        """Picks the fastest SpMV candidate for this batch: returns
        {"steps", "tile_order", "format", "us"}; later solves of the same
        precision and preconditioner use it."""
        st, order, fmt, us = c_i32(), c_i32(), c_i32(), c_dbl()
        check(lib().amie_gpu_batch_autotune(self.h, precision, precond, iters, ctypes.byref(st), ctypes.byref(order),
                                            ctypes.byref(fmt), ctypes.byref(us)))
        return {"steps": st.value, "tile_order": order.value, "format": fmt.value, "us": us.value}

    def set_spmv_kernel(self, precision, precond, steps, tile_order, format=0):
        """Selects an SpMV candidate directly (blocks per step, tile order,
        format 0 value planes / 1 value-indexed FIX48 matrix)."""
        check(lib().amie_gpu_batch_set_spmv_kernel(self.h, precision, precond, steps, tile_order, format))

    def element_strain(self, p=0, rhs=0):
        pr = self.problems[p]
        out = np.zeros((pr.conn.shape[0], VOIGT[self.etype]))
        check(lib().amie_gpu_batch_element_strain(self.h, p, rhs, _p(out)))
        return out

    def gauss_fields(self, p=0, rhs=0):
        """Strain and stress at every Gauss point: (nelem, ngp, 3|6) each."""
        pr = self.problems[p]
        ngp = 4 if self.etype == P2 else 1
        eps = np.zeros((pr.conn.shape[0], ngp, VOIGT[self.etype]))
        sig = np.zeros_like(eps)
        check(lib().amie_gpu_batch_gauss_fields(self.h, p, rhs, _p(eps), _p(sig)))
        return eps, sig

    def set_nonlocal(self, radius, kernel=0, threshold=1.0):
        """Non-local damage driving strain (linear elements); radius <= 0
        turns it off."""
        check(lib().amie_gpu_batch_set_nonlocal(self.h, radius, kernel, threshold))

    def smoothed_strain(self, p=0, rhs=0):
        pr = self.problems[p]
        out = np.zeros((pr.conn.shape[0], VOIGT[self.etype]))
        check(lib().amie_gpu_batch_smoothed_strain(self.h, p, rhs, _p(out)))
        return out

    def average_stress(self):
        out = np.zeros((len(self.problems), self.nrhs, VOIGT[self.etype]))
        check(lib().amie_gpu_batch_average_stress(self.h, _p(out)))
        return out

    def damage_step(self, kappa0=1e-4, kappaf=1e-3, dmax=1 - 1e-6):
        prm = DamageParams(kappa0, kappaf, dmax)
        g = c_i64()
        check(lib().amie_gpu_batch_damage_step(self.h, ctypes.byref(prm), ctypes.byref(g)))
        return g.value

    def event_step(self, criterion, mazars, radius, kernel=0, **kw):
        """One FeatureTree::step of the event-driven damage (NonLocalMazars
        + FiberBasedIsotropicLinearDamage, or IsotropicLinearDamage with
        damage_model=DAMAGE_ISOTROPIC_LINEAR) at the current Dirichlet
        values: criterion[m] (1/0) and mazars[m] = (threshold, E, nu, Gf) per
        material; kw: fibre_fraction, fractured_at, residual,
        score_tolerance, first_deferred, max_passes, damage_model,
        max_passes_per_step, and the solve options (precision, rtol,
        ref_eps, ...).  Returns the step's statistics."""
        prm = EventParams()
        lib().amie_gpu_default_event_params(ctypes.byref(prm))
        crit = np.ascontiguousarray(criterion, dtype=np.int32)
        maz = np.ascontiguousarray(np.asarray(mazars, dtype=np.float64).reshape(-1, 4))
        prm.nphase = len(crit)
        prm.criterion = crit.ctypes.data
        prm.mazars = maz.ctypes.data
        prm.radius = radius
        prm.kernel = kernel
        for k in ("fibre_fraction", "fractured_at", "residual", "score_tolerance", "first_deferred", "max_passes",
                  "damage_model", "max_passes_per_step"):
            if k in kw:
                setattr(prm, k, kw.pop(k))
        o = default_opts(**kw)
        st = EventStats()
        rc = lib().amie_gpu_batch_event_step(self.h, ctypes.byref(prm), ctypes.byref(o), ctypes.byref(st))
        check(rc, allow=(0, 1, -4))
        return st.as_dict()

    def scores(self, p=0):
        """Fracture scores of problem p's elements after the last pass."""
        out = np.zeros(self.problems[p].conn.shape[0])
        check(lib().amie_gpu_batch_get_scores(self.h, p, _p(out)))
        return out

    def set_damage(self, d, p=0):
        """Replace problem p's damage (one value per element); the next
        assembly uses it."""
        d = _c(d, np.float64)
        if d.shape != (self.problems[p].conn.shape[0],):
            raise ValueError("one damage value per element")
        check(lib().amie_gpu_batch_set_damage(self.h, p, _p(d)))

    def damage(self, p=0):
        ne = self.problems[p].conn.shape[0]
        d, k = np.zeros(ne), np.zeros(ne)
        check(lib().amie_gpu_batch_get_damage(self.h, p, _p(d), _p(k)))
        return d, k

    def scale_bc(self, f):
        check(lib().amie_gpu_batch_scale_bc(self.h, f))

    def set_bc_values(self, dofs, vals, p=0):
        """New values of existing Dirichlet conditions (next assembly)."""
        dofs = _c(dofs, np.int64)
        vals = _c(vals, np.float64)
        check(lib().amie_gpu_batch_set_bc_values(self.h, p, len(dofs), _p(dofs), _p(vals)))


class Server:
    def __init__(self, ctx, hbm_budget, max_slots, **opts):
        self.ctx = ctx
        self.h = vp()
        self.opts = default_opts(**opts)
        check(lib().amie_gpu_server_create(ctx.h, hbm_budget, max_slots, ctypes.byref(self.opts),
                                           ctypes.byref(self.h)))
        self._keep = {}

    def submit(self, problem, t_ms, key=0, damage=None):
        """Queues a request; key != 0 makes it a load step on that key's
        resident problem, damage = (kappa0, kappaf, dmax) its damage law."""
        t = c_u64()
        if damage is None:
            check(lib().amie_gpu_server_submit(self.h, ctypes.byref(problem.desc), key, t_ms, ctypes.byref(t)))
        else:
            prm = DamageParams(*damage)
            check(lib().amie_gpu_server_submit_step(self.h, ctypes.byref(problem.desc), key, ctypes.byref(prm), t_ms,
                                                    ctypes.byref(t)))
        return t.value

    @staticmethod
    def request_bytes(problem):
        """HBM set aside for a request at admission (estimate)."""
        b = c_i64()
        check(lib().amie_gpu_server_request_bytes(ctypes.byref(problem.desc), ctypes.byref(b)))
        return b.value

    def release(self, key):
        check(lib().amie_gpu_server_release(self.h, key))

    def damage(self, key, nelem):
        d, k = np.zeros(nelem), np.zeros(nelem)
        check(lib().amie_gpu_server_get_damage(self.h, key, _p(d), _p(k), nelem))
        return d, k

    def step(self, now_ms):
        na, nq = c_i32(), c_i32()
        check(lib().amie_gpu_server_step(self.h, now_ms, ctypes.byref(na), ctypes.byref(nq)))
        return na.value, nq.value

    def poll(self, maxn=4096):
        out = (RequestResult * maxn)()
        n = c_i32()
        check(lib().amie_gpu_server_poll(self.h, out, maxn, ctypes.byref(n)))
        return [{k: getattr(out[i], k) for k, _ in RequestResult._fields_} for i in range(n.value)]

    def solution(self, ticket, n):
        x = np.zeros(n)
        check(lib().amie_gpu_server_get_solution(self.h, ticket, _p(x), n))
        return x

    def close(self):
        if self.h:
            lib().amie_gpu_server_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
