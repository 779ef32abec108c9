"""The allocation-free voxel sampling (amie_gpu_voxelize_noalloc) against the
allocating entry point (amie_gpu_voxelize), both through the C ABI: the
mesh, the field and the output live in device arrays this test allocates
with the HIP runtime, and the library's call allocates nothing.  Two calls
in a row, the second on a smaller mesh and grid, each bitwise equal to
amie_gpu_voxelize; the largest grid the scratch array holds works, a larger
one is refused before any launch.
"""
import ctypes

import numpy as np
import pytest

import amie_gpu as ag

pytestmark = pytest.mark.gpu
SCRATCH_MAX = 1 << 22          # AMIE_VOXEL_SCRATCH_MAX
H2D, D2H = 1, 2                # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """Device arrays of the caller, through the HIP runtime the library
    itself is linked against."""

    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        self.L = L
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.live.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def ctx():
    c = ag.Context(0)
    yield c
    c.close()


def noalloc(ctx, hip, xyz, conn, field, box, n):
    """amie_gpu_voxelize_noalloc on device copies; returns ncomp x n^3."""
    field = np.ascontiguousarray(field, dtype=np.float64)
    ncomp = field.shape[1]
    dx, dc, df = hip.put(xyz.astype(np.float64)), hip.put(conn.astype(np.int32)), hip.put(field)
    dv = hip.empty(ncomp * n ** 3 * 8)
    ag.check(ag.lib().amie_gpu_voxelize_noalloc(ctx.h, None, xyz.shape[0], dx, conn.shape[0], dc, ncomp, df,
                                                box.ctypes.data, n, n, n, dv))
    hip.sync()   # the call queues its kernels on the context's stream and returns
    return hip.get(dv, (ncomp, n ** 3), np.float64)


def case(n_cells, ncomp, seed):
    xyz, conn = ag.structured_mesh(ag.TET4, n_cells)
    field = np.random.default_rng(seed).normal(size=(conn.shape[0], ncomp))
    return xyz, conn, field


def test_noalloc_voxelize_matches_allocating_entry_point(ctx):
    hip = Hip()
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    try:
        # first call, then a smaller mesh and grid: the array keeps the first
        # call's owners, which the second call must not see
        for n_cells, ncomp, n, seed in ((12, 3, 40, 1), (5, 2, 16, 2)):
            print("case", n_cells, ncomp, n, flush=True)
            xyz, conn, field = case(n_cells, ncomp, seed)
            ref = ag.voxelize_gpu(ctx, xyz, conn, field, box, n)
            got = noalloc(ctx, hip, xyz, conn, field, box, n)
            np.testing.assert_array_equal(got, ref)
        # the largest cubic grid the scratch array holds
        n = int(round(SCRATCH_MAX ** (1 / 3)))
        while (n + 1) ** 3 <= SCRATCH_MAX:
            n += 1
        while n ** 3 > SCRATCH_MAX:
            n -= 1
        print("case max", n, flush=True)
        xyz, conn, field = case(6, 1, 3)
        ref = ag.voxelize_gpu(ctx, xyz, conn, field, box, n)
        got = noalloc(ctx, hip, xyz, conn, field, box, n)
        np.testing.assert_array_equal(got, ref)
    finally:
        hip.free()


def test_noalloc_voxelize_refuses_grids_beyond_its_array(ctx):
    n = 1
    while n ** 3 <= SCRATCH_MAX:
        n += 1
    xyz, conn, field = case(2, 1, 4)
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    # refused on the host before any launch: host pointers are never read
    # as device memory because the size check comes first
    hip = Hip()
    try:
        dx, dc, df = hip.put(xyz), hip.put(conn.astype(np.int32)), hip.put(field)
        dv = hip.empty(8)
        with pytest.raises(ag.AmieError, match="at most"):
            ag.check(ag.lib().amie_gpu_voxelize_noalloc(ctx.h, None, xyz.shape[0], dx, conn.shape[0], dc, 1, df,
                                                        box.ctypes.data, n, n, n, dv))
    finally:
        hip.free()
