"""Benchmark of the FE-solve hot path on one MI355X.

Workload (one "step"): 512 periodic P1 RVEs (n ~ U{32..256}, seed 4, 2k-130k
DOFs each, 20-45 % circular inclusions, 12/60 GPa), three macro-strain load
cases each (1536 right-hand sides), resident in HBM.  A step runs the whole
hot path on the device: element integration + gather assembly + Dirichlet
pin + scaling (assembly), the mixed-precision CG (FP32 vectors, the scaled
matrix in 48-bit fixed point with FP64 accumulation, FP64 refinement) of all
1536 right-hand sides to relative residual 1e-8 in one batch, and the
volume-averaged stress (effective stiffness) of every RVE.

value = RVE load-case solves per second (1536 per step / step time).
roofline = the kernel with the larger share of the timed steps, the CG
update (k_update) or the SpMV (k_spmv), both timed with HIP events on the
library stream during the timed steps and both reported
(roofline_update, roofline_spmv).  k_update: the algorithmic bytes of every
launch (27 B per DOF and right-hand side on average: y += alpha p is added
every fourth iteration, four terms at once) over the kernel time.  k_spmv: the
algorithmic bytes of the FP32 block SpMV (nnzb (b^2 4 + 4) + 4 (nb + 1) +
2 n 4 per right-hand side) of every launch over the kernel time, the
physical bytes of the stored format beside it.
Beside it: the all-FP32 inner iteration, the plain FP64 CG (same kernels),
the displacement error against the CPU FP64 oracle on a sample of RVEs at
several stopping residuals, and the same step as one persistent launch per
solve pass.
cpu_baseline = the reference's own threaded ConjugateGradient (compiled from
its sources into oracle/_ref/ref_harness) on a bounded sample of the same
RVEs, solved to rel. res <= 1e-8, with its per-DOF rate, a 1-thread run and
the host's lscpu; the oracle's threaded FP64 Jacobi PCG beside it.
Side legs: the P1 128^2 patch test (FP64 CG + Jacobi to 1e-10, the reference
CG beside it), P2 400^2 and the Tet4 64^3 damage case (the reference CG
timed over a bounded number of iterations beside each), the served request
mix, mesh input (mid-edge nodes + renumbering) and voxel output on the
device.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xfem-amie_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_spmv_rve512.md")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def measured_traffic(kernel="k_spmv"):
    """HBM bytes / algorithmic bytes of a kernel, from the FETCH_SIZE and
    WRITE_SIZE counter runs summarised in the committed profile (scaled by
    the same counters' reading for a plain copy of known size)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("<!-- traffic:"):
                    return json.loads(line[len("<!-- traffic:"):].rsplit("-->", 1)[0])[kernel]
    except (OSError, ValueError, KeyError):
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes_8d(nrows, nnzb, b, nrhs):
    """Algorithmic bytes of one FP32 block SpMV over nrhs right-hand sides:
    values + block column indices, row pointer, x read and y written once
    per right-hand side (padding not counted)."""
    return nnzb * (b * b * 4 + 4) + 4 * (nrows + 1) + 2 * nrows * b * 4 * nrhs


def lscpu_summary():
    """The host's CPU as lscpu reports it, and the CPUs this process may use."""
    out = {}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in text.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                             "NUMA node(s)"):
                out[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    out["usable_cpus"] = len(os.sched_getaffinity(0))
    return out


def _write_system(path, A, f, eps, maxit):
    """A system in the layout the reference harness's --cg mode reads."""
    with open(path, "wb") as fh:
        fh.write(np.int32(A.stride).tobytes())
        fh.write(np.int64(A.nrows).tobytes())
        fh.write(np.ascontiguousarray(A.row_size, dtype=np.uint32).tobytes())
        fh.write(np.int64(len(A.column_index)).tobytes())
        fh.write(np.ascontiguousarray(A.column_index, dtype=np.uint32).tobytes())
        fh.write(np.ascontiguousarray(A.array, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(f, dtype=np.float64).tobytes())
        fh.write(np.float64(eps).tobytes())
        fh.write(np.int64(maxit).tobytes())


def reference_cg(tmp, A, f, eps, maxit, threads=None):
    """The reference's ConjugateGradient::solve (InverseDiagonal, nssor 32)
    on A x = f through oracle/_ref/ref_harness --cg: iterations, seconds of
    the solve, converged flag, OpenMP threads, solution."""
    path = os.path.join(tmp, "system.bin")
    _write_system(path, A, f, eps, maxit)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run([REF_HARNESS, "--cg", path], capture_output=True, text=True, timeout=900, env=env)
    if out.returncode != 0:
        raise RuntimeError(f"reference CG failed: {out.stderr[-300:]}")
    it, secs, conv, thr = out.stdout.strip().split("\n")[-1].split()
    return int(it), float(secs), int(conv), int(thr), np.fromfile(path + ".x")


def reference_solve_to(tmp, A, f, rtol, threads=None):
    """The reference CG solved until ||f - A x|| / ||f|| <= rtol: its own
    absolute stop sqrt(r.D^-1 r) <= eps starts at rtol sqrt(f.D^-1 f) / 4 and
    is tightened until the relative residual (the oracle's FP64 SpMV) meets
    rtol.  Returns iterations, seconds (the solve that met it), threads,
    relative residual."""
    import oracle as orc
    dinv = orc.inverse_diagonal(A)
    eps = 0.25 * rtol * np.sqrt(f @ (dinv * f))
    for _ in range(4):
        it, secs, conv, thr, x = reference_cg(tmp, A, f, eps, 10 * A.n, threads)
        rr = float(np.linalg.norm(f - orc.spmv(A, x)) / np.linalg.norm(f))
        if rr <= rtol:
            return it, secs, thr, rr
        eps *= 0.1
    raise RuntimeError(f"reference CG stopped at rel. res {rr:.2e} > {rtol:g}")


def rve_system(p):
    """The oracle's periodic RVE system (fluctuation pinned at DOF node 0)
    and its three macro-strain loads."""
    import oracle as orc
    A = orc.assemble_periodic(1, p.conn, p.dofmap, p.coords, p.cmats, p.phase)
    loads = [orc.macro_rhs(1, p.conn, p.dofmap, p.coords, p.cmats, p.phase, p.macro[lc], A.nrows)
             for lc in range(p.nrhs)]
    orc.apply_dirichlet(A, np.zeros(A.n), p.bc_dofs, p.bc_vals)
    for f in loads:
        f[p.bc_dofs] = p.bc_vals
    return A, loads


def cpu_baseline(problems, budget_s=20.0, rtol=1e-8):
    """The reference's own ConjugateGradient::solve (compiled from its
    sources into oracle/_ref/ref_harness, --cg mode: InverseDiagonal, nssor
    32), each solve to rel. res <= rtol (its absolute stop tightened until it
    gets there), on a sample drawn across the whole range of n: the RVEs at
    the 10 %, 30 %, 50 %, 70 % and 90 % points of n, then more in a seeded
    order, until the time budget is spent.  The first solve runs with one
    OpenMP thread and with every thread of the host; the faster setting
    (the reference's OpenMP tasks cost more than they gain on small rows)
    times the sample.  Beside it the rate per DOF, the host's lscpu, and the
    oracle's OpenMP PCG on the same solves.  Without the compiled reference,
    the oracle alone (kind "port")."""
    import oracle as orc
    use_ref = os.path.exists(REF_HARNESS)
    by_n = np.argsort([p.ndofnodes for p in problems], kind="stable")
    first = [int(by_n[int(q * (len(problems) - 1))]) for q in (.1, .3, .5, .7, .9)]
    rest = [int(k) for k in np.random.default_rng(11).permutation(len(problems)) if int(k) not in first]
    solved, t_ref, t_port, used, worst, iters, dofs = 0, 0.0, 0.0, [], 0.0, 0, 0
    threads, probe = None, None
    with tempfile.TemporaryDirectory() as tmp:
        for k in first + rest:
            p = problems[k]
            n = int(np.sqrt(p.ndofnodes))
            A, loads = rve_system(p)
            for f in loads:
                t0 = time.perf_counter()
                orc.pcg(A, f, rtol=rtol, maxit=200000)
                t_port += time.perf_counter() - t0
                if use_ref:
                    if threads is None:
                        it1, s1, thr1, _ = reference_solve_to(tmp, A, f, rtol, threads=1)
                        itn, sn, thrn, _ = reference_solve_to(tmp, A, f, rtol)
                        threads = 1 if s1 <= sn else thrn
                        probe = {"rve_n": n, "dofs": int(A.n), "one_thread_s": s1, "all_threads_s": sn,
                                 "all_threads": thrn, "iterations": it1}
                    it, secs, thr, rr = reference_solve_to(tmp, A, f, rtol, threads=threads)
                    worst = max(worst, rr)
                    t_ref += secs
                    iters += it
                dofs += A.n
            solved += len(loads)
            used.append(n)
            spent = t_ref if use_ref else t_port
            log(f"cpu baseline: rve n={n} {len(loads)} solves, reference {t_ref:.1f}s, oracle {t_port:.1f}s")
            if spent >= budget_s:
                break
    port = {"value": solved / t_port, "unit": "solves/s", "cores": orc.get_threads(), "kind": "port",
            "dof_solves_per_s": dofs / t_port,
            "sample": "the same solves with the oracle's FP64 Jacobi PCG (OpenMP)"}
    if not use_ref:
        port["sample"] = (f"{len(used)} of the {len(problems)} RVEs (n = {sorted(used)}), {solved} load-case solves, "
                          f"oracle FP64 Jacobi PCG to rel. res {rtol:g}, assembly excluded")
        port["lscpu"] = lscpu_summary()
        return port
    return {"value": solved / t_ref, "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": f"{len(used)} of the {len(problems)} RVEs (n = {sorted(used)}, drawn across the range of n), "
                      f"{solved} load-case solves by the reference's ConjugateGradient::solve (InverseDiagonal, "
                      f"nssor 32, {iters} iterations in all), each solved to rel. res <= {rtol:g} (worst reached "
                      f"{worst:.2e}), {threads} OpenMP thread(s), assembly and file transfer excluded",
            "dof_solves_per_s": dofs / t_ref, "threads_probe": probe, "lscpu": lscpu_summary(), "port": port}


def oracle_reference_solutions(problems, sample, rtol=1e-12):
    """FP64 solutions of the sampled RVEs by the CPU oracle (OpenMP PCG to
    rel. res rtol on the oracle's own assembly): {k: [x per load case]}."""
    import oracle as orc
    out, t0 = {}, time.perf_counter()
    for k in sample:
        A, loads = rve_system(problems[k])
        out[k] = [orc.pcg(A, f, rtol=rtol, maxit=500000)[0] for f in loads]
    log(f"oracle FP64 solutions of {len(sample)} RVEs (rel. res {rtol:g}): {time.perf_counter() - t0:.1f} s")
    return out


def rel_err(x, xref):
    x = np.asarray(x).reshape(len(xref), -1)
    return max(float(np.linalg.norm(x[lc] - xref[lc]) / np.linalg.norm(xref[lc])) for lc in range(len(xref)))


def ingest_input(ag, n):
    """Structured P1 mesh of 2 n^2 triangles, nodes relabelled and triangles
    shuffled (seed 13), as a mesher's element list would arrive."""
    coords, conn = ag.structured_mesh(ag.P1, n)
    rng = np.random.default_rng(13)
    perm = rng.permutation(coords.shape[0])
    xy = np.empty_like(coords)
    xy[perm] = coords
    return xy, perm[conn][rng.permutation(conn.shape[0])].astype(np.int32)


def ingest_leg(ag, ctx, n):
    """Mesh input on the device: P2 mid-edge nodes (DelaunayTree::
    addSharedNodes) then FeatureTree's first-touch renumbering of the P2 mesh,
    device time with the inputs in HBM (best of 3)."""
    xy, conn3 = ingest_input(ag, n)
    best_mid, best_ren = float("inf"), float("inf")
    for _ in range(3):
        c6, mid, ms_mid = ag.add_midpoints_gpu(ctx, xy, conn3, timed=True)
        ids, count, ms_ren = ag.renumber_gpu(ctx, c6, xy.shape[0] + mid.shape[0], timed=True)
        best_mid, best_ren = min(best_mid, ms_mid), min(best_ren, ms_ren)
    ntri = conn3.shape[0]
    return {"workload": f"structured P1 {n}x{n} ({ntri} triangles, {xy.shape[0]} vertices), nodes relabelled and "
                        f"triangles shuffled (seed 13): P2 mid-edge nodes, then first-touch renumbering of the "
                        f"{count} P2 nodes",
            "triangles": ntri, "p2_nodes": int(count), "add_midpoints_ms": best_mid, "renumber_ms": best_ren,
            "triangles_per_s": ntri / ((best_mid + best_ren) * 1e-3)}


def voxel_leg(ag, ctx, n, nvox):
    """VoxelWriter sampling on the device: a structured Tet4 mesh of n^3
    cells (6 n^3 tetrahedra) with a 6-component element field (seed 17) on
    an nvox^3 grid; device time with the inputs in HBM (best of 3)."""
    xyz, conn = ag.structured_mesh(ag.TET4, n)
    field = np.random.default_rng(17).normal(size=(conn.shape[0], 6))
    box = np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    best = float("inf")
    for _ in range(3):
        _, ms = ag.voxelize_gpu(ctx, xyz, conn, field, box, nvox, timed=True)
        best = min(best, ms)
    return {"workload": f"structured Tet4 {n}^3 ({conn.shape[0]} tetrahedra), 6-component field, {nvox}^3 voxels",
            "voxelize_ms": best, "voxels_per_s": nvox ** 3 / (best * 1e-3)}


def cpu_voxels(ag, n, nvox):
    """The oracle's serial voxel ownership on the same input."""
    import oracle as orc
    xyz, conn = ag.structured_mesh(ag.TET4, n)
    t0 = time.perf_counter()
    orc.voxel_owner(xyz, conn, np.array([0.5, 0.5, 0.5, 1.0, 1.0, 1.0]), nvox)
    dt = time.perf_counter() - t0
    return {"voxelize_ms": 1e3 * dt, "voxels_per_s": nvox ** 3 / dt, "cores": 1, "kind": "port"}


def cpu_ingest(ag, n):
    """The oracle's serial add_midpoints + renumber on the same input."""
    import oracle as orc
    xy, conn3 = ingest_input(ag, n)
    t0 = time.perf_counter()
    c6, mid = orc.add_midpoints(xy, conn3)
    t1 = time.perf_counter()
    orc.renumber(c6, xy.shape[0] + mid.shape[0])
    t2 = time.perf_counter()
    return {"add_midpoints_ms": 1e3 * (t1 - t0), "renumber_ms": 1e3 * (t2 - t1),
            "triangles_per_s": conn3.shape[0] / (t2 - t0), "cores": 1, "kind": "port"}


def serving(ag, ctx, reqs, precision, rtol, limit_s=90.0, **opts):
    """Replays a request trace against the server in wall-clock time and
    returns p50/p99 request latency (arrival to result) and solves/s.  The
    server autotunes each group's SpMV on its build thread."""
    S = ag.Server(ctx, 200 << 30, 64, precision=precision, rtol=rtol, autotune=1, **opts)
    t0 = time.perf_counter()
    lat, solves, nsub, bad = [], 0, 0, 0
    nrhs = {}
    while len(lat) < len(reqs):
        now = time.perf_counter() - t0
        if now > limit_s:
            break
        while nsub < len(reqs) and reqs[nsub][0] <= now:
            t_arr, kind, prob, key = reqs[nsub]
            tk = S.submit(prob, 1e3 * t_arr, key)
            nrhs[tk] = prob.nrhs
            nsub += 1
        S.step(1e3 * (time.perf_counter() - t0))
        for r in S.poll():
            lat.append(r["done_ms"] - r["submit_ms"])
            solves += nrhs[r["ticket"]]
            bad += r["status"] != 0
    elapsed = time.perf_counter() - t0
    S.close()
    lat = np.array(lat)
    return {"requests": len(reqs), "completed": int(len(lat)), "not_converged": int(bad),
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "solves_per_s": solves / elapsed, "wall_s": elapsed}


def certified_reference(ag, orc, B, A, b, rtol=1e-12):
    """The GPU's plain FP64 CG solved to rel. res rtol, its residual checked
    by the CPU oracle's FP64 SpMV on the oracle's own assembly: the FP64
    reference solution of a problem too large for a CPU solve inside the
    bench, and the residual the CPU certifies for it."""
    st = B.solve(precision=ag.FP64, precond=ag.JACOBI, rtol=rtol, use_graph=0, time_kernels=0)
    x = B.solution(0).copy()
    rr = float(np.linalg.norm(orc.spmv(A, x, b)) / np.linalg.norm(b))
    return x, rr, st["iterations_max"]


def reference_bounded(tmp, A, b, iters, threads=None):
    """The work of `iters` iterations of the reference CG's inner loop
    (ref_harness --cg-rate: the reference's own InverseDiagonal, dots, block
    SpMV and Kahan updates): ms per iteration."""
    path = os.path.join(tmp, "system.bin")
    _write_system(path, A, b, 0., iters)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run([REF_HARNESS, "--cg-rate", path, str(iters)], capture_output=True, text=True, timeout=900,
                         env=env)
    if out.returncode != 0:
        raise RuntimeError(f"reference CG rate failed: {out.stderr[-300:]}")
    it, secs, thr = out.stdout.strip().split("\n")[-1].split()
    return {"iterations": int(it), "ms_per_iteration": 1e3 * float(secs) / max(1, int(it)), "cores": int(thr),
            "kind": "reference"}


def config1_leg(ag, configs, ctx, no_cpu):
    """P1 128^2 uniaxial patch test (33 282 DOFs): plain FP64 CG + Jacobi to
    rel. res 1e-10 on the device (and the FIX48 mixed precision beside it),
    the error against the exact linear solution u_y = 1e-3 y, u_x = -nu 1e-3 x
    (plane stress), and the reference's ConjugateGradient on the same system
    (every thread and one, the faster kept), stopped at rel. res <= 1e-10."""
    import oracle as orc
    p = configs.config1(128)
    B = ag.Batch(ctx, [p])
    leg = {"workload": "P1 128^2 unit square, E 10 GPa, nu 0.2 plane stress, bottom u_y = 0, corner u_x = 0, top "
                       "u_y = 1e-3; Jacobi, rel. res 1e-10", "dofs": int(p.ndof)}
    exact = np.stack([-0.2 * 1e-3 * p.coords[:, 0], 1e-3 * p.coords[:, 1]], axis=1).ravel()
    for name, prec in (("fp64", ag.FP64), ("mixed", ag.MIXED_FIX48)):
        B.solve(precision=prec, rtol=1e-10, use_graph=0)   # warm
        t = time.perf_counter()
        st = B.solve(precision=prec, rtol=1e-10, use_graph=0)
        dt = time.perf_counter() - t
        x = B.solution(0)
        leg[name] = {"solve_ms": 1e3 * dt, "iterations": st["iterations_max"], "relres": st["relres_max"],
                     "status": st["status"],
                     "patch_test_max_err": float(np.abs(x - exact).max() / np.abs(exact).max())}
    B.close()
    if not no_cpu and os.path.exists(REF_HARNESS):
        A = orc.assemble(p.etype, p.conn, p.coords, p.cmats)
        b = np.zeros(A.n)
        orc.apply_dirichlet(A, b, p.bc_dofs, p.bc_vals)
        with tempfile.TemporaryDirectory() as tmp:
            it, secs, thr, rr = reference_solve_to(tmp, A, b, 1e-10)
            it1, secs1, thr1, rr1 = reference_solve_to(tmp, A, b, 1e-10, threads=1)
        best = min((secs, thr, it, rr), (secs1, thr1, it1, rr1))
        leg["cpu_reference"] = {"solve_ms": 1e3 * best[0], "iterations": best[2], "relres": best[3], "cores": best[1],
                                "kind": "reference", "one_thread_ms": 1e3 * secs1, "all_threads_ms": 1e3 * secs,
                                "all_threads": thr}
        leg["fp64_vs_cpu_reference"] = best[0] * 1e3 / leg["fp64"]["solve_ms"]
    log(f"P1 128^2: {leg}")
    return leg


def single_problem_legs(ag, configs, ctx, rtol, maxit=0, no_cpu=False, ref_iters=20):
    """The single-problem configurations beside the batch: P2 400^2 composite
    (1.28 M DOFs) solved once, and the Tet4 64^3 damage case (824 k DOFs)
    through 10 load steps (solve + damage update each) with the 3x3
    block-Jacobi preconditioner its configuration names, point Jacobi
    beside it; FIX48 mixed precision against the plain FP64 CG.  The solves
    run without per-launch events (a problem this size keeps every tile
    active, so the SpMV is timed on its own with amie_gpu_batch_time_spmv).
    Accuracy: the displacement error of the rel. res 1e-8 solutions against
    the FP64 solution to 1e-12, whose residual the CPU oracle certifies on
    its own assembly (a CPU FP64 solve of these sizes takes minutes).  The
    reference CG is timed over ref_iters iterations on the same system (one
    thread and every thread, the faster kept): ms per iteration."""
    import oracle as orc
    out = {}

    def timed_solve(B, precision, precond, **kw):
        return B.solve(precision=precision, precond=precond, rtol=rtol, use_graph=0, time_kernels=0, maxit=maxit, **kw)

    p2 = configs.config2(400)
    B = ag.Batch(ctx, [p2])
    leg = {"workload": f"P2 400^2 composite, 35 % inclusions (seed 2), uniaxial, Jacobi, rel.res {rtol:g}",
           "dofs": int(p2.ndof)}
    x = {}
    for name, prec in (("mixed", ag.MIXED_FIX48), ("fp64", ag.FP64)):
        B.assemble()
        tuned = B.autotune(prec, ag.JACOBI)
        t = time.perf_counter()
        st = timed_solve(B, prec, ag.JACOBI)
        dt = time.perf_counter() - t
        x[name] = B.solution(0).copy()
        us, by = B.time_spmv(prec, ag.JACOBI, 50)
        nr, nz = B.shape(0)
        by8 = spmv_bytes_8d(nr, nz, 2, 1)
        leg[name] = {"solve_ms": 1e3 * dt, "iterations": st["iterations_max"], "status": st["status"],
                     "spmv_kernel": tuned, "spmv_us": us, "spmv_gbs": by / (us * 1e-6) / 1e9,
                     "spmv_bytes_per_launch": by, "spmv_gbs_fp32_yardstick": by8 / (us * 1e-6) / 1e9}
    leg["mixed_vs_fp64"] = leg["fp64"]["solve_ms"] / leg["mixed"]["solve_ms"]
    A = orc.assemble(p2.etype, p2.conn, p2.coords, p2.cmats, p2.phase)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p2.bc_dofs, p2.bc_vals)
    xt, rr, its = certified_reference(ag, orc, B, A, b)
    leg["accuracy"] = {"reference": "FP64 CG to rel. res 1e-12, residual certified by the CPU oracle",
                       "reference_relres_cpu": rr, "reference_iterations": its,
                       "mixed_relres_cpu": float(np.linalg.norm(orc.spmv(A, x["mixed"], b)) / np.linalg.norm(b)),
                       "mixed_displacement_rel_err": float(np.linalg.norm(x["mixed"] - xt) / np.linalg.norm(xt)),
                       "fp64_displacement_rel_err": float(np.linalg.norm(x["fp64"] - xt) / np.linalg.norm(xt))}
    if not no_cpu and os.path.exists(REF_HARNESS):
        with tempfile.TemporaryDirectory() as tmp:
            r = min(reference_bounded(tmp, A, b, ref_iters, threads=1), reference_bounded(tmp, A, b, ref_iters),
                    key=lambda v: v["ms_per_iteration"])
        r["extrapolated_solve_s"] = r["ms_per_iteration"] * 1e-3 * leg["fp64"]["iterations"]
        r["note"] = ("the reference CG's iteration work timed over a bounded number of iterations (ref_harness "
                     "--cg-rate); the solve time is that rate times the device FP64 CG's iteration count (an "
                     "estimate: the reference's own restarts and Richardson smoothing are not counted)")
        leg["cpu_reference"] = r
    out["p2_400"] = leg
    log(f"P2 400^2: {leg}")
    B.close()

    p3 = configs.config3(64)
    leg = {"workload": f"Tet4 64^3 damage case (d0 field seed 3), 10 load steps of 2e-5, block-Jacobi (3x3) as the "
                       f"configuration names it, and point Jacobi beside it, rel.res {rtol:g}",
           "dofs": int(p3.ndof)}

    def damage_run(prec, precond):
        B = ag.Batch(ctx, [p3])
        tuned = B.autotune(prec, precond)
        us_alone, by_alone = B.time_spmv(prec, precond, 50)
        nr, nz = B.shape(0)
        by8 = spmv_bytes_8d(nr, nz, 3, 1)
        t = time.perf_counter()
        its = 0
        for s in range(1, 11):
            if s > 1:
                B.scale_bc(s / (s - 1))
            B.assemble()
            st = timed_solve(B, prec, precond, use_x0=1 if s > 1 else 0)
            its += st["iterations_max"]
            B.damage_step()
        dt = time.perf_counter() - t
        d, _ = B.damage()
        B.close()
        return {"ten_steps_ms": 1e3 * dt, "iterations": its, "damage_mean": float(d.mean()), "spmv_kernel": tuned,
                "spmv_us": us_alone, "spmv_gbs": by_alone / us_alone / 1e3, "spmv_bytes_per_launch": by_alone,
                "spmv_gbs_fp32_yardstick": by8 / us_alone / 1e3}

    for name, prec in (("mixed", ag.MIXED_FIX48), ("fp64", ag.FP64)):
        leg[name] = damage_run(prec, ag.BLOCK_JACOBI)
    leg["mixed_vs_fp64"] = leg["fp64"]["ten_steps_ms"] / leg["mixed"]["ten_steps_ms"]
    leg["jacobi"] = {name: damage_run(prec, ag.JACOBI) for name, prec in (("mixed", ag.MIXED_FIX48), ("fp64", ag.FP64))}
    leg["jacobi"]["mixed_vs_fp64"] = leg["jacobi"]["fp64"]["ten_steps_ms"] / leg["jacobi"]["mixed"]["ten_steps_ms"]
    # accuracy of the first load step against the CPU-certified FP64 solution
    B = ag.Batch(ctx, [p3])
    xs = {}
    for name, prec in (("mixed", ag.MIXED_FIX48), ("fp64", ag.FP64)):
        timed_solve(B, prec, ag.BLOCK_JACOBI)
        xs[name] = B.solution(0).copy()
    A = orc.assemble(p3.etype, p3.conn, p3.coords, p3.cmats, None, p3.damage)
    b = np.zeros(A.n)
    orc.apply_dirichlet(A, b, p3.bc_dofs, p3.bc_vals)
    xt, rr, its = certified_reference(ag, orc, B, A, b)
    leg["accuracy_step1"] = {"reference": "FP64 CG to rel. res 1e-12, residual certified by the CPU oracle",
                             "reference_relres_cpu": rr, "reference_iterations": its,
                             "mixed_displacement_rel_err": float(np.linalg.norm(xs["mixed"] - xt) / np.linalg.norm(xt)),
                             "fp64_displacement_rel_err": float(np.linalg.norm(xs["fp64"] - xt) / np.linalg.norm(xt))}
    B.close()
    if not no_cpu and os.path.exists(REF_HARNESS):
        with tempfile.TemporaryDirectory() as tmp:
            r = min(reference_bounded(tmp, A, b, ref_iters, threads=1), reference_bounded(tmp, A, b, ref_iters),
                    key=lambda v: v["ms_per_iteration"])
        r["extrapolated_solve_s"] = r["ms_per_iteration"] * 1e-3 * leg["accuracy_step1"]["reference_iterations"]
        r["note"] = ("the reference CG's iteration work on the first load step, timed over a bounded number of "
                     "iterations (ref_harness --cg-rate); the solve time is that rate times the device FP64 CG's "
                     "iteration count to rel. res 1e-12 (an estimate)")
        leg["cpu_reference_step1"] = r
    out["tet_64_damage"] = leg
    log(f"Tet4 64^3: {leg}")
    return out


def event_leg(ag, configs, ctx, n, n_cpu, no_cpu, loads=(0.8e-4, 1.2e-4)):
    """Event-driven damage (FeatureTree's solve / score / damage-pass loop
    with NonLocalMazars and fibre-based damage) on the event square: the
    device at n x n, and at n_cpu x n_cpu beside the oracle's loop on the
    same case (the oracle's damage must come out equal).  The same square
    with IsotropicLinearDamage (its Newton branch) at n x n, and its first
    load step at n_cpu x n_cpu beside the oracle's (equal to 1e-8; later
    steps amplify solver round-off and are not compared)."""
    import oracle as orc

    def device(m, model=ag.DAMAGE_FIBRE, steps=loads):
        prob, dofs, nfix, crit, maz, radius = configs.event_square(m)
        B = ag.Batch(ctx, [prob])
        rows, t0 = [], time.perf_counter()
        for v in steps:
            vals = np.zeros(len(dofs))
            vals[nfix:] = v
            B.set_bc_values(dofs, vals)
            st = B.event_step(crit, maz, radius, precision=ag.MIXED_FIX48, rtol=1e-12, damage_model=model)
            rows.append(st)
        wall = time.perf_counter() - t0
        d = B.damage()[0]
        B.close()
        passes = sum(r["passes"] for r in rows)
        return {"elements": int(prob.conn.shape[0]), "radius": radius, "passes": passes,
                "events": sum(r["events"] for r in rows), "fractured": int(rows[-1]["fractured"]),
                "cg_iterations": sum(r["cg_iterations"] for r in rows), "wall_s": wall,
                "ms_per_pass": 1e3 * wall / max(1, passes),
                "solve_ms": sum(r["solve_ms"] for r in rows), "damage_ms": sum(r["damage_ms"] for r in rows),
                "all_converged": all(r["converged"] == 1 for r in rows)}, d

    big, _ = device(n)
    log(f"event damage {n}^2: {big}")
    out = {"workload": f"P1 unit square n x n, elastic 40 GPa disc r = 0.18 at the centre in a 10 GPa matrix "
                       f"(StiffnessAndFracture + NonLocalMazars(1e-4, Gf 100), radius 3/n, fibre-based damage), "
                       f"u_x = {loads} on the right edge, one FeatureTree step each; FIX48 solves to rel. res 1e-12",
           f"n{n}": big}
    small, d_dev = device(n_cpu)
    out[f"n{n_cpu}"] = small
    if not no_cpu:
        prob, dofs, nfix, crit, maz, radius = configs.event_square(n_cpu)
        d = np.zeros(prob.conn.shape[0])
        passes, t0 = 0, time.perf_counter()
        for v in loads:
            vals = np.zeros(len(dofs))
            vals[nfix:] = v
            d, _, _, pp = orc.event_damage(orc.P1, prob.conn, prob.coords, prob.cmats, prob.phase, crit, maz, dofs,
                                           vals, d, radius)
            passes += pp
        wall = time.perf_counter() - t0
        out["cpu_oracle"] = {"n": n_cpu, "passes": passes, "wall_s": wall, "ms_per_pass": 1e3 * wall / max(1, passes),
                             "cores": orc.get_threads(), "kind": "port",
                             "damage_equal_to_device": bool(np.abs(d - d_dev).max() <= 1e-12),
                             "note": "the oracle's loop: OpenMP PCG solves, brute-force neighbourhoods"}
        out["device_vs_oracle_same_case"] = wall / max(1e-9, small["wall_s"])
        log(f"event damage, oracle {n_cpu}^2: {out['cpu_oracle']}")
    newton, _ = device(n, ag.DAMAGE_ISOTROPIC_LINEAR)
    out[f"isotropic_linear_n{n}"] = newton
    log(f"event damage, IsotropicLinearDamage {n}^2: {newton}")
    if not no_cpu:
        first, d_dev = device(n_cpu, ag.DAMAGE_ISOTROPIC_LINEAR, loads[:1])
        prob, dofs, nfix, crit, maz, radius = configs.event_square(n_cpu)
        vals = np.zeros(len(dofs))
        vals[nfix:] = loads[0]
        t0 = time.perf_counter()
        d, _, _, pp = orc.event_damage(orc.P1, prob.conn, prob.coords, prob.cmats, prob.phase, crit, maz, dofs,
                                       vals, np.zeros(prob.conn.shape[0]), radius, model="newton")
        wall = time.perf_counter() - t0
        out[f"isotropic_linear_first_step_n{n_cpu}"] = {
            "device": first, "cpu_oracle": {"passes": pp, "wall_s": wall, "ms_per_pass": 1e3 * wall / max(1, pp),
                                            "cores": orc.get_threads(), "kind": "port"},
            "damage_max_diff": float(np.abs(d - d_dev).max()),
            "device_vs_oracle_same_case": wall / max(1e-9, first["wall_s"])}
        log(f"event damage, IsotropicLinearDamage first step {n_cpu}^2: {out[f'isotropic_linear_first_step_n{n_cpu}']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nrve", type=int, default=512)
    ap.add_argument("--nmax", type=int, default=256)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--rtol-sweep", default="5e-9,1e-9,1e-10", help="tighter stopping residuals timed and checked "
                                                                       "for the 1e-7 displacement target ('' to skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-persistent", action="store_true", help="skip the persistent-launch step")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--maxit", type=int, default=0, help="cap on CG iterations (profiling runs only)")
    ap.add_argument("--serve-rates", default="10:8,50:8,200:2,2000:0.2",
                    help="serving legs as RATE:SECONDS (Poisson arrivals in req/s over an arrival window)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the P1 128^2, P2 400^2 and Tet4 64^3 legs")
    ap.add_argument("--ingest-n", type=int, default=2048, help="mesh-input leg: structured P1 n x n (0: skip)")
    ap.add_argument("--event-n", type=int, default=32, help="event-driven damage leg: P1 n x n (0: skip)")
    ap.add_argument("--voxel-n", type=int, default=64, help="voxel-output leg: structured Tet4 n^3 on a (2n)^3 grid "
                                                            "(0: skip)")
    ap.add_argument("--spmv-kernel", default="", help="STEPS,ORDER[,FORMAT]: set the batch's FIX48 SpMV candidate "
                                                      "instead of autotuning (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import amie_gpu as ag
    import configs

    t0 = time.perf_counter()
    problems = configs.config4(args.nrve, seed=4, nmax=args.nmax)
    ndof = sum(p.ndof for p in problems)
    log(f"built {len(problems)} RVE meshes ({ndof} DOFs) in {time.perf_counter() - t0:.1f}s")
    ctx = ag.Context(0)
    t0 = time.perf_counter()
    B = ag.Batch(ctx, problems)
    log(f"batch resident in HBM: {B.nbytes / 2**30:.2f} GiB, setup {time.perf_counter() - t0:.1f}s")
    shapes = [B.shape(k) for k in range(len(problems))]
    nnzb = sum(s[1] for s in shapes)
    bytes8d = np.array([spmv_bytes_8d(nr, nz, 2, problems[0].nrhs) for nr, nz in shapes])
    hbm_bytes = B.nbytes
    # set-up: the autotuner picks the SpMV kernel of each precision once
    if args.spmv_kernel:
        sk = [int(v) for v in args.spmv_kernel.split(",")] + [0]
        B.set_spmv_kernel(ag.MIXED_FIX48, ag.JACOBI, sk[0], sk[1], sk[2])
        tuned = {"fix48": {"steps": sk[0], "tile_order": sk[1], "format": sk[2], "us": None}}
    else:
        tuned = {name: B.autotune(prec, ag.JACOBI) for name, prec in
                 (("fix48", ag.MIXED_FIX48), ("fp32", ag.MIXED), ("fp64", ag.FP64))}
    log(f"autotuned SpMV kernels: {tuned}")

    def step(precision, time_kernels=1, rtol=None, **kw):
        B.assemble()
        st = B.solve(precision=precision, rtol=rtol or args.rtol, time_kernels=time_kernels, use_graph=0,
                     maxit=args.maxit, **kw)
        ceff = B.average_stress()
        return st, ceff

    def launches_8d():
        """Algorithmic (FP32 yardstick) bytes of every SpMV launch of the last
        solve: a problem's launches are its iterations plus the one that
        finished it."""
        tot = 0.0
        for k in range(len(problems)):
            info = B.problem_info(k)
            tot += bytes8d[k] * (info["iterations"] + (1 if info["status"] == 0 else 0))
        return tot

    headline = ag.MIXED_FIX48
    for w in range(args.warmup):
        st, _ = step(headline)
        log(f"warmup {w}: {st['solve_ms']:.1f} ms solve, iters max {st['iterations_max']}, status {st['status']}")
    sts, b8 = [], 0.0
    elapsed = 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        st, ceff = step(headline)
        elapsed += time.perf_counter() - t0
        sts.append(st)
        b8 += launches_8d()
        log(f"step {k}: solve {st['solve_ms']:.1f} ms, iters max {st['iterations_max']}, relres max "
            f"{st['relres_max']:.2e}, converged {st['nproblems_converged']}/{len(problems)}")
    ms_per_step = 1e3 * elapsed / args.steps
    solves = len(problems) * problems[0].nrhs
    value = solves * args.steps / elapsed
    spmv_ms = sum(s["spmv_kernel_ms"] for s in sts)
    spmv_bytes = sum(s["spmv_bytes"] for s in sts)
    launches = sum(s["spmv_launches"] for s in sts)
    upd_ms = sum(s["update_kernel_ms"] for s in sts)
    upd_bytes = sum(s["update_bytes"] for s in sts)
    upd_launches = sum(s["update_launches"] for s in sts)
    achieved = b8 / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    achieved_phys = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    ok = all(s["status"] == 0 for s in sts)
    # accuracy sample: six RVEs at the quantiles of n, against the CPU oracle
    by_n = np.argsort([p.ndofnodes for p in problems], kind="stable")
    sample = sorted({int(by_n[int(q * (len(problems) - 1))]) for q in (0., .2, .4, .6, .8, 1.)})
    x_at = {args.rtol: {k: B.solution(k).copy() for k in sample}}
    rate_at = {args.rtol: value}

    def other(precision, name, **kw):
        t1 = time.perf_counter()
        st2, ceff2 = step(precision, time_kernels=0, **kw)
        dt = time.perf_counter() - t1
        log(f"{name} step: {1e3 * dt:.1f} ms, iters max {st2['iterations_max']}, status {st2['status']}")
        return {"solves_per_s": solves / dt, "ms_per_step": 1e3 * dt, "iterations_max": st2["iterations_max"],
                "iterations_sum": st2["iterations_sum"], "status": st2["status"]}, ceff2

    fp32, fp64, accuracy, sweep = None, None, None, None
    x_fp64 = None
    if not args.no_fp64:
        fp32, _ = other(ag.MIXED, "all-FP32 inner iteration")
        fp64, ceff64 = other(ag.FP64, "plain FP64 CG")
        x_fp64 = {k: B.solution(k).copy() for k in sample}
    for r in [float(v) for v in args.rtol_sweep.split(",") if v.strip()]:
        row, _ = other(headline, f"mixed at rel. res {r:g}", rtol=r)
        rate_at[r] = row["solves_per_s"]
        x_at[r] = {k: B.solution(k).copy() for k in sample}
    planes_step = None
    if tuned.get("fix48", {}).get("format") == 1:
        # the same step without the value index (what a mesh with more than
        # 65 536 distinct blocks, any damaged or unstructured one, runs)
        tf = tuned["fix48"]
        B.set_spmv_kernel(headline, ag.JACOBI, tf["steps"], tf["tile_order"], 0)
        planes_step, _ = other(headline, "mixed, FIX48 value planes (no value index)")
        B.set_spmv_kernel(headline, ag.JACOBI, tf["steps"], tf["tile_order"], 1)
    persistent = None
    if not args.no_persistent:
        # the same step as one persistent launch per solve pass (every
        # resident workgroup; device-wide barriers between the phases)
        row, _ = other(headline, "mixed, persistent launch", persistent=1 << 20)
        persistent = dict(row, note="one cooperative launch per solve pass, all resident workgroups, device-wide "
                                    "barriers between the phases; bitwise the launch path's results")
    if not args.no_cpu:
        xref = oracle_reference_solutions(problems, sample)
        accuracy = {"reference": "CPU oracle FP64 Jacobi PCG to rel. res 1e-12 (OpenMP, the oracle's own assembly)",
                    "sample_rves": [int(np.sqrt(problems[k].ndofnodes)) for k in sample],
                    "mixed_displacement_rel_err_max": {f"{r:g}": max(rel_err(x_at[r][k], xref[k]) for k in sample)
                                                       for r in x_at},
                    "solves_per_s": {f"{r:g}": rate_at[r] for r in rate_at}}
        if x_fp64 is not None:
            accuracy["fp64_displacement_rel_err_max"] = {
                f"{args.rtol:g}": max(rel_err(x_fp64[k], xref[k]) for k in sample)}
            accuracy["ceff_rel_diff_mixed_vs_fp64"] = float(np.abs(ceff - ceff64).max() / np.abs(ceff64).max())
        meets = sorted((r for r in x_at if accuracy["mixed_displacement_rel_err_max"][f"{r:g}"] <= 1e-7), reverse=True)
        accuracy["loosest_rtol_meeting_1e-7"] = meets[0] if meets else None
        accuracy["solves_per_s_at_loosest_rtol_meeting_1e-7"] = rate_at[meets[0]] if meets else None
        log(f"accuracy vs CPU oracle: {accuracy}")

    tr = measured_traffic()
    avg_bytes = spmv_bytes / max(1, launches)
    traffic = tr["ratio"] * avg_bytes if tr else None
    tr_u = measured_traffic("k_update")
    avg_upd = upd_bytes / max(1, upd_launches)
    upd_gbs = upd_bytes / (upd_ms * 1e-3) / 1e9 if upd_ms > 0 else 0.0
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 20)
    log(f"stream copy: {copy_gbs:.0f} GB/s")
    us_full, by_full = B.time_spmv(headline, ag.JACOBI, 20)
    by8_full = float(bytes8d.sum())
    log(f"k_spmv with every tile active: {us_full:.1f} us, {by_full / 1e9:.3f} GB stored format "
        f"({by_full / us_full / 1e3:.0f} GB/s), {by8_full / 1e9:.3f} GB FP32 yardstick ({by8_full / us_full / 1e3:.0f} GB/s)")
    tf = tuned["fix48"]
    planes = None
    if tf.get("format") == 1:
        # the same product from the FIX48 value planes, for comparison
        B.set_spmv_kernel(headline, ag.JACOBI, 4, 8, 0)
        us_pl, by_pl = B.time_spmv(headline, ag.JACOBI, 20)
        B.set_spmv_kernel(headline, ag.JACOBI, tf["steps"], tf["tile_order"], 1)
        planes = {"us": us_pl, "bytes": by_pl, "gbs": by_pl / us_pl / 1e3, "gbs_fp32_yardstick": by8_full / us_pl / 1e3}
        log(f"k_spmv from the FIX48 planes: {us_pl:.1f} us, {by_pl / 1e9:.3f} GB, {by_pl / us_pl / 1e3:.0f} GB/s")
    B.close()
    cfg1 = None if args.no_single else config1_leg(ag, configs, ctx, args.no_cpu)
    single = None if args.no_single else single_problem_legs(ag, configs, ctx, args.rtol, args.maxit, args.no_cpu)
    ingest = None
    if args.ingest_n > 0:
        ingest = ingest_leg(ag, ctx, args.ingest_n)
        log(f"mesh input: {ingest}")
    events = None
    if args.event_n > 0:
        events = event_leg(ag, configs, ctx, args.event_n, 24, args.no_cpu)
    voxels = None
    if args.voxel_n > 0:
        voxels = voxel_leg(ag, ctx, args.voxel_n, 2 * args.voxel_n)
        log(f"voxel output: {voxels}")
    serve = None
    if not args.no_serve:
        serve = {"workload": f"served mix (seed 5): Poisson arrivals, 1/3 each 2D P1 RVE n~U{{64..512}} (3 load "
                             f"cases), 3D Tet4 RVE n~U{{16..48}}, damage step of Tet4 32^3 on 4 persistent states; "
                             f"rel.res {args.rtol:g}; latency = arrival to result", "rates": []}
        for i, leg in enumerate(args.serve_rates.split(",")):
            rate, secs = (float(v) for v in leg.split(":"))
            reqs = configs.config5(rate, secs)
            kinds = {k: sum(1 for r in reqs if r[1] == k) for k in ("rve2d", "rve3d", "damage")}
            row = {"rate": rate, "seconds": secs, "requests": kinds}
            for name, prec in (("mixed", ag.MIXED_FIX48), ("fp64", ag.FP64)):
                row[name] = serving(ag, ctx, reqs, prec, args.rtol)
                log(f"serving {rate:g} req/s {name}: {row[name]}")
            if i == 0 and not args.no_persistent:
                row["mixed_persistent"] = serving(ag, ctx, reqs, ag.MIXED_FIX48, args.rtol, persistent=256)
                log(f"serving {rate:g} req/s mixed, persistent launches (256 workgroups): {row['mixed_persistent']}")
            serve["rates"].append(row)
    cpu = None if args.no_cpu else cpu_baseline(problems, args.cpu_budget, args.rtol)
    if cpu:
        cpu["gpu_dof_solves_per_s"] = ndof * problems[0].nrhs * args.steps / elapsed
    if ingest and not args.no_cpu:
        ingest["cpu_baseline"] = cpu_ingest(ag, args.ingest_n)
        log(f"mesh input, oracle: {ingest['cpu_baseline']}")
    if voxels and not args.no_cpu:
        voxels["cpu_baseline"] = cpu_voxels(ag, args.voxel_n, 2 * args.voxel_n)
        log(f"voxel output, oracle: {voxels['cpu_baseline']}")
    out = {
        "metric": "RVE solves/s + p50/p99 latency @rel.res 1e-8; mixed vs FP64 CG; SpMV GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(problems)} periodic P1 RVEs n~U{{32..{args.nmax}}} (seed 4), 3 macro-strain "
                               f"load cases each, mixed-precision CG (FP32 vectors, 48-bit fixed-point scaled "
                               f"matrix, FP64 accumulation and refinement) to rel.res {args.rtol:g}, "
                               f"assembly + solve + effective stiffness per step",
                   "rves": len(problems), "load_cases": problems[0].nrhs, "dofs": ndof, "blocks": nnzb,
                   "hbm_bytes": hbm_bytes, "converged": ok},
        "roofline": None,   # the dominant kernel's, set below
        "roofline_spmv": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes": "algorithmic bytes of the FP32 block SpMV per launch: nnzb (b^2 4 + 4) + 4 (nb + 1) + "
                              "2 n 4 per right-hand side, over the problems each launch works on",
                     "achieved_physical": achieved_phys,
                     "physical_bytes": "the stored format's own bytes per launch (2-byte value index + table share, "
                                       "the column indices it loads, row pointers, p, r read and q written)",
                     "traffic_source": (f"FETCH_SIZE+WRITE_SIZE / stored-format bytes = {tr['ratio']} at full "
                                        f"occupancy ({os.path.relpath(TRAFFIC_PROFILE, ROOT)}), times the average "
                                        f"stored-format bytes per launch" if tr else None),
                     "stream_copy_gbs": copy_gbs,
                     "kernel": "k_spmv (block SpMV q = Ks p, 48-bit fixed-point values"
                               + (", value-indexed: 16-bit index per block into a table of the distinct blocks"
                                  if tf.get("format") == 1 else "") + ", fused p.q, q.z, q.Mq)",
                     "autotuned": tuned,
                     "launches": launches,
                     "avg_launch_us": 1e3 * spmv_ms / max(1, launches),
                     "avg_bytes_per_launch": b8 / max(1, launches),
                     "avg_physical_bytes_per_launch": avg_bytes,
                     "full_occupancy": {"us": us_full, "bytes_fp32_yardstick": by8_full, "bytes_physical": by_full,
                                        "gbs": by8_full / us_full / 1e3, "gbs_physical": by_full / us_full / 1e3},
                     "fix48_planes_full_occupancy": planes},
        "cpu_baseline": cpu,
        "fp64_batched": fp64,
        "fp32_inner_batched": fp32,
        "mixed_vs_fp64": (value / fp64["solves_per_s"]) if fp64 else None,
        "accuracy": accuracy,
        "persistent_batched": persistent,
        "fix48_planes_batched": planes_step,
        "p1_128_patch": cfg1,
        "single_problems": single,
        "serving": serve,
        "mesh_input": ingest,
        "voxel_output": voxels,
        "event_damage": events,
    }
    update = {"bound": "hbm", "achieved": upd_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": upd_gbs / HBM_PEAK_GBS,
              "traffic": tr_u["ratio"] * avg_upd if tr_u else None,
              "bytes": "algorithmic bytes per launch: per DOF and right-hand side, p, q, r read and r, p written "
                       "(FP32), and every fourth launch p_{k-3} .. p_{k-1} read and y read and written (FP64, "
                       "y += alpha p grouped): 20 and 48 B, 27 B on average; a refinement's slot update, which "
                       "closes a group, 32 B (p_{k-3} .. p_k read, y read and written); over the problems each "
                       "launch works on",
              "traffic_source": (f"FETCH_SIZE+WRITE_SIZE / algorithmic bytes = {tr_u['ratio']} at full occupancy "
                                 f"({os.path.relpath(TRAFFIC_PROFILE, ROOT)}), times the average algorithmic bytes "
                                 f"per launch" if tr_u else None),
              "stream_copy_gbs": copy_gbs,
              "kernel": "k_update (CG update: y += alpha p, r -= alpha q, p = r + beta p, fused r.r partials)",
              "launches": upd_launches, "avg_launch_us": 1e3 * upd_ms / max(1, upd_launches),
              "avg_bytes_per_launch": avg_upd, "kernel_ms": upd_ms, "spmv_kernel_ms": spmv_ms}
    out["roofline_update"] = update
    # the roofline line is the kernel with the larger share of the timed steps
    out["roofline"] = dict(update) if upd_ms > spmv_ms else dict(out["roofline_spmv"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
