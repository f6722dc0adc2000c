"""Benchmark of the colored LU-SGS hot path on one MI355X.

Headline workload (the configuration the metric is quoted on): colored
LU-SGS, 3-D Euler (nvars = 5), FP64, on a 96^3 hex grid split into 6 Kuhn
tets per hex (5,308,416 cells, nface = 4), cell order randomly permuted from
the seed, normals / areas / volumes from jittered vertex coordinates, fields
as in ucfd_sparse_amd.mesh.flow_fields.  The mesh is coloured by first-fit in
breadth-first order (2 colours: the Kuhn face graph is bipartite), stored
colour-contiguously in Hilbert order and declared to the library
(ucfd_declare_colored_layout).  One step = one legacy-API pass:
parallel_pre_lusgs, ns_parallel_lower_sweep for every colour,
ns_parallel_upper_sweep for every colour in reverse and
lusgs_parallel_ns_update, all on device-resident arrays on one stream.  The
two-array API keeps the right-hand side intact, so no restore copy sits in
the timed region.

Beside the headline the line carries: the dominant kernel's roofline, the
pass on two other colourings (the first-fit routine of the library and a
4-colour layout), the FP64 and mixed FP32/FP64 defect-correction solves to a
1e-10 residual drop with per-kernel bandwidth, the CPU baseline (the oracle's
OpenMP pass, two builds) with the GPU pass checked against the oracle, and
summaries of the block LU-SGS (96^3 x 6), Krylov (70^3 x 6) and adjoint
(55^3 x 6) workloads.  --quick keeps only the headline pass.

Prints one JSON line.  value = cell-iterations/s of the FP64 path.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# OpenMP placement of the CPU baseline: set before any OpenMP runtime loads
os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_PROC_BIND", "close")
# the CPUs this process may use, read before an OpenMP runtime binds the
# initial thread to its first place
_AFFINITY = sorted(os.sched_getaffinity(0))

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "LU-SGS cell-iters/s & HBM GB/s, FP32-mixed vs FP64 at equal residual drop"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(F, V=5, D=3, s=8):
    """Per-cell bytes of each legacy call: every input array element read
    once, every output written once, indices 4 B (colour arrays excluded)."""
    pre = s * (2 * F + 2)
    sweep = 4 * F + s * (2 * F + F * D + 1 + 4 * V)
    update = 3 * V * s
    return pre, sweep, update


def read_traffic(path):
    """key = value lines written from the rocprofv3 --pmc runs (profiles/):
    numbers as floats, anything else as text."""
    out = {}
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if "=" in line:
            k, v = (t.strip() for t in line.split("=", 1))
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out


def cpu_info():
    """CPU model, logical CPUs this process may use and the physical cores
    among them (sysfs topology)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cpus = _AFFINITY
    cores = set()
    for c in cpus:
        try:
            pkg = open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id").read()
            core = open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id").read()
            cores.add((pkg.strip(), core.strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores)}


def cpu_baseline(lay, seconds, gpu_pass=None, nvars=5):
    """The CPU baseline and the parity check of the measured GPU pass.

    Parity: one pass of the oracle's parity build (oracle/liboracle.so: -O2,
    no fast-math, no FMA contraction, bit-identical to the reference's own
    code) from the same initial state as gpu_pass (diag, dQ*, dQ, upts of one
    GPU pass), compared element by element.
    Baseline: the oracle's OpenMP coloured pass on the same renumbered arrays,
    pre + lower + upper + update per iteration, built twice: -O3
    -march=x86-64-v3 -fopenmp (liboracle_fast.so) and with the reference's
    own flags -O3 -ffast-math -fopenmp (liboracle_ref.so, Makefile.inc:11-16);
    3 warm-up iterations, then at least 10 timed iterations (about seconds/2
    per build), median reported, rhsb restored before each iteration."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as O
    m, f = lay.mesh, lay.flow
    N, F = m.neles, m.nface
    parity = None
    if gpu_pass is not None:
        od, os_, oq, ou = O.colored_pass(m.nei_ele, lay.icolor, lay.lcolor, m.fnorm_vol,
                                         m.vec_fnorm, f.fspr, f.dt, f.upts, f.rhsb)
        errs = {}
        bitwise = True
        for name, g, o in (("diag", gpu_pass[0], od), ("dstar", gpu_pass[1], os_),
                           ("dq", gpu_pass[2], oq), ("upts", gpu_pass[3], ou)):
            errs[name] = float(np.abs(g - o).max() / max(np.abs(o).max(), 1e-300))
            bitwise = bitwise and bool(np.array_equal(g, o))
        parity = {"max_rel_err": max(errs.values()), "per_array": errs, "bitwise": bitwise,
                  "checked": f"one pass of the timed configuration ({N} cells, "
                             f"{lay.ncolors} colours) against the oracle's parity build"}
    off = np.ascontiguousarray(lay.off, np.int32)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(_AFFINITY)
    builds = {}
    for key, build, flags in (("fast", "fast", "-O3 -march=x86-64-v3 -fopenmp"),
                              ("reference_flags", "ref", "-O3 -ffast-math -fopenmp")):
        lib = O.lib(build=build)
        upts = f.upts.copy()
        dub = f.rhsb.copy()
        diag = np.empty(N)
        args = (N, F, 3, 5, nvars, lay.ncolors, off.ctypes.data, m.nei_ele.ctypes.data,
                lay.icolor.ctypes.data, lay.lcolor.ctypes.data, m.fnorm_vol.ctypes.data,
                m.vec_fnorm.ctypes.data, upts.ctypes.data, dub.ctypes.data, diag.ctypes.data,
                f.fspr.ctypes.data, f.dt.ctypes.data)
        for _ in range(3):
            dub[...] = f.rhsb
            lib.orc_colored_iteration(*args)
        times = []
        while sum(times) < seconds / 2 or len(times) < 10:
            dub[...] = f.rhsb
            t0 = time.perf_counter()
            lib.orc_colored_iteration(*args)
            times.append(time.perf_counter() - t0)
        med = sorted(times)[len(times) // 2]
        builds[key] = {"cell_iters_per_s": N / med, "median_ms": med * 1e3,
                       "iterations": len(times), "flags": flags,
                       "gbs_algorithmic": 888 * N / med / 1e9}
    info = cpu_info()
    fast = builds["fast"]
    return {"value": fast["cell_iters_per_s"], "unit": "cell-iter/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {fast['iterations']} coloured FP64 iterations (pre + lower + "
                      f"upper + update) of the full {N}-cell workload, oracle/lusgs_oracle.c "
                      f"built {fast['flags']}, {threads} OpenMP threads "
                      f"(OMP_PLACES={os.environ.get('OMP_PLACES')}, "
                      f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')})",
            "cpu_model": info["model"], "logical_cpus": info["logical_cpus"],
            "physical_cores": info["physical_cores"], "builds": builds}, parity


# Algorithmic bytes per cell of one pass of the defect-correction kernels
# (every array a kernel touches counted once per pass; neighbour reads of the
# array the kernel writes counted as one more read, as for the legacy sweeps;
# s = bytes per real, packed face records of 4 reals per face).
def solver_kernel_bytes(F, s):
    lower = 4 * F + s * (4 * F + 26)      # nei, geo, diag, rhs, e, qb, dstar read + write
    upper = 4 * F + s * (4 * F + 31)      # nei, geo, diag, qb, dstar, delta r+w, e r+w
    return lower, upper


def precision_forms(lay, T, rtol, repeats, log):
    """Time-to-rtol of the FP64 and mixed FP32/FP64 defect-correction forms on
    the same device-resident system; error of the mixed result vs FP64; HIP
    event times of one pass of each solver kernel (FP64 and FP32) with their
    algorithmic bandwidth."""
    import ctypes

    import torch

    from ucfd_sparse_amd import lib as L
    from ucfd_sparse_amd.solver import LUSGSSolver
    m = lay.mesh
    N, F = m.neles, m.nface
    s = LUSGSSolver(N, F, lay.off, T["nei"], T["fnv"], T["vfn"], T["fsp"], T["dt"], T["upts"])
    b = T["rhs"]
    out = {}
    res = {}
    for mode in ("fp64", "mixed"):
        dq = torch.zeros_like(b)
        st = s.solve(b, dq, mode, rtol, 400)          # warm-up
        times = []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = s.solve(b, dq, mode, rtol, 400)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        times.sort()
        res[mode] = dq
        out[mode] = {"ms": times[len(times) // 2], "passes": st.passes, "outer": st.outer,
                     "rel_residual": st.rel_residual, "converged": bool(st.converged)}
        log(f"{mode}: {out[mode]}")
    ms = (ctypes.c_double * 5)()
    L.check(L.load().ucfd_lusgs_solver_time_kernels(s.handle, b.data_ptr(), 10, ms),
            "ucfd_lusgs_solver_time_kernels")
    s.close()
    a, bm = res["fp64"], res["mixed"]
    err = float((bm - a).abs().max() / a.abs().max())
    k64 = out["fp64"]["passes"]
    kern = {}
    # one pass = the lower launches of every colour but the first (residual
    # and lower sweep; with more than one colour the last colour's launch
    # also does its upper step) and the upper launches of every colour but
    # the last (upper sweep, update and, in the first colour's launch, that
    # colour's residual and lower step for the next pass): bandwidth over
    # the whole pass
    for key, lo, sz in (("pass_fp64", 0, 8), ("pass_fp32", 2, 4)):
        lb, ub = solver_kernel_bytes(F, sz)
        nb = (lb + ub) * N
        t = ms[lo] + ms[lo + 1]
        gbs = nb / (t * 1e-3) / 1e9
        kern[key] = {"ms_per_pass": t, "lower_launches_ms": ms[lo], "upper_launches_ms": ms[lo + 1],
                     "algorithmic_bytes": nb, "gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                     "launches": max(2 * lay.ncolors - 2, 1)}
    rb = (4 * F + 8 * (4 * F + 1 + 15)) * N        # nei, geo, diag, q, b, dq
    kern["residual_fp64"] = {"ms_per_pass": ms[4], "algorithmic_bytes": rb,
                             "gbs": rb / (ms[4] * 1e-3) / 1e9,
                             "frac": rb / (ms[4] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    p64 = ms[0] + ms[1]
    p32 = ms[2] + ms[3]
    log("solver kernels: " + ", ".join(f"{k} {v['ms_per_pass']:.3f} ms {v['gbs']:.0f} GB/s"
                                       for k, v in kern.items()))
    return {
        "rtol": rtol,
        "fp64": out["fp64"],
        "mixed": out["mixed"],
        "speedup_mixed_vs_fp64": out["fp64"]["ms"] / out["mixed"]["ms"],
        "fp64_equivalent_cell_iters_per_s": {"fp64": N * k64 / (out["fp64"]["ms"] * 1e-3),
                                             "mixed": N * k64 / (out["mixed"]["ms"] * 1e-3)},
        "mixed_max_rel_err_vs_fp64": err,
        "kernels": kern,
        "pass_ms": {"fp64": p64, "fp32": p32, "fp32_speedup": p64 / p32},
        "note": "defect correction A(dq) = b with one coloured LU-SGS pass as preconditioner; "
                "mixed = FP32 sweeps inside an FP64 outer residual loop, dq accumulated in FP64; "
                "algorithmic bytes per pass: lower 4F + s(4F+26) plus upper 4F + s(4F+31) per "
                "cell (s = 8 FP64, 4 FP32)",
    }


def krylov_bytes(A, nnz, n, fp32):
    """Algorithmic bytes of one BSR SpMV (beta = 0) and of one coloured block
    LU-SGS apply: every matrix entry and index read once, vectors read or
    written once per pass that needs them."""
    bn, blk = A.bn, A.blk
    b2 = blk * blk
    spmv = nnz * (b2 * 8 + 4) + (bn + 1) * 4 + 2 * n * 8
    vs = 4 if fp32 else 8
    off = nnz - bn
    apply = (off * (b2 * vs + 4)            # off-diagonal blocks + column indices
             + 2 * bn * (b2 * 8 + 4 + 2 * 4)  # diagonal LU, diagslots, rowptr pair, per sweep
             + n * (8 + vs) + n * (vs + vs + 8)   # own rows: lower b->x, upper x->x,b
             + 2 * n * vs)                    # neighbour values, once per sweep
    return spmv, apply


def krylov_workload(args, log):
    """GMRES(30) and BiCGSTAB on the block Jacobian of a 70^3 x 6 tet mesh
    (2,058,000 cells, 5x5 blocks), right-preconditioned by the coloured block
    LU-SGS with the FP32 apply (FP64 diagonal LU), FP64 Krylov vectors, dot
    products and Givens rotations, rtol 1e-10, 4 seeded N(0,1) right-hand
    sides; all arrays device-resident.  Also times the SpMV and the apply
    kernels, the all-FP64 preconditioner and the coloured block ILU(0)
    (factorisation, apply, and as preconditioner of the same solves) for
    comparison."""
    import numpy as np
    import torch

    from ucfd_sparse_amd import krylov as K
    from ucfd_sparse_amd import lib as L
    from ucfd_sparse_amd import mesh as M
    from ucfd_sparse_amd.layout import colored_layout

    m, seed = args.krylov_m, 3
    t0 = time.time()
    mesh = M.tet_mesh(m, seed=seed)
    flow = M.flow_fields(mesh, seed=seed)
    mesh.key = M.hilbert_keys(mesh, "tet", (m,))
    lay = colored_layout(mesh, flow, coloring=args.coloring)
    A = M.block_system(lay.mesh, lay.flow, blk=5, seed=seed)
    n = A.bn * A.blk
    nnz = int(A.rowptr[-1])
    log(f"krylov system: bn={A.bn} nnz={nnz} colours={lay.ncolors}, setup {time.time() - t0:.1f} s")
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    class D:
        pass
    Dv = D()
    Dv.bn, Dv.blk = A.bn, A.blk
    Dv.rowptr, Dv.colidx, Dv.diagslots, Dv.values = (up(A.rowptr), up(A.colidx),
                                                     up(A.diagslots), up(A.values))
    rhs = [up(M.normal_fill(n, 1000 + k)) for k in range(args.krylov_rhs)]
    del A.values
    mat = K.SpMat.from_bsr(Dv)
    pcs = {"fp32": K.Precon.from_bsr(Dv, icolors=lay.off, fp32=True),
           "fp64": K.Precon.from_bsr(Dv, icolors=lay.off)}
    # block ILU(0) factors its values in place: a copy of the matrix values
    Di = D()
    Di.bn, Di.blk, Di.rowptr, Di.colidx, Di.diagslots = Dv.bn, Dv.blk, Dv.rowptr, Dv.colidx, \
        Dv.diagslots
    Di.values = Dv.values.clone()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pcs["pbilu"] = K.Precon.from_bsr(Di, icolors=lay.off, kind="bilu")
    torch.cuda.synchronize()
    ilu_setup_ms = (time.perf_counter() - t1) * 1e3

    # a later Prepare on new values (as once per Newton step): the schedule is
    # reused, only the values are read
    def reprepare(pc, fresh=None):
        if fresh is not None:
            Di.values.copy_(Dv.values)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        pc.prepare()
        torch.cuda.synchronize()
        return (time.perf_counter() - t2) * 1e3
    reprep = {"blusgs_fp64": reprepare(pcs["fp64"]), "blusgs_fp32": reprepare(pcs["fp32"]),
              "pbilu": reprepare(pcs["pbilu"], fresh=True)}
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    L.call("ucfd_set_stream", stream.cuda_stream)

    def timed(fn, reps):
        fn()
        L.call("ucfd_synchronize")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        L.call("ucfd_synchronize")
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    x = torch.zeros(n, dtype=torch.float64, device=dev)
    y = torch.zeros_like(x)
    b_spmv, b_ap32 = krylov_bytes(Dv, nnz, n, True)
    _, b_ap64 = krylov_bytes(Dv, nnz, n, False)
    # the block ILU(0) solve reads no diagonal block in its lower sweep
    b_ilu = b_ap64 - Dv.bn * Dv.blk * Dv.blk * 8
    kern = {}
    ms = timed(lambda: mat.mult(rhs[0], y), 20)
    kern["spmv_bsr5"] = {"ms": ms, "algorithmic_bytes": b_spmv, "gbs": b_spmv / ms / 1e6}
    w = rhs[0].clone()
    for key, name, bb in (("blusgs_apply_fp32", "fp32", b_ap32),
                          ("blusgs_apply_fp64", "fp64", b_ap64),
                          ("pbilu_apply_fp64", "pbilu", b_ilu)):
        ms = timed(lambda: pcs[name].apply(w), 10)
        kern[key] = {"ms": ms, "algorithmic_bytes": bb, "gbs": bb / ms / 1e6,
                     "launches": 2 * lay.ncolors}
    for k, v in kern.items():
        log(f"{k}: {v['ms']:.3f} ms, {v['gbs']:.0f} GB/s algorithmic")
    kern["pbilu_prepare"] = {"ms": ilu_setup_ms, "launches": lay.ncolors,
                             "note": "wall time of the first UCFDPreconPrepare: structure "
                                     "checks, schedule and the in-place factorisation"}
    kern["reprepare_wall_ms"] = reprep
    log(f"pbilu prepare (schedule + factorisation): {ilu_setup_ms:.1f} ms; later Prepares "
        + ", ".join(f"{k} {v:.2f} ms" for k, v in reprep.items()))

    # the solves and the residual checks below run on the default stream
    # (library and torch alike)
    L.call("ucfd_set_stream", None)
    torch.cuda.synchronize()
    solves = {}
    for kind in ("gmres", "bicgstab"):
        for pname in ("fp32", "fp64", "pbilu"):
            s = K.Solver(kind, n, 200 if kind == "gmres" else 2000, args.rtol, m=30, right=True)
            rec = []
            for k, b in enumerate(rhs):
                x.zero_()
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                st, it, res = s.solve(pcs[pname], mat, x, b)
                torch.cuda.synchronize()
                el = (time.perf_counter() - t1) * 1e3
                r = b.clone()
                mat.mult(x, r, alpha=-1.0, beta=1.0)
                L.call("ucfd_synchronize")
                rel = float(r.norm() / b.norm())
                rec.append({"ms": el, "iterations": it, "status": st, "true_rel_residual": rel})
            s.close()
            key = f"{kind}_{pname}"
            solves[key] = {"ms_mean": sum(r["ms"] for r in rec) / len(rec), "per_rhs": rec}
            rels = ", ".join(f"{r['true_rel_residual']:.1e}" for r in rec)
            log(f"{key}: {solves[key]['ms_mean']:.1f} ms mean, iterations "
                f"{[r['iterations'] for r in rec]}, status {[r['status'] for r in rec]}, "
                f"true rel residuals {rels}")
    g32 = solves["gmres_fp32"]["ms_mean"]
    rec = {
        "metric": "Krylov time to rtol (GMRES(30), right-preconditioned coloured block LU-SGS)",
        "value": g32,
        "unit": "ms per solve",
        "n_gpus": 1,
        "steps": len(rhs),
        "warmup": 0,
        "ms_per_step": g32,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (Krylov), f32 (preconditioner products)",
        "data": "synthetic (seeded Kuhn-tet mesh, block Jacobian ucfdgen_bsr_fill, N(0,1) rhs)",
        "config": {"workload": f"GMRES(30)/BiCGSTAB, {m}^3 hex x 6 tets, BSR 5x5, coloured "
                               f"block LU-SGS right preconditioner, rtol {args.rtol}",
                   "bn": Dv.bn, "nnz_blocks": nnz, "ncolors": lay.ncolors},
        "kernels": kern,
        "solves": solves,
    }
    return rec


def flowsys_bytes(npos, N, nv, fp32):
    """Algorithmic bytes of one FlowSys block sweep (all levels) with nv x nv
    blocks: per position the column index, the packed neighbour block, fv and
    the neighbour's du row (FP32 copies in the FP32 form); per cell rowptr,
    its rhs row, its FP64 diagonal LU and the du row written (plus the FP32
    shadow row in the FP32 form)."""
    vs = 4 if fp32 else 8
    per_pos = 4 + vs * nv * nv + vs + vs * nv
    per_cell = 4 + 8 * nv + 8 * nv * nv + 8 * nv + (4 * nv if fp32 else 0)
    return npos * per_pos + N * per_cell + 4


def blusgs_workload(args, log):
    """Block LU-SGS at scale: the FlowSys API (NS 5x5 blocks plus the 1x1
    Spalart-Allmaras turbulence system) on the 96^3 x 6 tet mesh, FP64
    factors with FP64 sweeps vs FP32 sweep products, and the Krylov-path
    coloured block LU-SGS apply on the BSR block Jacobian of the same mesh
    (FP64 vs FP32 apply).  All arrays device-resident; colour-contiguous
    layout, so each sweep is one launch per colour."""
    import numpy as np
    import torch

    from ucfd_sparse_amd import flowsys as FS
    from ucfd_sparse_amd import krylov as K
    from ucfd_sparse_amd import lib as L
    from ucfd_sparse_amd import mesh as M
    from ucfd_sparse_amd.layout import colored_layout
    from ucfd_sparse_amd.mesh import FlowSysMesh

    t0 = time.time()
    mesh = M.tet_mesh(args.m, seed=args.seed)
    flow = M.flow_fields(mesh, seed=args.seed)
    mesh.key = M.hilbert_keys(mesh, "tet", (args.m,))
    lay = colored_layout(mesh, flow, coloring=args.coloring)
    del mesh, flow
    m, f = lay.mesh, lay.flow
    N, nc = m.neles, lay.ncolors
    fs = M.flowsys_system(m, f, nv=5, seed=args.seed)
    tj = M.flowsys_system(m, f, nv=1, seed=args.seed).jmat
    el = M.flowsys_elements(m, f, nvars=6, seed=args.seed)
    npos = int(fs.rowptr[-1])
    log(f"blusgs: {N} cells, {npos} positions, {nc} colours, setup {time.time() - t0:.1f} s")
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fs_nfaces = fs.nfaces
    fsd = FlowSysMesh(fs.nlocal, fs.nfaces, up(fs.rowptr), up(fs.colidx), up(fs.slots),
                      up(fs.sides), up(fs.face_area), up(fs.face_normal), up(fs.rcp_vol),
                      up(fs.jmat))
    tjd = up(tj)
    del fs, tj
    stream = torch.cuda.Stream(device=dev)

    def ev():
        return torch.cuda.Event(enable_timing=True)

    flows = {}
    du_after = {}
    for prec in ("fp64", "fp32"):
        fp32 = prec == "fp32"
        S = FS.FlowSys(fsd, 6, 5, tjmat=tjd, fp32=fp32)
        S.set_element(0, {k: up(v) for k, v in el.items()})
        torch.cuda.synchronize()
        L.call("ucfd_set_stream", stream.cuda_stream)
        # the first Prepare also validates the structure and builds the sweep
        # schedule on the host (once per handle); the second one is timed
        t1 = time.perf_counter()
        S.call("UCFDBLUSGS_Pack", 0, 0.3)
        S.call("UCFDBLUSGS_SA_Pack", 0, 1.5, 0.3)
        S.call("UCFDBLUSGS_NSPrepare")
        S.call("UCFDBLUSGS_RANSPrepare")
        L.call("ucfd_synchronize")
        first_ms = (time.perf_counter() - t1) * 1e3
        e = [ev() for _ in range(4)]
        e[0].record(stream)
        S.call("UCFDBLUSGS_Pack", 0, 0.3)
        S.call("UCFDBLUSGS_SA_Pack", 0, 1.5, 0.3)
        e[1].record(stream)
        S.call("UCFDBLUSGS_NSPrepare")
        e[2].record(stream)
        S.call("UCFDBLUSGS_RANSPrepare")
        e[3].record(stream)
        L.call("ucfd_synchronize")
        prep = {"first_pack_prepare_with_structure_ms": first_ms,
                "pack_ms": e[0].elapsed_time(e[1]),
                "ns_prepare_ms": e[1].elapsed_time(e[2]),
                "rans_prepare_ms": e[2].elapsed_time(e[3])}
        S.call("UCFDBLUSGS_Reset")

        def it(evs=None):
            if evs:
                evs[0].record(stream)
            S.call("UCFDBLUSGS_NSLowerSweep")
            if evs:
                evs[1].record(stream)
            S.call("UCFDBLUSGS_NSUpperSweep")
            if evs:
                evs[2].record(stream)
            S.call("UCFDBLUSGS_RANSLowerSweep")
            S.call("UCFDBLUSGS_RANSUpperSweep")
            if evs:
                evs[3].record(stream)
        it()
        L.call("ucfd_synchronize")
        du_after[(prec, 1)] = S.export(FS.DU)
        for _ in range(4):
            it()
        L.call("ucfd_synchronize")
        du_after[(prec, 5)] = S.export(FS.DU)
        reps = args.steps
        evs = [[ev() for _ in range(4)] for _ in range(reps)]
        for r in range(reps):
            it(evs[r])
        L.call("ucfd_synchronize")
        torch.cuda.synchronize()
        L.call("ucfd_set_stream", None)
        lo = sum(x[0].elapsed_time(x[1]) for x in evs) / reps
        upp = sum(x[1].elapsed_time(x[2]) for x in evs) / reps
        rans = sum(x[2].elapsed_time(x[3]) for x in evs) / reps
        S.close()
        b_ns = flowsys_bytes(npos, N, 5, fp32)
        b_rans = flowsys_bytes(npos, N, 1, fp32)
        flows[prec] = dict(prep, **{
            "ns_lower_ms": lo, "ns_upper_ms": upp, "rans_lower_upper_ms": rans,
            "ns_cell_iters_per_s": N / ((lo + upp) * 1e-3),
            "ns_sweep_algorithmic_bytes": b_ns,
            "ns_sweep_gbs": b_ns / (0.5 * (lo + upp)) / 1e6,
            "ns_launch_ms": 0.5 * (lo + upp) / nc,
            "rans_sweep_algorithmic_bytes": b_rans,
            "rans_sweep_gbs": 2 * b_rans / rans / 1e6})
        log(f"flowsys {prec}: NS lower {lo:.3f} upper {upp:.3f} ms "
            f"({flows[prec]['ns_sweep_gbs']:.0f} GB/s alg), RANS lower+upper {rans:.3f} ms, "
            f"pack {prep['pack_ms']:.2f}, prepare NS {prep['ns_prepare_ms']:.2f} RANS "
            f"{prep['rans_prepare_ms']:.2f} ms (first call with structure "
            f"{prep['first_pack_prepare_with_structure_ms']:.0f} ms)")

    def rel(a, b, sl):
        a = a.reshape(N, 6)[:, sl]
        b = b.reshape(N, 6)[:, sl]
        return float(np.linalg.norm(a - b) / np.linalg.norm(a))
    errs = {f"{part}_after_{k}": rel(du_after[("fp64", k)], du_after[("fp32", k)], sl)
            for k in (1, 5) for part, sl in (("ns", slice(0, 5)), ("rans", slice(5, 6)))}
    log(f"flowsys fp32 vs fp64 du: {errs}")
    del tjd
    torch.cuda.empty_cache()

    # the FlowSys scalar LU-SGS method on the same structure: per-slot
    # spectral radii, NS rows + 1 SA row
    rng = np.random.default_rng(args.seed + 11)
    fspr_s = up(rng.uniform(0.5, 2.0, size=fsd.nfaces))
    tfspr_s = up(rng.uniform(0.2, 1.0, size=fsd.nfaces))
    SL = FS.FlowSysLUSGS(fsd, 6, 5, fspr_s, tfspr_s)
    SL.set_element(0, {k: up(v) for k, v in el.items()})
    torch.cuda.synchronize()
    L.call("ucfd_set_stream", stream.cuda_stream)
    SL.call("UCFDLUSGS_Pack", 0, 1.5, 0.3)
    SL.call("UCFDLUSGS_NSPrepare", 1.0)
    SL.call("UCFDLUSGS_RANSPrepare", 1.0)
    L.call("ucfd_synchronize")
    ep = [ev() for _ in range(4)]
    ep[0].record(stream)
    SL.call("UCFDLUSGS_Pack", 0, 1.5, 0.3)
    ep[1].record(stream)
    SL.call("UCFDLUSGS_NSPrepare", 1.0)
    ep[2].record(stream)
    SL.call("UCFDLUSGS_RANSPrepare", 1.0)
    ep[3].record(stream)
    L.call("ucfd_synchronize")
    scalar_prep = {"pack_ms": ep[0].elapsed_time(ep[1]), "ns_prepare_ms": ep[1].elapsed_time(ep[2]),
                   "rans_prepare_ms": ep[2].elapsed_time(ep[3])}

    def lit(evs=None):
        if evs:
            evs[0].record(stream)
        SL.call("UCFDLUSGS_NSLowerSweep", 1.0)
        SL.call("UCFDLUSGS_NSUpperSweep", 1.0)
        if evs:
            evs[1].record(stream)
        SL.call("UCFDLUSGS_RANSLowerSweep", 1.0)
        SL.call("UCFDLUSGS_RANSUpperSweep", 1.0)
        if evs:
            evs[2].record(stream)
    for _ in range(3):
        lit()
    L.call("ucfd_synchronize")
    evl = [[ev() for _ in range(3)] for _ in range(args.steps)]
    for r in range(args.steps):
        lit(evl[r])
    # Update (uptsb += du per element), timed once after the sweeps
    eu = [ev() for _ in range(2)]
    eu[0].record(stream)
    SL.call("UCFDLUSGS_Update", 0)
    eu[1].record(stream)
    L.call("ucfd_synchronize")
    torch.cuda.synchronize()
    L.call("ucfd_set_stream", None)
    ns_l = sum(x[0].elapsed_time(x[1]) for x in evl) / args.steps
    rans_l = sum(x[1].elapsed_time(x[2]) for x in evl) / args.steps
    scalar_prep["update_ms"] = eu[0].elapsed_time(eu[1])
    SL.close()
    # algorithmic bytes of one NS scalar sweep, every array element counted
    # once: per cell rowptr, rcp_vol and the 5 rows of diag, u, du (read) and
    # du (written); per position colidx, slot, side; per face slot the normal
    # (3), area and spectral radius
    b_scalar = N * (4 + 8 + 8 * 5 * 4) + npos * (4 + 4 + 1) + fs_nfaces * 8 * 5
    # the pack: per element the 6 rows of upts, rhs and dsrc and dt read,
    # the rank rows of u, du and diag written, cell_ids read
    b_pack = N * (4 + 8 + 6 * 8 * 3 + 6 * 8 * 3)
    scalar_prep["pack_algorithmic_bytes"] = b_pack
    scalar_prep["pack_gbs"] = b_pack / (scalar_prep["pack_ms"] * 1e-3) / 1e9
    scalar = {"ns_lower_upper_ms": ns_l, "rans_lower_upper_ms": rans_l, **scalar_prep,
              "ns_cell_iters_per_s": N / (ns_l * 1e-3),
              "ns_sweep_algorithmic_bytes": b_scalar,
              "ns_sweep_gbs": b_scalar / (0.5 * ns_l) / 1e6}
    log(f"flowsys scalar LU-SGS: NS lower+upper {ns_l:.3f} ms "
        f"({scalar['ns_sweep_gbs']:.0f} GB/s alg), RANS lower+upper {rans_l:.3f} ms, pack "
        f"{scalar['pack_ms']:.3f} ms ({scalar['pack_gbs']:.0f} GB/s alg), prepare NS "
        f"{scalar['ns_prepare_ms']:.3f} RANS {scalar['rans_prepare_ms']:.3f} ms, update "
        f"{scalar['update_ms']:.3f} ms")
    del fsd
    torch.cuda.empty_cache()

    # Krylov-path coloured block LU-SGS apply on the BSR block Jacobian
    A = M.block_system(m, f, blk=5, seed=args.seed)
    nnz = int(A.rowptr[-1])

    class Dv:
        pass
    Dv.bn, Dv.blk = A.bn, A.blk
    Dv.rowptr, Dv.colidx, Dv.diagslots, Dv.values = (up(A.rowptr), up(A.colidx),
                                                     up(A.diagslots), up(A.values))
    del A
    n = Dv.bn * Dv.blk
    b = up(M.normal_fill(n, 2024))
    L.call("ucfd_set_stream", stream.cuda_stream)
    applies = {}
    outs = {}
    for prec in ("fp64", "fp32"):
        pc = K.Precon.from_bsr(Dv, icolors=lay.off, fp32=prec == "fp32")
        w = b.clone()
        pc.apply(w)
        L.call("ucfd_synchronize")
        outs[prec] = w.clone()
        e0, e1 = ev(), ev()
        e0.record(stream)
        for _ in range(args.steps):
            pc.apply(w)
        e1.record(stream)
        L.call("ucfd_synchronize")
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        bb = krylov_bytes(Dv, nnz, n, prec == "fp32")[1]
        applies[prec] = {"ms": ms, "algorithmic_bytes": bb, "gbs": bb / ms / 1e6,
                         "launches": 2 * nc}
        pc.close()
        log(f"block LU-SGS apply {prec}: {ms:.3f} ms, {bb / ms / 1e6:.0f} GB/s algorithmic")
    L.call("ucfd_set_stream", None)
    apply_err = float((outs["fp32"] - outs["fp64"]).norm() / outs["fp64"].norm())
    f64, f32 = flows["fp64"], flows["fp32"]
    ns_ms64 = f64["ns_lower_ms"] + f64["ns_upper_ms"]
    ns_ms32 = f32["ns_lower_ms"] + f32["ns_upper_ms"]
    launch_ms = f64["ns_launch_ms"]
    lb = f64["ns_sweep_algorithmic_bytes"] / nc
    rec = {
        "metric": "block LU-SGS cell-iters/s (NS 5x5 lower + upper sweeps), FP32 sweeps vs FP64",
        "value": f64["ns_cell_iters_per_s"],
        "unit": "cell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 5,
        "ms_per_step": ns_ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (factors and sweeps); fp32 sweep form reported beside",
        "data": "synthetic (seeded Kuhn-tet mesh; FlowSys face-slot Jacobian ucfdgen_flowsys_fill; "
                "BSR block Jacobian ucfdgen_bsr_fill)",
        "config": {"workload": f"FlowSys block LU-SGS (NS 5x5 + SA 1x1), {args.m}^3 hex x 6 tets, "
                               f"{args.coloring} coloured, colour-ordered numbering",
                   "neles": N, "positions": npos, "ncolors": nc, "seed": args.seed},
        "flowsys": flows,
        "fp32_speedup_ns_sweeps": ns_ms64 / ns_ms32,
        "fp32_speedup_rans_sweeps": f64["rans_lower_upper_ms"] / f32["rans_lower_upper_ms"],
        "fp32_rel_err_vs_fp64": errs,
        "flowsys_scalar_lusgs": scalar,
        "krylov_precon_apply": applies,
        "krylov_precon_apply_fp32_rel_err": apply_err,
        "roofline": {
            "kernel": "FlowSys NS block sweep (k_sweep_w<5,false>), one launch per colour",
            "bound": "hbm", "achieved": lb / (launch_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": lb / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": lb, "avg_launch_ms": launch_ms},
    }
    return rec


def adjoint_workload(args, log):
    """Differentiation form: the FP64 coloured LU-SGS pass (pre + lower +
    upper) on a 55^3 x 6 tet mesh (998,250 cells) behind
    torch.autograd.Function with the hand-written adjoint; J = sum(w dQ),
    w ~ N(0,1).  Times forward + backward with dQ* stored and recomputed,
    reports the peak memory of each, and checks the gradients with respect to
    rhsb, dt and fspr against central differences (h = 1e-6) of the forward
    on 64 seeded directions per input."""
    import numpy as np
    import torch

    from ucfd_sparse_amd import mesh as M
    from ucfd_sparse_amd.autograd import ColoredSystem, lusgs_pass
    from ucfd_sparse_amd.layout import colored_layout

    m, seed = args.adjoint_m, 4
    t0 = time.time()
    mesh = M.tet_mesh(m, seed=seed)
    flow = M.flow_fields(mesh, seed=seed)
    mesh.key = M.hilbert_keys(mesh, "tet", (m,))
    lay = colored_layout(mesh, flow, coloring=args.coloring)
    del mesh, flow
    N = lay.mesh.neles
    log(f"adjoint mesh {N} cells, {lay.ncolors} colours, setup {time.time() - t0:.1f} s")
    S = ColoredSystem(lay)
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    f = lay.flow
    rhs0, dt0, fs0 = up(f.rhsb[:5]), up(f.dt), up(f.fspr)
    w = up(M.normal_fill(5 * N, 4000).reshape(5, N))
    runs = {}
    grads = {}
    for recompute in (False, True):
        def step(mem=None):
            r = rhs0.clone().requires_grad_()
            d = dt0.clone().requires_grad_()
            q = fs0.clone().requires_grad_()
            if mem is not None:
                torch.cuda.synchronize()
                mem["before"] = torch.cuda.memory_allocated()
            J = (w * lusgs_pass(S, r, d, q, recompute=recompute)).sum()
            if mem is not None:
                torch.cuda.synchronize()
                mem["saved"] = torch.cuda.memory_allocated() - mem["before"]
            J.backward()
            return r.grad, d.grad, q.grad
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        mem = {}
        torch.cuda.reset_peak_memory_stats()
        step(mem)
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated() - mem["before"]
        reps = args.adjoint_reps
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g = step()
        e1.record()
        torch.cuda.synchronize()
        runs["recompute" if recompute else "stored"] = {
            "ms_forward_backward": e0.elapsed_time(e1) / reps,
            "bytes_held_forward_to_backward": mem["saved"],
            "peak_bytes_forward_backward": peak}
        grads[recompute] = g
        log(f"recompute={recompute}: {e0.elapsed_time(e1) / reps:.3f} ms fwd+bwd, held after "
            f"forward {mem['saved'] / 2**20:.1f} MiB, peak {peak / 2**20:.1f} MiB")
    same = all(torch.equal(a, b) for a, b in zip(grads[False], grads[True]))
    # forward alone and adjoint alone (stored dQ*)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    dq, dstar, _ = S.forward_arrays(rhs0, dt0, fs0)
    e1.record()
    S.adjoint(dt0, fs0, dstar, dq, w)
    e2.record()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
    # central differences, h = 1e-6, on seeded directions per input (unit
    # scale for rhsb, relative x100 for dt and fspr):
    #  * Gaussian directions: the scale-free error |fd - ad| / (|g| |d|) and
    #    the raw |fd - ad| / |<g, d>|.  The raw error is the difference
    #    quotient's per-entry accuracy amplified by sum|g_i d_i| / |<g, d>|,
    #    about sqrt(N) for random-sign directions.
    #  * gradient-signed directions d_i = sign(g_i) |z_i|: <g, d> has no
    #    cancellation, so the raw error is meaningful; a wrong sign or
    #    magnitude in any entry of g still shows.
    h = 1e-6
    rng = np.random.default_rng(40)
    base_in = {"rhsb": rhs0, "dt": dt0, "fspr": fs0}
    g_in = dict(zip(("rhsb", "dt", "fspr"), grads[False]))

    def scaled(name, z):
        return z if name == "rhsb" else z * base_in[name] * 100.0

    def fd_of(name, d):
        xp = dict(base_in, **{name: base_in[name] + h * d})
        xm = dict(base_in, **{name: base_in[name] - h * d})
        dqp, _, _ = S.forward_arrays(xp["rhsb"], xp["dt"], xp["fspr"])
        dqm, _, _ = S.forward_arrays(xm["rhsb"], xm["dt"], xm["fspr"])
        return float((w * (dqp - dqm)).sum()) / (2 * h)
    fd_err = {}
    for name in ("rhsb", "dt", "fspr"):
        shape = tuple(base_in[name].shape)
        g = g_in[name]
        gnorm = float(g.norm())
        worst_n = worst_r = worst_s = 0.0
        for _ in range(args.adjoint_dirs):
            z = torch.from_numpy(rng.normal(size=shape)).to(dev)
            d = scaled(name, z)
            fd, ad = fd_of(name, d), float((g * d).sum())
            worst_n = max(worst_n, abs(fd - ad) / (gnorm * float(d.norm())))
            worst_r = max(worst_r, abs(fd - ad) / abs(ad))
            d = scaled(name, torch.sign(g) * z.abs())
            fd, ad = fd_of(name, d), float((g * d).sum())
            worst_s = max(worst_s, abs(fd - ad) / abs(ad))
        fd_err[name] = {"gaussian_scale_free": worst_n, "gaussian_raw": worst_r,
                        "gradient_signed_raw": worst_s}
        log(f"FD check {name}: gaussian |fd-ad|/(|g||d|) {worst_n:.2e} (raw {worst_r:.2e}), "
            f"gradient-signed raw {worst_s:.2e}, {args.adjoint_dirs} directions each")
    S.close()
    st = runs["stored"]["ms_forward_backward"]
    rec = {
        "metric": "differentiation form: forward + adjoint time of one coloured LU-SGS pass",
        "value": st,
        "unit": "ms per forward+backward",
        "n_gpus": 1,
        "steps": args.adjoint_reps,
        "warmup": 2,
        "ms_per_step": st,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kuhn-tet mesh and Euler fields, ucfd_sparse_amd.mesh)",
        "config": {"workload": f"FP64 coloured LU-SGS pass behind torch.autograd.Function, "
                               f"{m}^3 hex x 6 tets, J = sum(w dQ)",
                   "neles": N, "ncolors": lay.ncolors},
        "forward_ms": fwd_ms,
        "adjoint_ms": adj_ms,
        "modes": runs,
        "recompute_gradients_identical": same,
        "fd_check": {"h": h, "directions_per_input": args.adjoint_dirs,
                     "direction_scale": {"rhsb": "N(0,1)", "dt": "N(0,1) x 100 dt",
                                         "fspr": "N(0,1) x 100 fspr"},
                     "worst": fd_err,
                     "pass_1e-6": all(v["gaussian_scale_free"] <= 1e-6 and
                                      v["gradient_signed_raw"] <= 1e-6
                                      for v in fd_err.values())},
    }
    return rec


class LegacyPass:
    """One legacy-API pass on a colour-contiguous layout, device arrays,
    enqueued on its own stream."""

    def __init__(self, lay):
        import numpy as np
        import torch

        from ucfd_sparse_amd import lib as L
        self.L, self.lib = L, L.load()
        self.lay = lay
        dev = torch.device("cuda:0")

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        m, f = lay.mesh, lay.flow
        self.N, self.F, self.nc = m.neles, m.nface, lay.ncolors
        self.T = dict(nei=up(m.nei_ele), ic=up(lay.icolor), lc=up(lay.lcolor),
                      fnv=up(m.fnorm_vol), vfn=up(m.vec_fnorm), fsp=up(f.fspr), dt=up(f.dt),
                      upts=up(f.upts), rhs=up(f.rhsb))
        self.upts0 = up(f.upts)
        self.T["dstar"] = torch.zeros_like(self.T["rhs"])
        self.T["dq"] = torch.zeros_like(self.T["rhs"])
        self.T["diag"] = torch.zeros(self.N, dtype=torch.float64, device=dev)
        self.P = {k: v.data_ptr() for k, v in self.T.items()}
        self.off = [int(x) for x in lay.off]
        L.call("ucfd_set_config", 5, 5, 3)
        L.call("ucfd_declare_colored_layout", self.N, self.T["ic"], self.T["lc"], None)
        self.stream = torch.cuda.Stream(device=dev)
        torch.cuda.synchronize()

    def step(self, ev=None, cev=None):
        lib, P, N, F, off, nc = self.lib, self.P, self.N, self.F, self.off, self.nc
        if ev:
            ev[0].record(self.stream)
        lib.parallel_pre_lusgs(N, F, 1.0, P["fnv"], P["dt"], P["diag"], P["fsp"])
        if ev:
            ev[1].record(self.stream)
        for c in range(nc):
            lib.ns_parallel_lower_sweep(off[c], off[c + 1], N, F, P["nei"], P["ic"], P["lc"],
                                        P["fnv"], P["vfn"], P["upts"], P["rhs"], P["dstar"],
                                        P["diag"], P["fsp"])
            if cev:
                cev[0][c].record(self.stream)
        if ev:
            ev[2].record(self.stream)
        for c in reversed(range(nc)):
            lib.ns_parallel_upper_sweep(off[c], off[c + 1], N, F, P["nei"], P["ic"], P["lc"],
                                        P["fnv"], P["vfn"], P["upts"], P["dq"], P["dstar"],
                                        P["diag"], P["fsp"])
            if cev:
                cev[1][c].record(self.stream)
        if ev:
            ev[3].record(self.stream)
        lib.lusgs_parallel_ns_update(N, P["upts"], P["dq"])
        if ev:
            ev[4].record(self.stream)

    def run(self, fn):
        """fn() with the library on this pass's stream; waits and checks."""
        import torch
        self.L.call("ucfd_set_stream", self.stream.cuda_stream)
        try:
            r = fn()
            self.L.call("ucfd_synchronize")
        finally:
            self.L.call("ucfd_set_stream", None)
        torch.cuda.synchronize()
        return r

    def time(self, steps, warmup):
        """ms per step over `steps` timed steps (events on the pass stream)."""
        import torch

        def go():
            for _ in range(warmup):
                self.step()
            self.L.call("ucfd_synchronize")
            evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record(self.stream)
            for k in range(steps):
                self.step(evs[k])
            t1.record(self.stream)
            return evs, t0, t1
        evs, t0, t1 = self.run(go)
        total = t0.elapsed_time(t1)
        phases = [sum(e[j].elapsed_time(e[j + 1]) for e in evs) / steps for j in range(4)]
        return total / steps, dict(zip(("pre", "lower", "upper", "update"), phases))

    def launch_times(self, reps=3):
        """Per-colour launch times from separate instrumented steps."""
        import torch
        nc = self.nc
        cevs = [[[torch.cuda.Event(enable_timing=True) for _ in range(nc)] for _ in range(2)]
                for _ in range(reps)]
        pevs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(reps)]

        def go():
            for r in range(reps):
                self.step(pevs[r], cevs[r])
        self.run(go)
        lower = [sum((pevs[r][1] if c == 0 else cevs[r][0][c - 1]).elapsed_time(cevs[r][0][c])
                     for r in range(reps)) / reps for c in range(nc)]
        upper = [sum((pevs[r][2] if c == nc - 1 else cevs[r][1][c + 1]).elapsed_time(
            cevs[r][1][c]) for r in range(reps)) / reps for c in range(nc)]
        return lower, upper

    def fresh_pass(self):
        """One pass from the initial state into fresh outputs; host copies of
        (diag, dQ*, dQ, upts) for the parity check."""
        import torch
        T = self.T
        T["upts"].copy_(self.upts0)
        for k in ("dstar", "dq", "diag"):
            T[k].zero_()
        # the copies ran on torch's stream, the pass runs on its own
        torch.cuda.synchronize()
        self.run(self.step)
        return tuple(T[k].cpu().numpy() for k in ("diag", "dstar", "dq", "upts"))

    def close(self):
        self.L.call("ucfd_release_colored_layout", self.T["ic"], self.T["lc"])
        self.T.clear()


def layout_rate(mesh, flow, coloring, steps, warmup, log):
    """The same pass on another colouring of the same mesh."""
    import numpy as np
    import torch

    from ucfd_sparse_amd.layout import colored_layout
    t0 = time.time()
    lay = colored_layout(mesh, flow, coloring=coloring)
    lp = LegacyPass(lay)
    ms, phases = lp.time(steps, warmup)
    lower, upper = lp.launch_times()
    lp.close()
    del lp
    torch.cuda.empty_cache()
    N = lay.mesh.neles
    log(f"layout {coloring}: {lay.ncolors} colours, {ms:.3f} ms per pass "
        f"({N / ms / 1e6:.2f} G cell-iter/s), setup {time.time() - t0:.1f} s")
    return {"ncolors": lay.ncolors, "colour_sizes": [int(x) for x in np.diff(lay.off)],
            "ms_per_step": ms, "cell_iters_per_s": N / (ms * 1e-3),
            "hbm_gbs_algorithmic": 888 * N / (ms * 1e-3) / 1e9,
            "phase_ms": phases, "sweep_launch_ms": {"lower": lower, "upper": upper}}


def krylov_summary(r):
    k = r["kernels"]
    out = {"config": r["config"]["workload"], "bn": r["config"]["bn"]}
    for key in ("spmv_bsr5", "blusgs_apply_fp32", "blusgs_apply_fp64", "pbilu_apply_fp64"):
        out[key] = {"ms": k[key]["ms"], "gbs": k[key]["gbs"],
                    "frac": k[key]["gbs"] / HBM_PEAK_GBS}
    for key, v in r["solves"].items():
        out[key] = {"ms_mean": v["ms_mean"],
                    "iterations": [x["iterations"] for x in v["per_rhs"]],
                    "status": [x["status"] for x in v["per_rhs"]],
                    "max_true_rel_residual": max(x["true_rel_residual"] for x in v["per_rhs"])}
    return out


def blusgs_summary(r):
    f64, f32 = r["flowsys"]["fp64"], r["flowsys"]["fp32"]
    sc = r["flowsys_scalar_lusgs"]
    return {"config": r["config"]["workload"],
            "ns_cell_iters_per_s": {"fp64": f64["ns_cell_iters_per_s"],
                                    "fp32": f32["ns_cell_iters_per_s"]},
            "ns_sweep_gbs": {"fp64": f64["ns_sweep_gbs"], "fp32": f32["ns_sweep_gbs"]},
            "ns_sweep_frac_fp64": f64["ns_sweep_gbs"] / HBM_PEAK_GBS,
            "fp32_speedup_ns_sweeps": r["fp32_speedup_ns_sweeps"],
            "fp32_speedup_rans_sweeps": r["fp32_speedup_rans_sweeps"],
            "fp32_rel_err_vs_fp64": r["fp32_rel_err_vs_fp64"],
            "prepare_ms": {"pack": f64["pack_ms"], "ns": f64["ns_prepare_ms"],
                           "rans": f64["rans_prepare_ms"]},
            "scalar_lusgs": {"ns_lower_upper_ms": sc["ns_lower_upper_ms"],
                             "ns_sweep_gbs": sc["ns_sweep_gbs"],
                             "ns_sweep_frac": sc["ns_sweep_gbs"] / HBM_PEAK_GBS,
                             "rans_lower_upper_ms": sc["rans_lower_upper_ms"],
                             "pack_ms": sc["pack_ms"], "pack_gbs": sc["pack_gbs"],
                             "update_ms": sc["update_ms"],
                             "ns_prepare_ms": sc["ns_prepare_ms"]},
            "krylov_precon_apply": {k: {"ms": v["ms"], "gbs": v["gbs"],
                                        "frac": v["gbs"] / HBM_PEAK_GBS}
                                    for k, v in r["krylov_precon_apply"].items()},
            "roofline": r["roofline"]}


def adjoint_summary(r):
    mib = 2.0 ** 20
    return {"config": r["config"]["workload"], "forward_ms": r["forward_ms"],
            "adjoint_ms": r["adjoint_ms"],
            "stored": {"ms": r["modes"]["stored"]["ms_forward_backward"],
                       "peak_mib": r["modes"]["stored"]["peak_bytes_forward_backward"] / mib,
                       "held_mib": r["modes"]["stored"]["bytes_held_forward_to_backward"] / mib},
            "recompute": {"ms": r["modes"]["recompute"]["ms_forward_backward"],
                          "peak_mib": r["modes"]["recompute"]["peak_bytes_forward_backward"] / mib,
                          "held_mib": r["modes"]["recompute"]["bytes_held_forward_to_backward"]
                          / mib},
            "recompute_gradients_identical": r["recompute_gradients_identical"],
            "fd_directions_per_input": r["fd_check"]["directions_per_input"],
            "fd_worst": r["fd_check"]["worst"], "fd_pass_1e-6": r["fd_check"]["pass_1e-6"]}


def lusgs_headline(args, log):
    """The headline pass (timed), its roofline, the other colourings, the
    precision forms and the CPU baseline with the parity check."""
    import numpy as np
    import torch

    from ucfd_sparse_amd import mesh as M
    from ucfd_sparse_amd.layout import colored_layout

    t0 = time.time()
    mesh = M.tet_mesh(args.m, seed=args.seed)
    flow = M.flow_fields(mesh, seed=args.seed)
    mesh.key = M.hilbert_keys(mesh, "tet", (args.m,),
                              tet_type_major=args.cell_order == "tet-type")
    lay = colored_layout(mesh, flow, coloring=args.coloring)
    N, F = lay.mesh.neles, lay.mesh.nface
    nc = lay.ncolors
    log(f"mesh {N} cells, {nc} colours, setup {time.time() - t0:.1f} s")

    lp = LegacyPass(lay)
    ms_step, phases = lp.time(args.steps, args.warmup)
    lower_c, upper_c = lp.launch_times()
    log(f"per step {ms_step:.3f} ms: " + " ".join(f"{k} {v:.3f}" for k, v in phases.items()))
    log("launch ms per colour: lower " + " ".join(f"{x:.3f}" for x in lower_c)
        + " | upper " + " ".join(f"{x:.3f}" for x in upper_c))
    b_pre, b_sweep, b_upd = algorithmic_bytes(F)
    bytes_cell = b_pre + 2 * b_sweep + b_upd
    gbs = bytes_cell * N / (ms_step * 1e-3) / 1e9
    # dominant kernel: the longest lower-sweep launch (on the 2-colour layout
    # the last colour, every face taken; the first colour's launch takes no
    # face and reads 88 of the 344 algorithmic bytes per cell)
    kc = max(range(nc), key=lambda c: lower_c[c])
    off = lp.off
    heavy_cells = off[kc + 1] - off[kc]
    lower_launch_ms = lower_c[kc]
    lower_bytes_launch = b_sweep * heavy_cells
    kname = f"k_sweep<0,{3 if (kc == nc - 1 and F in (4, 6)) else 2},3,{F}>"
    achieved = lower_bytes_launch / (lower_launch_ms * 1e-3) / 1e9
    tpath = os.path.join("profiles", "traffic_lusgs_pass_tet96.txt")
    tr = read_traffic(os.path.join(REPO, tpath))
    traffic = pass_traffic = None
    traffic_source = "not measured for this configuration"
    if (tr and tr.get("m") == args.m and tr.get("ncolors") == nc and tr.get("seed") == args.seed
            and tr.get("colour") == kc and tr.get("coloring") == args.coloring
            and args.cell_order == "hex"):
        traffic = tr["hbm_bytes_per_launch"]
        pass_traffic = tr["pass_hbm_bytes"]
        traffic_source = (f"{tpath}: calibrated FETCH_SIZE + WRITE_SIZE of this configuration "
                          f"from separate rocprofv3 --pmc runs (not measured in this run)")
    log("parity: one pass from the initial state ...")
    gpu_pass = lp.fresh_pass()
    T = {k: lp.T[k] for k in ("nei", "fnv", "vfn", "fsp", "dt", "rhs")}
    T["upts"] = lp.upts0
    mixed = None
    if not args.no_mixed:
        log("precision forms (FP64 vs mixed) ...")
        mixed = precision_forms(lay, T, args.rtol, 3, log)
    lp.close()
    del lp, T
    torch.cuda.empty_cache()

    rates = {}
    if not args.quick:
        for col in ("greedy", "greedy-key-reduced"):
            rates[col] = layout_rate(mesh, flow, col, args.steps, args.warmup, log)
    del mesh, flow

    cpu = parity = None
    if not args.no_cpu:
        log("cpu baseline and parity check ...")
        cpu, parity = cpu_baseline(lay, args.cpu_seconds, gpu_pass)
        log(f"parity: max rel err {parity['max_rel_err']:.3e}, bitwise {parity['bitwise']}")

    rec = {
        "metric": METRIC,
        "value": N / (ms_step * 1e-3),
        "unit": "cell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kuhn-tet mesh and Euler fields, ucfd_sparse_amd.mesh)",
        "config": {
            "workload": f"colored LU-SGS FP64 legacy API pass (pre + lower + upper + update), "
                        f"{args.m}^3 hex x 6 Kuhn tets, permuted, {args.coloring} coloured, "
                        f"colour-contiguous Hilbert-ordered layout ({args.cell_order} order)",
            "coloring": args.coloring, "cell_order": args.cell_order,
            "colour_sizes": [int(x) for x in np.diff(lay.off)],
            "neles": N, "nface": F, "ncolors": nc, "seed": args.seed,
            "parallelism": "single GPU",
        },
        "hbm_gbs_algorithmic": gbs,
        "bytes_per_cell_iter": bytes_cell,
        "phase_ms": phases,
        "sweep_launch_ms": {"lower": lower_c, "upper": upper_c},
        "parity_max_rel_err": None if parity is None else parity["max_rel_err"],
        "parity": parity,
        "roofline": {
            "kernel": f"lower sweep of colour {kc} ({kname}, {heavy_cells} cells)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "hbm_gbs_from_traffic": (traffic / (lower_launch_ms * 1e-3) / 1e9
                                     if traffic else None),
            "algorithmic_bytes_per_launch": lower_bytes_launch,
            "avg_launch_ms": lower_launch_ms,
            "pass_algorithmic_gbs": gbs,
            "pass_frac": gbs / HBM_PEAK_GBS,
            "pass_traffic_bytes": pass_traffic,
            "pass_traffic_gbs": (pass_traffic / (ms_step * 1e-3) / 1e9 if pass_traffic
                                 else None),
            "pass_traffic_frac": (pass_traffic / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS
                                  if pass_traffic else None),
        },
        "other_colourings": rates,
        "cpu_baseline": cpu,
        "precision_forms": mixed,
    }
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--m", type=int, default=96, help="hex cells per axis (6 m^3 tets)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline time budget (split over the two builds)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the parity check")
    ap.add_argument("--no-mixed", action="store_true", help="skip the precision-form comparison")
    ap.add_argument("--quick", action="store_true",
                    help="headline pass only: no other colourings, no block LU-SGS / Krylov / "
                         "adjoint summaries")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--coloring", choices=("greedy", "greedy-key", "greedy-key-reduced",
                                           "greedy-bfs", "balanced"),
                    default="greedy-bfs",
                    help="colouring of the headline mesh: first-fit in cell id "
                         "(ucfd_greedy_coloring), first-fit along the Hilbert curve "
                         "(ucfd_greedy_coloring_ordered, -reduced: then Kempe-chain "
                         "reduction), first-fit in breadth-first order (ucfd_bfs_order) "
                         "or balanced")
    ap.add_argument("--cell-order", choices=("hex", "tet-type"), default="hex",
                    help="numbering inside each colour of the headline mesh: along the "
                         "Hilbert curve of the cubes, a cube's tets together (hex), or "
                         "tet index first, then the curve (tet-type)")
    ap.add_argument("--workload", choices=("lusgs", "krylov", "adjoint", "blusgs"),
                    default="lusgs",
                    help="lusgs: the headline line (default, with the other workloads' "
                         "summaries unless --quick); krylov / adjoint / blusgs: that "
                         "workload's own full line")
    ap.add_argument("--adjoint-m", type=int, default=55)
    ap.add_argument("--adjoint-reps", type=int, default=10)
    ap.add_argument("--adjoint-dirs", type=int, default=64)
    ap.add_argument("--krylov-m", type=int, default=70)
    ap.add_argument("--krylov-rhs", type=int, default=4)
    ap.add_argument("--calibrate-copy-mb", type=float, default=0.0,
                    help="only run a plain elementwise copy kernel (out = in + 0) over this "
                         "many MiB, 5 times: the reference reading for the PMC byte counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures the colored LU-SGS path on one GPU only; multi-GPU runs "
              "(domain decomposition with halo exchange) are not part of this build. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy_mb > 0:
        n = int(args.calibrate_copy_mb * 2**20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.add(a, 0.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel_bytes_read": 8 * n, "copy_kernel_bytes_written": 8 * n}))
        return

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    side = {"krylov": krylov_workload, "adjoint": adjoint_workload, "blusgs": blusgs_workload}
    if args.workload in side:
        print(json.dumps(side[args.workload](args, log)), flush=True)
        return

    t_start = time.time()
    rec = lusgs_headline(args, log)
    if not args.quick:
        summarize = {"blusgs": blusgs_summary, "krylov": krylov_summary,
                     "adjoint": adjoint_summary}
        for name in ("blusgs", "krylov", "adjoint"):
            t0 = time.time()
            log(f"{name} workload ...")
            rec[name] = summarize[name](side[name](args, log))
            torch.cuda.empty_cache()
            log(f"{name} done in {time.time() - t0:.1f} s")
    rec["bench_wall_s"] = time.time() - t_start
    print(json.dumps(rec), flush=True)
    if rec["parity_max_rel_err"] is not None and rec["parity_max_rel_err"] > 1e-12:
        print(f"parity check failed: max rel err {rec['parity_max_rel_err']:.3e} > 1e-12",
              file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
