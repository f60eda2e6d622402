"""Benchmark of the FDFD hot path on one MI355X.

Workload (the mixed-precision configuration): 128³ Yee grid, 10-cell SC-PML,
background ε=1 with 40 seeded axis-aligned boxes (ε ~ U[1,12]),
subpixel-smoothed diagonal ε, μ=1, one seeded dipole of random orientation,
ω = 2π/24.  One step = one solve of (Cₘ μ⁻¹ Cₑ − ω² ε) e = −iω j from e = 0 to
FP64 relative residual 1e-8 in mixed precision (complex-FP32 inner SQMR cycles
with FP64 residual/update refinement; cycles that stop paying off continue in
FP64), with b and e resident in HBM.

Also reported: the plain FP64 solve of the same system, the curl-curl apply
bandwidth, per-kernel bandwidth of one solver iteration against the 8 TB/s
HBM roofline, and the oracle's CPU FP64 path timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maxwellfdm.jl_amd"))

HBM_PEAK_GBPS = 8000.0
METRIC = "solve time to rel. residual 1e-8 + curl-curl apply GB/s; mixed vs FP64 on GPU"
KERNELS = ("stencil", "upd_r", "upd_dzq")
KERNEL_NAMES = {  # device kernels as rocprofv3 lists them, template arguments abridged
    "stencil": "k_apply_tiled<{R}, 1, 2, ...> (Krylov stencil, qᵀPq epilogue)",
    "upd_r": "k_upd_r_dot<{R}, ...> (r update, α and ‖r‖², rᵀr)",
    "upd_dzq": "k_upd_dzq<{R}, ...> (d, z, q update)",
}


def bytes_per_cell(prec):
    """Algorithmic HBM bytes per grid cell of each kernel of one SQMR iteration
    (3 complex unknowns per cell, each stored operand counted once)."""
    sc = 16 if prec == "fp64" else 8           # complex vector element
    return {
        "stencil": 3 * sc * 2 + 3 * 16,          # read q, write t, read the FP64 complex Jacobi factors
        "upd_r": 3 * sc * 3,                     # read r, t; write r
        "upd_dzq": 3 * sc * 7,                   # read d, z, q, r; write d, z, q
    }


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic_c2.json"))
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the per-kernel timing launches and a 256 MiB device copy (for hardware-counter runs)")
    ap.add_argument("--tiles", default=None,
                    help="fixed stencil tiles 'TY,KZ,order/TY,KZ,order' (FP64/FP32) instead of the autotuner, "
                         "so that separate hardware-counter passes launch the same kernels")
    ap.add_argument("--workload", default="random-boxes",
                    choices=["random-boxes", "batched-rhs", "uneven-batch", "adjoint", "large-aniso", "next-rows",
                             "mixed-fgmres"],
                    help="random-boxes: the metric line; the others time the batched-RHS, uneven-problem "
                         "and adjoint configurations")
    ap.add_argument("--nrhs", type=int, default=4, help="batched-rhs: right-hand sides")
    ap.add_argument("--next-n", type=int, default=64,
                    help="next-rows: edge of the random-box cube for the EE, HH and boundft = HH solves")
    ap.add_argument("--count", type=int, default=32, help="uneven-batch: problems")
    ap.add_argument("--voxels", type=int, default=16, help="adjoint: finite-difference voxels")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="batched-rhs, uneven-batch: solver precision (mixed: complex-FP32 inner cycles)")
    ap.add_argument("--maxit", type=int, default=2000, help="large-aniso, batched-rhs --bounded: solver steps")
    ap.add_argument("--method", default=None, choices=["sqmr", "cocg", "bicgstab"],
                    help="large-aniso: Krylov method (default: SQMR for a complex-symmetric operator)")
    ap.add_argument("--inner-steps", default="1000,4000",
                    help="mixed-fgmres: FP32 SQMR steps per preconditioner application (comma-separated)")
    ap.add_argument("--restart", type=int, default=30, help="mixed-fgmres: FGMRES restart length")
    ap.add_argument("--bounded", action="store_true",
                    help="batched-rhs: stop at --maxit steps, report the residual history and the projection")
    return ap.parse_args()


def host_cores():
    """Cores this process may use: the OpenMP share the host sets for it
    (OMP_NUM_THREADS), else the CPUs it is allowed to run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env), "OMP_NUM_THREADS"
    return len(os.sched_getaffinity(0)), "sched_getaffinity"


def oracle_spec(m, om):
    import types
    g = m.grid
    return types.SimpleNamespace(lprim=g.lprim, isbloch=g.isbloch, boundft=(0, 0, 0), npml=m.Npml, omega_pml=om,
                                 kbloch=(0, 0, 0), pml_m=4.0, pml_lnR=-16.0, order_cmpfirst=True)


def oracle_shapes(m):
    shapes, et, mt = [], [], []
    for o in m.objs:
        s = o.shape
        if hasattr(s, "h"):
            shapes.append(("cylinder", tuple(s.c), s.r, s.h))
        elif hasattr(s, "r"):
            shapes.append(("ball", tuple(s.c), s.r))
        else:
            shapes.append(("cuboid", tuple(s.c), tuple(s.s)))
        et.append(o.mat.eps)
        mt.append(o.mat.mu)
    return shapes, et, mt


def cpu_baseline(m, om, fp64_iters, iters):
    """CPU baselines on the GPU box's host cores, on bounded samples:
    - the oracle's C FP64 path on the same system: CSC assembly (1 thread),
      then `iters` SQMR iterations with an OpenMP CSR product, on every core
      this process may use and on 1 thread; `value` is the all-core time per
      iteration times the FP64 GPU solve's iteration count (an extrapolation,
      labelled so);
    - the reference's own `A \\ b`, stood in by scipy's SuperLU (one thread),
      on the plain-method sphere generator at 20³ (24,000 unknowns): at the
      32³ size a factorisation did not finish within 15 minutes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import scipy.sparse.linalg as sla
    import oracle as O
    from maxwellfdfd import workloads as W
    spec = oracle_spec(m, om)
    shapes, et, mt = oracle_shapes(m)
    t0 = time.perf_counter()
    ea, ma = O.matparams(spec, shapes, np.array(et), np.array(mt))
    A = O.create_A(spec, om, ea, ma)
    t_asm = time.perf_counter() - t0
    N = m.grid.N
    j = np.zeros(N + (3,), dtype=np.complex128, order="F")
    for src in getattr(m, "_bench_srcs", []):
        O.add_pointsrc(spec, 0, src[0], src[1], 1.0, j)
    je = O.field_arr2vec(spec, j)
    b = O.create_b(spec, om, ma, je, np.zeros_like(je))
    ncores, how = host_cores()
    per_it = {}
    for nth in (ncores, 1):
        t0 = time.perf_counter()
        _, it, _ = O.solve_sqmr(spec, A, b, tol=1e-8, maxit=iters, nthreads=nth)
        per_it[nth] = (time.perf_counter() - t0) / max(it, 1)
    # the direct solve the reference's users run, at a size it finishes
    md, omd = W.sphere_dipole(N=20, npml=5, r=3.75, lam=12.5)
    sd = oracle_spec(md, omd)
    shd, etd, mtd = oracle_shapes(md)
    ead, mad = O.matparams(sd, shd, np.array(etd), np.array(mtd))
    Ad = O.to_scipy(O.create_A(sd, omd, ead, mad)).tocsc()
    jd = np.zeros((20, 20, 20, 3), dtype=np.complex128, order="F")
    O.add_pointsrc(sd, 0, (10.0, 10.0, 10.0), (0.0, 0.0, 1.0), 1.0, jd)
    bd = O.create_b(sd, omd, mad, O.field_arr2vec(sd, jd), np.zeros(Ad.shape[0], complex))
    t0 = time.perf_counter()
    xd = sla.splu(Ad).solve(bd)
    t_lu = time.perf_counter() - t0
    res_lu = float(np.linalg.norm(bd - Ad @ xd) / np.linalg.norm(bd))
    return dict(value=round(per_it[ncores] * fp64_iters, 3), unit="s", cores=ncores, kind="port",
                sample=f"oracle C FP64 on the same {N[0]}x{N[1]}x{N[2]} system: CSC assembly ({t_asm:.1f} s, "
                       f"1 thread) + {iters} SQMR iterations (CSR SpMV) on {ncores} threads ({how}): "
                       f"{per_it[ncores] * 1e3:.1f} ms/iteration; value = that time x the {fp64_iters} "
                       f"iterations of the FP64 GPU solve (extrapolated, not run to 1e-8)",
                assembly_s=round(t_asm, 3), s_per_iteration=per_it[ncores],
                one_thread={"s_per_iteration": per_it[1], "solve_s_extrapolated": round(per_it[1] * fp64_iters, 3)},
                direct_solve={"what": "scipy SuperLU factor + solve (1 thread), A \\ b stand-in, plain-method sphere "
                                      "generator at 20^3 (24,000 unknowns; the 32^3 factorisation did not finish "
                                      "in 15 min)",
                              "s": round(t_lu, 3), "relres": res_lu})


def kernel_table(op, cells, iters=50):
    """Per-kernel device time and algorithmic bandwidth of one solver iteration."""
    out = {}
    for prec in ("fp64", "mixed"):
        key = "fp64" if prec == "fp64" else "fp32"
        ms = op.bench_iteration(1, prec, iters)
        bpc = bytes_per_cell(key)
        out[key] = {k: {"ms": round(ms[i], 5), "bytes": bpc[k] * cells,
                        "GBps": round(bpc[k] * cells / (ms[i] * 1e-3) / 1e9, 1)}
                    for i, k in enumerate(KERNELS)}
        out[key]["iteration_ms"] = round(sum(ms), 5)
    return out


def copy_rates(sizes_mib=(64, 256, 1024, 4096), iters=20):
    """Device-to-device copy rate (GB/s, bytes read + written) at several sizes:
    the practical streaming ceiling next to the 8 TB/s peak.  Working sets up
    to the 256 MiB Infinity Cache run above the HBM rate."""
    import torch
    dev = torch.device("cuda:0")
    out = {}
    for mb in sizes_mib:
        src = torch.ones(mb * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        out[f"{mb}MiB"] = round(2 * mb * 1048576 / (ms * 1e-3) / 1e9, 1)
        del src, dst
    torch.cuda.empty_cache()
    return out


def run_batched_rhs(a):
    """N³ random boxes (the 40-box generator) with nrhs seeded dipoles solved in
    one batched launch (FP64 or mixed), against one single-RHS solve of the
    first dipole in the same precision; with --bounded, the batched solve stops
    at --maxit steps and reports its residual history, the per-RHS step cost
    against a single RHS (solver-iteration kernels in context) and the
    projected time to 1e-8."""
    import numpy as np
    import torch
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W
    N = a.N
    m, om = W.random_boxes(N=N, npml=10, nbox=40, lam=24.0, seed=1, ndip=0)
    M.calc_matparams(m)
    B = W.dipole_rhs(m, om, W.random_dipoles(N, 10, a.nrhs, 2024))
    op = M.MaxwellOperator(m, om)
    op.autotune(a.nrhs, 5)
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(B).to(dev)
    e_d = torch.zeros_like(b_d)
    torch.cuda.synchronize()
    if a.bounded:
        it_batch = op.bench_iteration(a.nrhs, a.precision, 5)
        op.autotune(1, 5)
        it_one = op.bench_iteration(1, a.precision, 10)
        op.autotune(a.nrhs, 5)
        print(f"[batched-rhs] {N}^3, {a.nrhs} right-hand sides: step {sum(it_batch):.3f} ms "
              f"({sum(it_batch) / a.nrhs:.4f} per RHS) vs {sum(it_one):.4f} ms for one", file=sys.stderr, flush=True)

        def solve():
            t0 = time.perf_counter()
            st = op.solve_dev(b_d.data_ptr(), e_d.data_ptr(), a.nrhs, tol=1e-8, maxit=a.maxit,
                              precision=a.precision, raise_on_noconv=False)
            torch.cuda.synchronize()
            return st, time.perf_counter() - t0

        (st, t_solve), lines = traced(solve)
        op.close()
        hist = residual_history(lines)
        hist.append((int(st.iterations), float(st.relres), "end"))
        ms_step = 1e3 * t_solve / max(st.iterations, 1)
        return {"workload": f"batched-rhs {N}^3 random boxes, {a.nrhs} dipoles, {a.precision} SQMR bounded at "
                            f"{a.maxit} steps",
                "iteration_kernels_ms_batch": [round(x, 4) for x in it_batch],
                "iteration_ms_per_rhs_batch": round(sum(it_batch) / a.nrhs, 4),
                "iteration_ms_single_rhs": round(sum(it_one), 4),
                "solve_s": round(t_solve, 3), "steps": int(st.iterations), "steps_fp32": int(st.iterations_fp32),
                "relres_reached": st.relres, "converged": bool(st.converged), "ms_per_step_all_rhs": round(ms_step, 3),
                "residual_history": [[int(k), float(r), w] for k, r, w in hist],
                "projection": project(hist, st.converged, ms_step)}
    print(f"[batched-rhs] {N}^3, {a.nrhs} right-hand sides: batched solve", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    st = op.solve_dev(b_d.data_ptr(), e_d.data_ptr(), a.nrhs, tol=1e-8, precision=a.precision)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # the same right-hand sides one at a time (every one up to 8, else the first),
    # with the tile shape tuned for one RHS
    nsingle = a.nrhs if a.nrhs <= 8 else 1
    print(f"[batched-rhs] batched solve {t_batch:.2f} s; {nsingle} single-RHS solve(s)", file=sys.stderr, flush=True)
    t1 = op.autotune(1, 5)
    # the single-RHS stencil at this size: FP64 A apply (144 B/cell) and the
    # Krylov step's kernels, with the tile tuned for one RHS
    ms_apply = op.bench_apply(1, "fp64", iters=50)
    it_one = op.bench_iteration(1, "fp64", 20)
    cells = m.grid.M
    singles, errs = [], []
    for r in range(nsingle):
        e1 = torch.zeros_like(b_d[r])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st1 = op.solve_dev(b_d[r].data_ptr(), e1.data_ptr(), 1, tol=1e-8, precision=a.precision)
        torch.cuda.synchronize()
        singles.append((time.perf_counter() - t0, int(st1.iterations)))
        errs.append(float(torch.linalg.vector_norm(e_d[r] - e1) / torch.linalg.vector_norm(e1)))
        print(f"[batched-rhs] single RHS {r}: {singles[-1][0]:.2f} s, {singles[-1][1]} steps", file=sys.stderr,
              flush=True)
    op.close()
    out = {"workload": f"batched-rhs {N}^3 random boxes, {a.nrhs} dipoles, {a.precision} SQMR to 1e-8",
           "batched_s": round(t_batch, 3), "iterations_max": int(st.iterations),
           "iterations_fp32": int(st.iterations_fp32), "relres_max": st.relres,
           "single_rhs_s": [round(t, 3) for t, _ in singles], "single_rhs_iterations": [k for _, k in singles],
           "s_per_rhs_batched": round(t_batch / a.nrhs, 3), "rel_diff_vs_single": errs,
           "single_rhs_tiles": [list(t1[0]), list(t1[1])], "fp64_A_apply_ms": round(ms_apply, 4),
           "fp64_A_apply_GBps": round(144.0 * cells / (ms_apply * 1e-3) / 1e9, 1),
           "iteration_kernels_ms_1rhs_fp64": [round(x, 4) for x in it_one]}
    if nsingle == a.nrhs:
        out["singles_total_s"] = round(sum(t for t, _ in singles), 3)
        out["speedup_batched_vs_singles"] = round(out["singles_total_s"] / t_batch, 3)
    return out


def run_uneven_batch(a):
    """`count` independent problems (N ~ U{48..128}, own boxes, own ω) in one
    lockstep graph-launched batched solve (FP64 or mixed), against the same
    problems solved one after another in the same precision."""
    import numpy as np
    import torch
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W
    probs = W.uneven_problems(count=a.count, nmin=48, nmax=128, npml=10, nbox=40, seed=7)
    log = lambda *x: print("[uneven-batch]", *x, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    ops, bds, eds, cells = [], [], [], 0
    for m, om in probs:
        M.calc_matparams(m)
        je, jm = M.create_srcs(m)
        b = M.create_b(M.EE, om, m, je, jm)
        op = M.MaxwellOperator(m, om)
        op.autotune(1, 5)
        ops.append(op)
        bds.append(torch.from_numpy(b).to(dev))
        eds.append(torch.zeros_like(bds[-1]))
        cells += m.grid.M
        log("set up problem", len(ops), "N =", m.grid.N[0])
    torch.cuda.synchronize()
    log("batched solve of", len(ops), "problems,", cells, "cells")
    os.environ["MFD_SOLVE_TRACE"] = "1"   # the solver reports each problem as it converges
    t0 = time.perf_counter()
    st = ops[0].solve_batch_dev(ops[1:], [x.data_ptr() for x in bds], [x.data_ptr() for x in eds], tol=1e-8,
                                precision=a.precision)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    del os.environ["MFD_SOLVE_TRACE"]
    log("batched solve done in", round(t_batch, 2), "s; sequential solves")
    t0 = time.perf_counter()
    its = []
    for op, bd, ed in zip(ops, bds, eds):
        ed.zero_()
        its.append(int(op.solve_dev(bd.data_ptr(), ed.data_ptr(), 1, tol=1e-8, precision=a.precision).iterations))
        log("sequential problem", len(its), "iterations", its[-1])
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for op in ops:
        op.close()
    return {"workload": f"uneven-batch: {a.count} problems, N ~ U{{48..128}}, omega ~ U[2pi/30, 2pi/16], "
                        f"{a.precision} SQMR to 1e-8",
            "sizes": [int(m.grid.N[0]) for m, _ in probs], "cells_total": int(cells),
            "batched_s": round(t_batch, 3), "sequential_s": round(t_seq, 3),
            "speedup_batched_vs_sequential": round(t_seq / t_batch, 3),
            "iterations": [int(s.iterations) for s in st], "converged": all(bool(s.converged) for s in st),
            "iterations_fp32": [int(s.iterations_fp32) for s in st],
            "iterations_sequential": its}


def traced(fn):
    """Runs fn() with the solver's trace on (MFD_SOLVE_TRACE: one line per
    inner cycle and per true-residual probe, written by the library to fd 2)
    and returns (fn's result, the trace lines).  fd 2 goes through a pipe
    whose reader thread echoes every line as it arrives, so a long solve
    keeps printing its progress."""
    import threading
    os.environ["MFD_SOLVE_TRACE"] = "1"
    sys.stderr.flush()
    saved = os.dup(2)
    rfd, wfd = os.pipe()
    lines = []

    def pump():
        with os.fdopen(rfd, "rb") as r:
            for raw in r:
                os.write(saved, raw)
                lines.append(raw.decode("utf-8", "replace").rstrip("\n"))

    th = threading.Thread(target=pump, daemon=True)
    th.start()
    os.dup2(wfd, 2)
    os.close(wfd)
    try:
        out = fn()
    finally:
        sys.stderr.flush()
        os.dup2(saved, 2)   # closes the pipe's last write end: the reader ends
        th.join(timeout=10)
        os.close(saved)
        del os.environ["MFD_SOLVE_TRACE"]
    return out, lines


def residual_history(lines):
    """(Krylov steps, FP64 relative residual) pairs from a solver trace: the
    residual at each cycle start and at each true-residual probe."""
    import re
    hist, steps = [], 0
    for line in lines:
        m = re.search(r"\[mfd solve\] cycle (\d+) prec (\w+) residual ([0-9.e+-]+) target [0-9.e+-]+ iterations (\d+)", line)
        if m:
            hist.append((steps, float(m.group(3)), m.group(2)))
            steps += int(m.group(4))
            continue
        m = re.search(r"\[mfd probe\] it (\d+) true residual ([0-9.e+-]+)", line)
        if m:
            hist.append((steps + int(m.group(1)), float(m.group(2)), "probe"))
    return hist


def project(hist, converged, ms_step):
    """Steps and seconds to relative residual 1e-8, extrapolated from the last
    two points of a residual history (None if it converged or made no progress)."""
    import math
    pts = [(k, r) for k, r, _ in hist if r > 0]
    if converged or len(pts) < 2:
        return None
    (k0, r0), (k1, r1) = pts[-2], pts[-1]
    if not (r1 < r0 and k1 > k0):
        return None
    spd = (k1 - k0) / math.log10(r0 / r1)
    steps_needed = k1 + spd * math.log10(r1 / 1e-8)
    return {"steps_per_decade": round(spd, 1), "steps_to_1e-8": int(steps_needed),
            "s_to_1e-8": round(steps_needed * ms_step * 1e-3, 1)}


def run_large_aniso(a):
    """384×384×256 anisotropic case (200 smoothed spheres, full 3×3 ε, 8
    dipoles): assembly on the device, apply bandwidth of the full-tensor
    stencil, and a bounded batched solve (a.maxit steps, a.precision) with its
    FP64 residual history, the time per step and the time to 1e-8 projected
    from the history's last decade of progress."""
    import numpy as np
    import torch
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W
    log = lambda *x: print("[large-aniso]", *x, file=sys.stderr, flush=True)
    m, om, dips = W.random_spheres(ndip=a.nrhs)
    t0 = time.perf_counter()
    M.calc_matparams(m)
    t_mat = time.perf_counter() - t0
    full = bool(np.any(m.εarr[..., 0, 1] != 0))
    log("matparams", round(t_mat, 2), "s; off-diagonal ε present:", full)
    B = W.dipole_rhs(m, om, dips)
    op = M.MaxwellOperator(m, om)
    cells = m.grid.M
    t0 = time.perf_counter()
    tiles = op.autotune(a.nrhs, 3)
    log("tiles", tiles, "autotune", round(time.perf_counter() - t0, 1), "s")
    ms_apply = op.bench_apply(1, "fp64", iters=10)
    bpc = 6 * 16 + 6 * 16                   # x + y, and the six stored ε entries (3 diagonal + 3 corner)
    log("apply", round(ms_apply, 3), "ms")
    it_ms = op.bench_iteration(a.nrhs, a.precision, 5)
    log("iteration kernels", it_ms)
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(B).to(dev)
    e_d = torch.zeros_like(b_d)
    torch.cuda.synchronize()

    def solve():
        t0 = time.perf_counter()
        st = op.solve_dev(b_d.data_ptr(), e_d.data_ptr(), a.nrhs, tol=1e-8, maxit=a.maxit, precision=a.precision,
                          method=a.method, raise_on_noconv=False)
        torch.cuda.synchronize()
        return st, time.perf_counter() - t0

    (st, t_solve), lines = traced(solve)
    op.close()
    hist = residual_history(lines)
    hist.append((int(st.iterations), float(st.relres), "end"))
    ms_step = 1e3 * t_solve / max(st.iterations, 1)
    proj = project(hist, st.converged, ms_step)
    return {"workload": f"large-aniso {m.grid.N[0]}x{m.grid.N[1]}x{m.grid.N[2]}, 200 smoothed spheres "
                        f"(full eps tensors: {full}), {a.nrhs} dipoles, {a.precision} {(a.method or 'sqmr').upper()} "
                        f"bounded at {a.maxit} steps",
            "cells": int(cells), "matparams_s": round(t_mat, 3), "tiles": [list(tiles[0]), list(tiles[1])],
            "apply_ms": round(ms_apply, 4),
            "apply_GBps": round(bpc * cells / (ms_apply * 1e-3) / 1e9, 1),
            "iteration_kernels_ms": [round(x, 4) for x in it_ms],
            "solve_s": round(t_solve, 3), "steps": int(st.iterations), "steps_fp32": int(st.iterations_fp32),
            "relres_reached": st.relres, "converged": bool(st.converged), "ms_per_step_all_rhs": round(ms_step, 3),
            "residual_history": [[int(k), float(r), w] for k, r, w in hist], "projection": proj}


def run_adjoint(a):
    """128×128×64 adjoint case: forward solve, adjoint solve and ∂f/∂ε, with
    central finite differences (h = 1e-4, perturbation solves) on seeded design voxels."""
    import numpy as np
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W
    m, om, b, mon, sl = W.adjoint_design()
    op = M.MaxwellOperator(m, om)
    op.autotune(1, 5)
    t0 = time.perf_counter()
    e, pullback = op.solve_with_pullback(b, tol=1e-11)
    t_fwd = time.perf_counter() - t0
    g = np.zeros_like(e)
    g[mon] = np.conj(e[mon])
    t0 = time.perf_counter()
    grad = pullback(g)
    t_adj = time.perf_counter() - t0
    op.close()
    N = m.grid.N
    rng = np.random.default_rng(1)
    h = 1e-4
    errs = []
    print(f"[adjoint] forward {t_fwd:.2f} s, adjoint + gradient {t_adj:.2f} s; finite differences",
          file=sys.stderr, flush=True)
    # Each perturbed field is e + δ with A(ε ± h) δ = b − A(ε ± h) e, solved to
    # 1e-12 relative to its own right-hand side: the solver's error then scales
    # with δ = O(h), not with e, so it stays far below the O(h²) truncation of
    # the central difference (a warm-started solve of A(ε ± h) x = b to a
    # tolerance relative to b leaves an error of order tol·|e|/h in the quotient).
    t0 = time.perf_counter()
    for v in range(a.voxels):
        c = int(rng.integers(3))
        i, j, k = (int(rng.integers(s.start, s.stop)) for s in sl)
        fs = []
        for sgn in (1, -1):
            ea = m.εarr.copy()
            ea[i, j, k, c, c] += sgn * h
            o2 = M.MaxwellOperator(m, om, εarr=ea)
            d, _ = o2.solve(b - o2.apply(e), tol=1e-12)
            o2.close()
            x = e + d
            fs.append(np.sum(np.abs(x[mon]) ** 2))
        fd = (fs[0] - fs[1]) / (2 * h)
        dof = c + 3 * (i + N[0] * (j + N[1] * k))
        errs.append(abs(grad[dof] - fd) / max(abs(fd), 1e-300))
        print(f"[adjoint] voxel {v} {(i, j, k, c)}: fd {fd:.6e} adjoint {grad[dof]:.6e} rel {errs[-1]:.2e}",
              file=sys.stderr, flush=True)
    t_fd = time.perf_counter() - t0
    return {"workload": "adjoint: 128x128x64, 10-cell PML, 64x64x16 design region eps = 1 + 11 sigmoid(G*xi), "
                        "z-dipole, |E_z|^2 over a 4x4x1 monitor",
            "forward_s": round(t_fwd, 3), "adjoint_and_grad_s": round(t_adj, 3),
            "fd": "central, h = 1e-4, perturbed fields e + A(eps +- h)^-1 (b - A(eps +- h) e) to 1e-12",
            "fd_voxels": a.voxels, "fd_s": round(t_fd, 1), "fd_rel_err_max": max(errs), "fd_rel_err": errs}


def run_mixed_fgmres(a):
    """A mixed-precision alternative to the automatic FP32→FP64 hand-over,
    measured on the N³ random-box system: FP64 flexible GMRES(m) whose
    preconditioner is k steps of the library's FP32-storage SQMR from a zero
    start (the mixed solve with maxit = k), against the plain FP64 SQMR solve.
    The outer method runs in FP64 on the device (torch vectors, classical
    Gram-Schmidt applied twice, the small least-squares problem on the host);
    it stops at FP64 relative residual 1e-8 or after about 150,000 inner steps."""
    import numpy as np
    import torch
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W
    log = lambda *x: print("[mixed-fgmres]", *x, file=sys.stderr, flush=True)
    N, mres = a.N, a.restart
    m, om = W.random_boxes(N=N, npml=10, nbox=40, lam=24.0, seed=1, ndip=1)
    m._bench_srcs = [(pos, v) for pos, v in W.random_dipoles(N, 10, 1, 1001)]
    M.calc_matparams(m)
    je, jm = M.create_srcs(m)
    b = M.create_b(M.EE, om, m, je, jm)
    op = M.MaxwellOperator(m, om)
    op.autotune(1, 10)
    dev = torch.device("cuda:0")
    bd = torch.from_numpy(b).to(dev)
    torch.cuda.synchronize()
    bn = torch.linalg.vector_norm(bd).item()

    def A(x):
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        op.apply_dev(x.data_ptr(), y.data_ptr(), 1, "fp64")
        return y

    count = {"inner": 0, "fp32": 0}

    def precondition(v, k):
        z = torch.zeros_like(v)
        torch.cuda.synchronize()
        st = op.solve_dev(v.data_ptr(), z.data_ptr(), 1, tol=1e-30, precision="mixed", maxit=k, inner_tol=1e-30,
                          raise_on_noconv=False)
        count["inner"] += int(st.iterations)
        count["fp32"] += int(st.iterations_fp32)
        return z

    e = torch.zeros_like(bd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = op.solve_dev(bd.data_ptr(), e.data_ptr(), 1, tol=1e-8)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    log(f"FP64 SQMR {t64:.2f} s, {st.iterations} steps")

    def fgmres(k, max_outer, tol=1e-8):
        x = torch.zeros_like(bd)
        outer, hist = 0, []
        while True:
            r = bd - A(x)
            beta = torch.linalg.vector_norm(r).item()
            hist.append([outer, count["inner"], beta / bn])
            log(f"k {k} outer {outer} inner steps {count['inner']} FP64 residual {beta / bn:.3e}")
            if beta / bn <= tol or outer >= max_outer:
                return x, outer, hist
            V, Z = [r / beta], []
            H = np.zeros((mres + 1, mres), complex)
            y = None
            for j in range(mres):
                Z.append(precondition(V[j], k))
                w = A(Z[j])
                for _ in range(2):
                    h = torch.stack([torch.vdot(V[i], w) for i in range(j + 1)]).cpu().numpy()
                    H[:j + 1, j] += h
                    for i in range(j + 1):
                        w = w - complex(h[i]) * V[i]
                hn = torch.linalg.vector_norm(w).item()
                H[j + 1, j] = hn
                outer += 1
                rhs = np.zeros(j + 2, complex)
                rhs[0] = beta
                y = np.linalg.lstsq(H[:j + 2, :j + 1], rhs, rcond=None)[0]
                est = np.linalg.norm(H[:j + 2, :j + 1] @ y - rhs) / bn
                if est <= 0.5 * tol or hn == 0 or outer >= max_outer:
                    break
                V.append(w / hn)
            for i in range(len(y)):
                x = x + complex(y[i]) * Z[i]

    runs = []
    for k in (int(v) for v in a.inner_steps.split(",")):
        count["inner"] = count["fp32"] = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, outer, hist = fgmres(k, max(2, 150000 // k))
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        runs.append({"inner_steps_per_application": k, "s": round(tf, 3), "applications": outer,
                     "inner_steps": count["inner"], "inner_steps_fp32": count["fp32"], "relres": hist[-1][2],
                     "converged": hist[-1][2] <= 1e-8, "time_vs_fp64_sqmr": round(tf / t64, 3),
                     "history": hist})
    op.close()
    return {"workload": f"mixed-fgmres {N}^3 random boxes: FP64 FGMRES({mres}) preconditioned by k FP32-storage "
                        f"SQMR steps, against FP64 SQMR, to FP64 rel. residual 1e-8",
            "fp64_sqmr": {"s": round(t64, 3), "steps": int(st.iterations)}, "fgmres": runs}


def run_next_rows(a):
    """The rows beyond the plain EE path, on the N³ random-box model (N =
    --next-n, 64 by default; FP64 to 1e-8): the EE solve (tiled stencil) next to the HH formulation and a
    boundft = HH z axis (the tiled stencil on the reflected grid), h_from_e /
    e_from_h, create_Mcs, and a 1024² 2D TE solve (the single-z-cell
    operator restricted to Ex, Ey), with the default and the autotuned tile."""
    import numpy as np
    import maxwellfdfd as M
    from maxwellfdfd import twod as T
    from maxwellfdfd import workloads as W
    log = lambda *x: print("[next-rows]", *x, file=sys.stderr, flush=True)
    N = a.next_n
    out = {"workload": f"next rows: {N}^3 random boxes (EE, HH formulation, boundft HH on z), field conversion, "
                       "create_Mcs, 1024^2 2D TE; FP64 SQMR to 1e-8"}

    def solve(op, b):
        t0 = time.perf_counter()
        x, st = op.solve(b, tol=1e-8)
        return x, {"s": round(time.perf_counter() - t0, 3), "device_s": round(st.seconds, 3),
                   "iterations": int(st.iterations), "relres": st.relres}

    m, om = W.random_boxes(N=N, npml=10, nbox=40, lam=24.0, seed=1, ndip=1)
    M.calc_matparams(m)
    je, jm = M.create_srcs(m)
    op = M.MaxwellOperator(m, om)
    e, out["ee_solve"] = solve(op, M.create_b(M.EE, om, m, je, jm))
    out["ee_apply_ms"] = round(op.bench_apply(1, "fp64", iters=20), 4)
    log("EE", out["ee_solve"])
    t0 = time.perf_counter()
    h_ee = op.h_from_e(e)
    out["h_from_e_s"] = round(time.perf_counter() - t0, 4)
    t0 = time.perf_counter()
    e_back = op.e_from_h(h_ee, je)   # e = (−i/ω) ε⁻¹ (Cₘ h − jₑ) returns the solved e
    out["e_from_h_s"] = round(time.perf_counter() - t0, 4)
    out["e_from_h_of_h_from_e_rel"] = float(np.linalg.norm(e_back - e) / np.linalg.norm(e))
    op.close()
    oph = M.MaxwellOperator(m, om, ft=M.HH)
    h, out["hh_solve"] = solve(oph, M.create_b(M.HH, om, m, je, jm))
    out["hh_apply_ms"] = round(oph.bench_apply(1, "fp64", iters=20), 4)
    out["hh_vs_h_from_e_rel"] = float(np.linalg.norm(h - h_ee) / np.linalg.norm(h_ee))
    oph.close()
    log("HH formulation", out["hh_solve"])
    m2, _ = W.random_boxes(N=N, npml=10, nbox=40, lam=24.0, seed=1, ndip=0)
    M.set_boundft(m2, (M.EE, M.EE, M.HH))   # before the deposit: it fixes the E grid types
    for pos, v in W.random_dipoles(N, 10, 1, 1001):
        M.add_srce(m2, M.PointSrc(pos, v, 1.0))
    M.calc_matparams(m2)
    je2, jm2 = M.create_srcs(m2)
    op2 = M.MaxwellOperator(m2, om)
    _, out["boundft_hh_z_solve"] = solve(op2, M.create_b(M.EE, om, m2, je2, jm2))
    out["boundft_hh_z_apply_ms"] = round(op2.bench_apply(1, "fp64", iters=20), 4)
    op2.close()
    log("boundft HH on z", out["boundft_hh_z_solve"])
    m128, _ = W.random_boxes(N=128, npml=10, nbox=40, lam=24.0, seed=1, ndip=0)
    M.calc_matparams(m128)
    t0 = time.perf_counter()
    Mce, Mcm = M.create_Mcs(m128)
    out["create_Mcs_128_s"] = round(time.perf_counter() - t0, 3)
    out["create_Mcs_128_nnz"] = [int(Mce.nnz), int(Mcm.nnz)]
    del Mce, Mcm
    log("create_Mcs 128^3", out["create_Mcs_128_s"], "s")
    n2 = 1024
    lp = [np.arange(n2 + 1, dtype=np.float64)] * 2
    m2d = T.ModelTE(T.Grid2(lp, (False, False)))
    T.set_ωpml(m2d, om)
    T.set_Npml(m2d, ((10, 10), (10, 10)))
    T.add_obj(m2d, "bg", T.Rect((n2 / 2, n2 / 2), (float(n2), float(n2))), ε=1.0)
    T.add_obj(m2d, "rod", T.Disk((n2 / 2 + 40.0, n2 / 2 - 25.0), 120.0), ε=9.0)
    T.add_srce(m2d, T.PointSrc2((n2 / 2 - 200.5, n2 / 2 + 0.5), (0.0, 1.0)))
    T.calc_matparams(m2d)
    je2d, jm2d = T.create_srcs(m2d)
    op2d = T.MaxwellOperator2D(m2d, om)
    b2d = T.create_b(M.EE, om, m2d, je2d, jm2d)
    _, out["te2d_1024_solve"] = solve(op2d, b2d)
    log("2D TE 1024^2", out["te2d_1024_solve"])
    t64, _, _ = op2d.op.autotune(1, 10)
    out["te2d_1024_tile_autotuned"] = list(t64)
    _, out["te2d_1024_solve_autotuned"] = solve(op2d, b2d)
    op2d.close()
    log("2D TE 1024^2 autotuned", t64, out["te2d_1024_solve_autotuned"])
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload != "random-boxes":
        if a.workload == "batched-rhs":
            if a.N == 128:
                a.N = 160
            out = run_batched_rhs(a)
        elif a.workload == "uneven-batch":
            out = run_uneven_batch(a)
        elif a.workload == "large-aniso":
            out = run_large_aniso(a)
        elif a.workload == "mixed-fgmres":
            out = run_mixed_fgmres(a)
        elif a.workload == "next-rows":
            out = run_next_rows(a)
        else:
            out = run_adjoint(a)
        print(json.dumps(out))
        return
    import numpy as np
    import torch
    import maxwellfdfd as M
    from maxwellfdfd import workloads as W

    N, npml, nbox, lam, seed = a.N, 10, 40, 24.0, 1
    m, om = W.random_boxes(N=N, npml=npml, nbox=nbox, lam=lam, seed=seed, ndip=1)
    m._bench_srcs = [(pos, v) for pos, v in W.random_dipoles(N, npml, 1, seed + 1000)]
    M.calc_matparams(m)
    je, jm = M.create_srcs(m)
    b = M.create_b(M.EE, om, m, je, jm)
    op = M.MaxwellOperator(m, om)
    if a.tiles:
        t64, t32 = ([int(v) for v in t.split(",")] for t in a.tiles.split("/"))
        op.set_tile("fp64", *t64)
        op.set_tile("mixed", *t32)
        tiles = (tuple(t64), tuple(t32), None)
    else:
        tiles = op.autotune(1, 10)
    cells = m.grid.M
    dev = torch.device("cuda:0")
    if a.kernels_only:
        # counter calibration: a plain copy of known size (256 MiB read + 256 MiB written)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        kt = kernel_table(op, cells, iters=10)
        print(json.dumps({"kernels_only": True, "kernels": kt, "tiles": tiles,
                          "copy_bytes": 2 * src.numel() * 4, "cells": cells}))
        op.close()
        return
    b_d = torch.from_numpy(b).to(dev)
    e_d = torch.zeros_like(b_d)
    torch.cuda.synchronize()

    def step(precision):
        e_d.zero_()
        torch.cuda.synchronize()
        return op.solve_dev(b_d.data_ptr(), e_d.data_ptr(), 1, tol=1e-8, precision=precision)

    for _ in range(a.warmup):
        step("mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(a.steps):
        stats.append(step("mixed"))
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / a.steps
    st_m = stats[-1]
    e_mixed = e_d.clone()
    # true FP64 residual of the last mixed solve, through the device operator
    r = b - op.apply(e_mixed.cpu().numpy())
    relres = float(np.linalg.norm(r) / np.linalg.norm(b))

    # plain FP64 solve of the same system (one untimed, one timed)
    step("fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_d = step("fp64")
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0
    err_vs_fp64 = float(torch.linalg.vector_norm(e_mixed - e_d) / torch.linalg.vector_norm(e_d))
    del e_mixed

    # curl-curl apply bandwidth (FP64 A, 144 B/cell) and the iteration's kernels
    ms_apply = op.bench_apply(1, "fp64", iters=50)
    copies = copy_rates()
    gbps_apply = 144.0 * cells / (ms_apply * 1e-3) / 1e9
    kt = kernel_table(op, cells)
    # dominant kernel: the largest share of the iteration in the precision
    # that ran most of the mixed solve's iterations
    dom_prec = "fp32" if 2 * st_m.iterations_fp32 > st_m.iterations else "fp64"
    dom = max(KERNELS, key=lambda k: kt[dom_prec][k]["ms"])
    kd = kt[dom_prec][dom]
    traffic = None
    if os.path.exists(a.traffic_json):
        try:
            traffic = json.load(open(a.traffic_json)).get(f"{dom_prec}_{dom}")
        except (OSError, ValueError):
            traffic = None

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(m, om, int(st_d.iterations), a.cpu_iters)

    rdev = "float" if dom_prec == "fp32" else "double"
    out = {
        "metric": METRIC,
        "value": round(t_mixed, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 arithmetic; c64 inner-cycle storage",
        "data": "synthetic",
        "config": {"workload": f"{N}^3 grid, {npml}-cell SC-PML, {nbox} seeded boxes eps~U[1,12] (seed {seed}), "
                               f"random dipole, omega=2pi/{lam:g}; mixed-precision SQMR solve to FP64 rel. "
                               f"residual 1e-8",
                   "grid": [N, N, N], "unknowns": 3 * cells, "tol": 1e-8},
        "mixed": {"s": round(t_mixed, 4), "iterations": int(st_m.iterations),
                  "iterations_fp32": int(st_m.iterations_fp32), "cycles": int(st_m.outer_cycles),
                  "converged": bool(st_m.converged), "relres_fp64": relres, "rel_err_vs_fp64_solution": err_vs_fp64},
        "fp64": {"s": round(t_fp64, 4), "iterations": int(st_d.iterations), "relres": st_d.relres,
                 "converged": bool(st_d.converged)},
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        "apply": {"fp64_A_ms": round(ms_apply, 5), "fp64_A_GBps": round(gbps_apply, 1),
                  "tiles_fp64_fp32": [list(tiles[0]), list(tiles[1])]},
        "device_copy_GBps": copies,
        "kernels": kt,
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom].format(R=rdev),
                     "achieved": kd["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(kd["GBps"] / HBM_PEAK_GBPS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    op.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
