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
KERNEL_NAMES = {  # device symbols as rocprofv3 lists them (prefix match)
    "stencil": "k_apply_tiled<{R}, 1, 1",
    "upd_r": "k_upd_r<{R}>",
    "upd_dzq": "k_upd_dzq<{R}>",
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
    return ap.parse_args()


def cpu_baseline(m, om, fp64_iters, nthreads, iters):
    """Oracle CPU FP64 path on a bounded sample: assembly of the CSC A plus
    `iters` SQMR iterations with the CSC product; the per-iteration time is
    scaled by the iteration count of the FP64 device solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import types
    import oracle as O
    g = m.grid
    spec = types.SimpleNamespace(lprim=g.lprim, isbloch=g.isbloch, boundft=(0, 0, 0), npml=m.Npml, omega_pml=om,
                                 kbloch=(0, 0, 0), pml_m=4.0, pml_lnR=-16.0, order_cmpfirst=True)
    shapes = []
    et, mt = [], []
    for o in m.objs:
        s = o.shape
        shapes.append(("ball", tuple(s.c), s.r) if hasattr(s, "r") else ("cuboid", tuple(s.c), tuple(s.s)))
        et.append(o.mat.eps)
        mt.append(o.mat.mu)
    t0 = time.perf_counter()
    ea, ma = O.matparams(spec, shapes, np.array(et), np.array(mt))
    A = O.create_A(spec, om, ea, ma)
    t_asm = time.perf_counter() - t0
    N = g.N
    j = np.zeros(N + (3,), dtype=np.complex128, order="F")
    for src in getattr(m, "_bench_srcs", []):
        O.add_pointsrc(spec, 0, src[0], src[1], 1.0, j)
    je = O.field_arr2vec(spec, j)
    b = O.create_b(spec, om, ma, je, np.zeros_like(je))
    t0 = time.perf_counter()
    _, it, rr = O.solve_sqmr(spec, A, b, tol=1e-8, maxit=iters, nthreads=nthreads)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    return dict(value=round(t_it * fp64_iters, 3), unit="s", cores=nthreads, kind="port",
                sample=f"oracle C FP64 on the same {N[0]}x{N[1]}x{N[2]} system: CSC assembly "
                       f"({t_asm:.1f} s, 1 thread) + {it} SQMR iterations (CSR SpMV, {nthreads} threads, "
                       f"{t_it * 1e3:.1f} ms/iteration) scaled to the {fp64_iters} iterations of the FP64 GPU "
                       f"solve", assembly_s=round(t_asm, 3), s_per_iteration=t_it)


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


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
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
        nth = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(m, om, int(st_d.iterations), nth, a.cpu_iters)

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
