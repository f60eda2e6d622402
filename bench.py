"""Benchmark of the FDFD hot path on one MI355X.

Workload (the mixed-precision benchmark configuration): 128³ Yee grid, 10-cell
SC-PML, background ε=1 with 40 seeded axis-aligned boxes (ε ~ U[1,12]),
subpixel-smoothed diagonal ε, μ=1, one seeded dipole of random orientation,
ω = 2π/24.  One step = one solve of (Cₘ μ⁻¹ Cₑ − ω² ε) e = −iω j from e = 0
to FP64 relative residual 1e-8 with complex-FP32 inner SQMR iterations and
FP64 residual/update refinement, b and e resident in HBM.

Also reported: the plain FP64 solve of the same system, the curl-curl apply
bandwidth, and the oracle's CPU FP64 path timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maxwellfdm.jl_amd"))

HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic_c2.json"))
    ap.add_argument("--apply-only", action="store_true",
                    help="only the stencil applies and a 256 MiB device copy (for hardware-counter runs)")
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
    return dict(value=t_it * fp64_iters, unit="s", cores=nthreads, kind="port",
                sample=f"oracle C FP64 on the same {N[0]}x{N[1]}x{N[2]} system: CSC assembly "
                       f"({t_asm:.1f} s, 1 thread) + {it} SQMR iterations (CSR SpMV, {nthreads} threads, "
                       f"{t_it * 1e3:.1f} ms/iteration) scaled to the {fp64_iters} iterations of the FP64 GPU "
                       f"solve", assembly_s=t_asm, s_per_iteration=t_it)


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
    dev = torch.device("cuda:0")
    if a.apply_only:
        # counter calibration: a plain copy of known size (256 MiB read + 256 MiB written)
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        ms32 = op.bench_apply(1, "mixed", iters=20)
        ms64 = op.bench_apply(1, "fp64", iters=20)
        print(json.dumps({"apply_only": True, "fp32_P_ms": ms32, "fp64_A_ms": ms64, "tiles": tiles,
                          "copy_bytes": 2 * src.numel() * 4, "cells": m.grid.M}))
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
    # true FP64 residual of the last mixed solve, through the device operator
    e_h = e_d.cpu().numpy()
    r = b - op.apply(e_h)
    relres = float(np.linalg.norm(r) / np.linalg.norm(b))

    # plain FP64 solve of the same system
    step("fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_d = step("fp64")
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0

    # dominant kernel: the complex-FP32 preconditioned curl-curl stencil of the
    # inner iterations; algorithmic bytes = (x + y + q) 3 complex FP32 each = 72 B/cell
    M_cells = m.grid.M
    ms32 = op.bench_apply(1, "mixed", iters=50)
    gbps32 = 72.0 * M_cells / (ms32 * 1e-3) / 1e9
    ms64 = op.bench_apply(1, "fp64", iters=50)
    gbps64 = 144.0 * M_cells / (ms64 * 1e-3) / 1e9
    traffic = None
    if os.path.exists(a.traffic_json):
        try:
            traffic = json.load(open(a.traffic_json)).get("stencil_fp32_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not a.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(m, om, int(st_d.iterations), nth, a.cpu_iters)

    out = {
        "metric": "solve time to rel. residual 1e-8 (mixed FP32/FP64 SQMR) + curl-curl apply GB/s",
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 inner / complex128 residual+update",
        "data": "synthetic",
        "config": {"workload": f"{N}^3 grid, {npml}-cell SC-PML, {nbox} seeded boxes eps~U[1,12] "
                               f"(seed {seed}), random dipole, omega=2pi/{lam:g}, mixed-precision SQMR to 1e-8",
                   "grid": [N, N, N], "unknowns": 3 * M_cells, "tol": 1e-8},
        "mixed": {"s": round(t_mixed, 5), "iterations": int(st_m.iterations), "cycles": int(st_m.outer_cycles),
                  "relres_fp64": relres},
        "fp64": {"s": round(t_fp64, 5), "iterations": int(st_d.iterations), "relres": st_d.relres},
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        "apply": {"fp32_P_ms": round(ms32, 4), "fp32_P_GBps": round(gbps32, 1), "fp64_A_ms": round(ms64, 4),
                  "fp64_A_GBps": round(gbps64, 1), "tiles_fp64_fp32": [list(tiles[0]), list(tiles[1])]},
        "roofline": {"bound": "hbm", "kernel": "k_apply<float, MODE_P> (FP32 stencil, 72 B/cell)",
                     "achieved": round(gbps32, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps32 / HBM_PEAK_GBPS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    op.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
