#!/usr/bin/env python3
"""Benchmark of the FP64 tensor-quadrature Coulomb3D apply on one MI355X.

Workload (the reference's config 1 shape, which fits one GPU): density = sum
of 4 normalised Gaussians (exponents U[0.5,5] bohr^-2, centres U[-2,2]^3) on a
97^3 cube over [-8,8]^3 built from 7-point LIP cells (16 cells, h = 1/6);
40-point t-quadrature (20 Gauss-Legendre points on [0,2] + 20 logarithmic on
[2,500]); coda on.  One step = one apply of the operator to the resident cube
(Operator3D_transform_cube, src/bubbles/function3d.F90:4230-4349) with input
and output already in HBM.

Prints one JSON line: throughput in Gpts/s (output points per second), the
roofline entry of the dominant kernel (algorithmic FP64 FLOPs per launch over
its event-timed average duration, against the 78.6 TF FP64 MFMA peak), the
error against the CPU oracle and the analytic potential, and the CPU baseline
(the oracle's single-threaded FP64 apply on this host).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Coulomb/Helmholtz apply Gpts/s & %MFMA peak; SCF time-to-1e-6 Ha vs plain FP64"
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# FP64 flops per (point, bubble) of the inject at lmax 4, nlip 7: 49 Horner,
# 25 x 7 radial dot products, ~75 for the solid-harmonic recursion, 25 scalings
FLOP_PER_POINT_BUBBLE = 2 * 49 + 2 * 25 * 7 + 75 + 2 * 25


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-reps", type=int, default=2)
    p.add_argument("--no-inject", dest="inject", action="store_false",
                   help="skip the bubbles-inject measurement")
    p.add_argument("--no-gbfmm", dest="gbfmm", action="store_false",
                   help="skip the GBFMM measurement")
    p.add_argument("--copy-calibration", action="store_true",
                   help="time a plain device copy of a 257^3 cube instead (the known-size "
                        "kernel that scales the FETCH_SIZE/WRITE_SIZE counters)")
    return p.parse_args()


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/hbm_traffic.json: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0,
    factors from the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(args):
    import numpy as np

    import dage_amd
    n = 257
    ctx = dage_amd.Context(0)
    a = dage_amd.DeviceCube(ctx, (n, n, n), np.ones((n, n, n), order="F"))
    b = dage_amd.DeviceCube(ctx, (n, n, n))
    for _ in range(args.warmup):
        b.copy_from(a)
    ctx.time_begin()
    for _ in range(args.steps):
        b.copy_from(a)
    ms = ctx.time_end() / args.steps
    nbytes = 8 * n ** 3
    print(json.dumps({"mode": "copy_calibration", "kernel": "cube_copy_kernel",
                      "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                      "ms_per_launch": round(ms, 4),
                      "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
    for obj in (a, b, ctx):
        obj.close()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how the apply would shard).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args)
        return

    import numpy as np

    import dage_amd
    import oracle as O

    n_cell, nlip, T = 16, 7, 40
    ax = O.Axis.ncell_fixed(-8.0, 8.0, nlip, n_cell)
    N = ax.n
    rng = np.random.default_rng(0xDA6E0001)
    alpha = rng.uniform(0.5, 5.0, 4)
    centers = rng.uniform(-2.0, 2.0, (4, 3))
    q = np.ones(4)
    rho = O.gauss_density(q, alpha, centers, ax.points, ax.points, ax.points)

    # the library's own context stream carries the kernels and the timing events
    ctx = dage_amd.Context(0)
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, nlip)
    quad = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    op = dage_amd.Coulomb3D(ctx, grid, grid, quad)

    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)

    for _ in range(args.warmup):
        op.transform_cuda_cube(d_in, d_out)
    ctx.synchronize()

    ctx.profile(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.time_begin()
    for _ in range(args.steps):
        op.transform_cuda_cube(d_in, d_out)
    ms_total = ctx.time_end()
    wall = time.perf_counter() - t0
    ms_step = ms_total / args.steps

    # mixed precision (split-BF16, 6 products) on the same operator and cube
    d_mix = dage_amd.DeviceCube(ctx, rho.shape)
    mixed = {}
    for mode, label in ((dage_amd.DAGE_PREC_BF16X6, "bf16x6"), (dage_amd.DAGE_PREC_BF16X3, "bf16x3")):
        for _ in range(args.warmup):
            op.transform_cuda_cube(d_in, d_mix, precision=mode)
        ctx.time_begin()
        for _ in range(args.steps):
            op.transform_cuda_cube(d_in, d_mix, precision=mode)
        mms = ctx.time_end() / args.steps
        mixed[label] = (mms, d_mix.download())

    kernels = {}
    for name in ("gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode", "splitk_reduce_kernel",
                 "gemm_mixed_xmode", "gemm_mixed_ymode", "gemm_mixed_zmode"):
        tot, n = ctx.profile_query(name)
        kernels[name] = (tot / max(n, 1), n // args.steps)
    ctx.profile(False)

    V = d_out.download()
    pts = float(N) ** 3
    gpts = pts / (ms_step * 1e-3) / 1e9
    flop_mode = 2.0 * T * float(N) ** 4  # each mode product, per apply
    flops_apply = 3.0 * flop_mode

    # dominant kernel: the FP64 GEMM pass with the largest time per apply
    per_apply = {k: v[0] * v[1] for k, v in kernels.items() if k.startswith("gemm_f64")}
    dom = max(per_apply, key=per_apply.get)
    dom_ms, dom_launches = kernels[dom]
    dom_flops = flop_mode / max(dom_launches, 1)
    achieved_tf = dom_flops / (dom_ms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma",
        "kernel": dom,
        "achieved": round(achieved_tf, 3),
        "peak": FP64_MFMA_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / FP64_MFMA_PEAK_TFLOPS, 4),
        "traffic": measured_traffic(dom),
        "avg_launch_ms": round(dom_ms, 4),
        "algorithmic_flops_per_launch": dom_flops,
    }

    # parity against the oracle (same operator, CPU) and the analytic potential
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, quad.tp)
    w = O.coulomb_weights(quad.tw)
    cpu_times = []
    reps = 1 if args.no_cpu_baseline else max(1, args.cpu_reps)
    ref = None
    for _ in range(reps):
        c0 = time.perf_counter()
        ref = O.transform_cube(fx, fy, fz, w, O.coda(500.0), rho)
        cpu_times.append(time.perf_counter() - c0)
    err_oracle = float(np.max(np.abs(V - ref)) / np.max(np.abs(ref)))
    Va = O.gauss_potential(q, alpha, centers, ax.points, ax.points, ax.points)
    err_analytic = float(np.max(np.abs(V - Va)) / np.max(np.abs(Va)))
    cpu_s = min(cpu_times)
    cpu_baseline = None
    if not args.no_cpu_baseline:
        cpu_baseline = {
            "value": round(pts / cpu_s / 1e9, 6),
            "unit": "Gpts/s",
            "cores": 1,
            "kind": "port",
            "sample": f"full 97^3 T=40 apply with the oracle's C restatement of "
                      f"Operator3D_transform_cube (1 thread, best of {reps}: {cpu_s:.2f} s)",
        }

    line = {
        "metric": METRIC,
        "value": round(gpts, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "config 1: plain Coulomb3D apply, FP64, 97^3 cube over [-8,8]^3, "
                        "nlip 7, T=40 (20 lin [0,2] + 20 log [2,500]), 4 Gaussians, coda on",
            "n": N, "T": T, "nlip": nlip, "precision": "fp64",
        },
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "apply_tflops": round(flops_apply / (ms_step * 1e-3) / 1e12, 3),
        "apply_mfma_frac": round(flops_apply / (ms_step * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4),
        "kernels_ms_per_launch": {k: round(v[0], 4) for k, v in kernels.items() if v[1] > 0},
        "mixed": {label: {"ms_per_apply": round(mms, 4),
                          "Gpts_s": round(pts / (mms * 1e-3) / 1e9, 4),
                          "speedup_vs_fp64": round(ms_step / mms, 3),
                          "dense_fp64_equiv_tflops": round(flops_apply / (mms * 1e-3) / 1e12, 2),
                          "max_rel_err_vs_oracle": float(np.max(np.abs(vm - ref)) / np.max(np.abs(ref)))}
                  for label, (mms, vm) in mixed.items()},
        "max_rel_err_vs_oracle": err_oracle,
        "max_rel_err_vs_analytic": err_analytic,
        "wall_s_timed": round(wall, 4),
    }
    for obj in (d_in, d_out, d_mix, op, grid):
        obj.close()
    if args.inject:
        line["inject"] = bench_inject(ctx, args)
    if args.gbfmm:
        line["gbfmm"] = bench_gbfmm(ctx, args)
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_inject(ctx, args):
    """Bubbles inject at the config-2 shape: 241^3 cube over [-12,12]^3,
    4 atoms (Z in 1..8, centres in [-4,4]^3), 199-point radial grids to 20
    bohr, lmax 4, f_lm = c_lm r^l exp(-Z r).  One fused launch per step."""
    import numpy as np

    import dage_amd
    import oracle as O
    rng = np.random.default_rng(0xDA6E0002)
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, 40)
    n = ax.n
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    lmax, natom = 4, 4
    bub = dage_amd.Bubbles(ctx, natom, lmax, 0, 7)
    Z = rng.integers(1, 9, natom).astype(float)
    centers = rng.uniform(-4.0, 4.0, (natom, 3))
    host = []
    for a in range(natom):
        st = dage_amd.radial_grid_steps(Z[a], 33, 7, 20.0)
        r = O.Axis(0.0, st, 7).points
        c = rng.uniform(-1.0, 1.0, (lmax + 1) ** 2)
        bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                bf[:, l * l + l + m] = c[l * l + l + m] * r ** l * np.exp(-Z[a] * r)
        bub.set_bubble(a, centers[a], Z[a], st, bf)
        host.append((st, bf))
    cube = dage_amd.DeviceCube(ctx, (n, n, n))
    cube.set_to_zero()
    for _ in range(max(1, args.warmup)):
        bub.inject(grid, cube)
    steps = max(3, args.steps // 2)
    ctx.profile(True)
    ctx.time_begin()
    for _ in range(steps):
        bub.inject(grid, cube)
    ms = ctx.time_end() / steps
    kms, kn = ctx.profile_query("bubbles_inject_kernel")
    ctx.profile(False)
    kernel_ms = kms / max(kn, 1)
    # parity on a slab of the cube: z-planes 0..7 through the oracle
    cube.set_to_zero()
    bub.inject(grid, cube)
    got = cube.download()
    bset = O.BubbleSet(centers, [h[0] for h in host], [h[1] for h in host], lmax, 0, 7)
    ref = O.bubbles_eval_grid(bset, ax.points, ax.points, ax.points[:8])
    err = float(np.max(np.abs(got[:, :, :8] - ref)) / np.max(np.abs(ref)))
    nbytes = 16.0 * n ** 3  # cube read-modify-write per launch
    for obj in (cube, bub, grid):
        obj.close()
    return {
        "workload": "config 2 bubbles: 241^3 cube over [-12,12]^3, 4 atoms, lmax 4, "
                    "199-point radial grids, fused all-atom inject",
        "ms_per_launch": round(kernel_ms, 4),
        "Gpts_s": round(n ** 3 / (ms * 1e-3) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(nbytes / (kernel_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nbytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("bubbles_inject_kernel")},
        "fp64_valu_tflops": round(FLOP_PER_POINT_BUBBLE * natom * n ** 3 / (kernel_ms * 1e-3) / 1e12, 2),
        "max_rel_err_vs_oracle_slab": err,
    }


def bench_gbfmm(ctx, args):
    """GBFMM Coulomb potential at the config-2 shape: 241^3 cube over
    [-12,12]^3 (nlip 7, 40 cells), 4x4x4 boxes (maxlevel 2, 10 cells = 61
    points per box), lmax 12, near field with GaussQuad(25, 25) (T = 50);
    density = 16 Gaussians, exponents log-uniform in [0.3, 30], centres in
    [-4,4]^3.  Compared with the analytic potential and with the direct
    FP64 Coulomb apply of the same quadrature on the whole cube."""
    import numpy as np

    import dage_amd
    import oracle as O
    rng = np.random.default_rng(0xDA6E0002)
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, 40)
    n = ax.n
    q = np.ones(16)
    alpha = np.exp(rng.uniform(np.log(0.3), np.log(30.0), 16))
    centers = rng.uniform(-4.0, 4.0, (16, 3))
    rho = O.gauss_density(q, alpha, centers, ax.points, ax.points, ax.points)
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    gq = dage_amd.GaussQuad(25, 25)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, grid, maxlevel=2, lmax=12, gauss=gq)
    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)
    for _ in range(max(1, args.warmup // 2)):
        fmm.apply_device(d_in, d_out)
    steps = max(2, args.steps // 5)
    ctx.profile(True)
    ctx.time_begin()
    for _ in range(steps):
        fmm.apply_device(d_in, d_out)
    ms = ctx.time_end() / steps
    kernels = {}
    for k in ("gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode", "splitk_reduce_kernel",
              "box_multipoles_kernel", "m2m_kernel", "m2l_kernel", "l2l_kernel",
              "local_expansion_kernel"):
        kms, kn = ctx.profile_query(k)
        if kn:
            kernels[k] = {"ms_per_apply": round(kms / steps, 4), "launches_per_apply": kn // steps}
    ctx.profile(False)
    pot = d_out.download()
    exact = O.gauss_potential(q, alpha, centers, ax.points, ax.points, ax.points)
    err_exact = float(np.max(np.abs(pot - exact)) / np.max(np.abs(exact)))
    # direct apply with the same quadrature (no coda, as the near field)
    op = dage_amd.Coulomb3D(ctx, grid, grid, gq, coda=False)
    op.transform_cuda_cube(d_in, d_out)
    ctx.time_begin()
    for _ in range(2):
        op.transform_cuda_cube(d_in, d_out)
    direct_ms = ctx.time_end() / 2
    direct = d_out.download()
    err_direct = float(np.max(np.abs(pot - direct)) / np.max(np.abs(direct)))
    # the reference applies one sub-operator per box: sum over boxes of
    # 2 T (nxo nxi nyi nzi + nxo nyo nyi nzi + nxo nyo nzo nzi); here the near
    # fields of all boxes are one whole-cube apply with banded matrices,
    # 6 n^4 T dense FLOPs (the MFMA work actually issued)
    T = gq.get_npoints()
    s = 10 * 6
    nf = [min(4, b + 2) * s - max(0, b - 1) * s + 1 for b in range(4)]
    ref_flops = 0.0
    for bx in range(4):
        for by in range(4):
            for bz in range(4):
                no = s + 1
                ref_flops += 2 * T * (no * nf[bx] * nf[by] * nf[bz] + no * no * nf[by] * nf[bz]
                                      + no * no * no * nf[bz])
    flops = 6.0 * n ** 4 * T
    # split-BF16 near field on the same input
    fmm.apply_device(d_in, d_out, dage_amd.DAGE_PREC_BF16X6)
    ctx.time_begin()
    for _ in range(steps):
        fmm.apply_device(d_in, d_out, dage_amd.DAGE_PREC_BF16X6)
    mixed_ms = ctx.time_end() / steps
    err_mixed = float(np.max(np.abs(d_out.download() - pot)) / np.max(np.abs(pot)))
    gemm_ms = sum(kernels.get(k, {}).get("ms_per_apply", 0.0)
                  for k in ("gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode"))
    mp_ms = kernels.get("box_multipoles_kernel", {}).get("ms_per_apply", 0.0)
    le_ms = kernels.get("local_expansion_kernel", {}).get("ms_per_apply", 0.0)
    for obj in (op, fmm, d_in, d_out, grid):
        obj.close()
    return {
        "workload": "config 2 GBFMM: 241^3 cube over [-12,12]^3, 4x4x4 boxes, lmax 12, "
                    "near field T=50, 16 Gaussians",
        "ms_per_apply": round(ms, 3),
        "Gpts_s": round(n ** 3 / (ms * 1e-3) / 1e9, 4),
        "nearfield_dense_flops": flops,
        "nearfield_per_box_reference_flops": ref_flops,
        "nearfield_tflops": round(flops / (gemm_ms * 1e-3) / 1e12, 2) if gemm_ms else None,
        "nearfield_mfma_frac": round(flops / (gemm_ms * 1e-3) / 1e12 / 78.6, 4) if gemm_ms else None,
        "bf16x6_nearfield": {"ms_per_apply": round(mixed_ms, 3),
                             "speedup_vs_fp64": round(ms / mixed_ms, 3),
                             "max_rel_diff_vs_fp64_gbfmm": err_mixed},
        "multipoles_GBps": round(8.0 * 64 * 61 ** 3 / (mp_ms * 1e-3) / 1e9, 1) if mp_ms else None,
        "local_expansion_GBps": round(16.0 * n ** 3 / (le_ms * 1e-3) / 1e9, 1) if le_ms else None,
        "kernels": kernels,
        "direct_apply_ms": round(direct_ms, 3),
        "max_rel_err_vs_analytic": err_exact,
        "max_rel_diff_vs_direct_apply": err_direct,
    }


if __name__ == "__main__":
    main()
