#!/usr/bin/env python3
"""Benchmarks of the operator-application hot path on one MI355X.

Headline (config 2 shapes): one step = the density of 16 Gaussians
(exponents log-uniform in [0.3, 30], centres in [-4,4]^3) plus 4 atoms'
bubbles (Z in 1..8, 199-point radial grids, lmax 4, c_lm r^l exp(-Z r))
injected onto a 241^3 cube over [-12,12]^3 (nlip 7, 40 cells), then its
Coulomb potential by the grid-based FMM (4x4x4 boxes, lmax 12, near field
with GaussQuad(25, 25), T = 50), all in FP64 with the cubes resident in HBM:
  Bubbles_eval_3dgrid / bubbles_inject_cuda  (bubbles.F90:2148-2211)
  GBFMMCoulomb3D_calculate_potential         (gbfmm_coulomb3d.F90:650-821)
`value` = 241^3 output points per second of wall time per step.

Also reported: config 1 (plain Coulomb3D apply, 97^3, T = 40, FP64 and
split-BF16, CPU oracle timed beside it), config 3 (32 Helmholtz
right-hand sides, k^2 in [0.1, 4], cubes of 65/97/129/161 points at h 0.15,
one graph launch per batch, autotuned tiles, against 32 separate FP64
applies) and config 4 (Helmholtz SCF of a 10-atom molecule on a 129^3 cube,
12 orbitals: iterations and wall time to 1e-6 Ha, plain vs the learned
preconditioner trained beforehand on another molecule).

Prints one JSON line.  The roofline entry is the dominant kernel of the
headline step: algorithmic FLOPs per launch over its event-timed average
duration, against the 78.6 TF FP64 MFMA peak; `traffic` is the HBM bytes per
launch from the committed rocprofv3 counter summary.
"""
import argparse
import ctypes
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
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-config1", dest="config1", action="store_false",
                   help="skip the config-1 apply measurement")
    p.add_argument("--no-config3", dest="config3", action="store_false",
                   help="skip the config-3 batched Helmholtz measurement")
    p.add_argument("--no-config4", dest="config4", action="store_false",
                   help="skip the config-4 SCF (plain vs learned) measurement")
    p.add_argument("--no-config2", dest="config2", action="store_false",
                   help="skip the config-2 GBFMM (241^3) measurement")
    p.add_argument("--config2-cpu", action="store_true",
                   help="also time the oracle's sample of a config-2 step")
    p.add_argument("--scf-max-iter", type=int, default=1500,
                   help="iteration cap of the timed plain and learned SCF runs")
    p.add_argument("--scf-train-iter", type=int, default=600,
                   help="plain iterations on the training molecule (targets: its last orbitals)")
    p.add_argument("--scf-record-iter", type=int, default=200,
                   help="training pairs come from every 4th of the first this-many iterations")
    p.add_argument("--headline-only", action="store_true",
                   help="run only the headline's warmup and timed steps (for profiler runs: "
                        "no parity, split-BF16, CPU-baseline or other-config legs)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="time a plain device copy of a 257^3 cube instead (the known-size "
                        "kernel that scales the FETCH_SIZE/WRITE_SIZE counters)")
    return p.parse_args()


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` in `workload` from the committed
    rocprofv3 counter summary (profiles/hbm_traffic.json: FETCH_SIZE x 2.0 +
    WRITE_SIZE x 1.0, factors from the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["workloads"][workload]["kernels"][kernel]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def measured_executed_fraction(workload: str, kernel: str):
    """Fraction of the dense-equivalent FLOPs that `kernel` executes (the
    rest are skipped all-zero panels), from the committed MFMA counter
    summary (profiles/hbm_traffic.json: SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 per
    dispatch over the dense count), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["workloads"][workload]["kernels"][kernel]["executed_fraction"]
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


def progress(msg):
    """One line on stderr per finished leg (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# ----------------------------------------------------------------- config 2
def config2_inputs(ctx):
    import numpy as np

    import dage_amd
    import oracle as O
    rng = np.random.default_rng(0xDA6E0002)
    ax = O.Axis.ncell_fixed(-12.0, 12.0, 7, 40)
    q = np.ones(16)
    alpha = np.exp(rng.uniform(np.log(0.3), np.log(30.0), 16))
    centers = rng.uniform(-4.0, 4.0, (16, 3))
    rho = O.gauss_density(q, alpha, centers, ax.points, ax.points, ax.points)
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    lmax, natom = 4, 4
    bub = dage_amd.Bubbles(ctx, natom, lmax, 0, 7)
    Z = rng.integers(1, 9, natom).astype(float)
    acent = rng.uniform(-4.0, 4.0, (natom, 3))
    host = []
    for a in range(natom):
        st = dage_amd.radial_grid_steps(Z[a], 33, 7, 20.0)
        r = O.Axis(0.0, st, 7).points
        c = rng.uniform(-1.0, 1.0, (lmax + 1) ** 2)
        bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                bf[:, l * l + l + m] = c[l * l + l + m] * r ** l * np.exp(-Z[a] * r)
        bub.set_bubble(a, acent[a], Z[a], st, bf)
        host.append((st, bf))
    bset = O.BubbleSet(acent, [h[0] for h in host], [h[1] for h in host], lmax, 0, 7)
    return ax, (q, alpha, centers), rho, grid, bub, bset


def bench_config2(ctx, args):
    import numpy as np

    import dage_amd
    import oracle as O
    ax, gauss, rho, grid, bub, bset = config2_inputs(ctx)
    n = ax.n
    gq = dage_amd.GaussQuad(25, 25)
    fmm = dage_amd.GBFMMCoulomb3D(ctx, grid, maxlevel=2, lmax=12, gauss=gq)
    d_rho = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_tot = dage_amd.DeviceCube(ctx, rho.shape)
    d_pot = dage_amd.DeviceCube(ctx, rho.shape)

    def step(precision=dage_amd.DAGE_PREC_FP64):
        d_tot.copy_from(d_rho)
        bub.inject(grid, d_tot)
        fmm.apply_device(d_tot, d_pot, precision)

    steps = max(args.steps, 10)  # a 17-ms step: time at least 10
    for _ in range(max(args.warmup, 2)):
        step()
    ctx.synchronize()
    ctx.profile(True)
    wall0 = time.perf_counter()
    ctx.time_begin()
    for _ in range(steps):
        step()
    ms = ctx.time_end() / steps
    wall = time.perf_counter() - wall0
    names = ("xy_fused_kernel", "gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode",
             "band_xmode", "band_ymode", "band_xymode", "band_zmode", "lowrank",
             "splitk_reduce_kernel",
             "bubbles_inject_kernel", "box_multipoles_kernel", "box_multipoles_gemm",
             "m2m_kernel", "m2l_gemm", "l2l_kernel", "local_expansion_kernel",
             "local_expansion_gemm")
    kernels = {}
    for k in names:
        kms, kn = ctx.profile_query(k)
        if kn:
            kernels[k] = {"ms_per_launch": round(kms / kn, 4), "launches_per_step": kn // steps,
                          "ms_per_step": round(kms / steps, 4)}
    ctx.profile(False)
    T = gq.get_npoints()
    # dominant kernel: largest time per step; a GEMM pass does 2 T n^4
    # dense-equivalent FLOPs per apply, the fused x/y pass twice that (the
    # zero K panels of the banded near-field matrices are skipped but counted)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dom]
    if dom.startswith("gemm_f64") or dom == "xy_fused_kernel":
        passes = 2.0 if dom == "xy_fused_kernel" else 1.0
        # the MFMA passes cover the t-points between the low-rank head and the
        # banded tail
        t_mfma = fmm.band_t0() - fmm.lowrank_t1() if dom.startswith("gemm_f64") else T
        per_launch = passes * 2.0 * t_mfma * float(n) ** 4 / dk["launches_per_step"]
        achieved = per_launch / (dk["ms_per_launch"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3),
                    "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4),
                    "traffic": measured_traffic("config2", dom),
                    "avg_launch_ms": dk["ms_per_launch"], "algorithmic_flops_per_launch": per_launch}
    else:
        per_launch = 16.0 * n ** 3
        achieved = per_launch / (dk["ms_per_launch"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic("config2", dom),
                    "avg_launch_ms": dk["ms_per_launch"], "algorithmic_bytes_per_launch": per_launch}
    # split-BF16 near field in the same step
    step(dage_amd.DAGE_PREC_BF16X6)
    ctx.time_begin()
    for _ in range(steps):
        step(dage_amd.DAGE_PREC_BF16X6)
    ms_mixed = ctx.time_end() / steps
    pot_mixed = d_pot.download()
    step()
    pot = d_pot.download()
    # parity: GBFMM of the Gaussians alone vs the analytic potential and vs
    # the direct apply; the inject on 8 z-planes vs the oracle
    fmm.apply_device(d_rho, d_pot)
    pot_g = d_pot.download()
    q, alpha, centers = gauss
    exact = O.gauss_potential(q, alpha, centers, ax.points, ax.points, ax.points)
    op = dage_amd.Coulomb3D(ctx, grid, grid, gq, coda=False)
    op.transform_cuda_cube(d_rho, d_pot)
    ctx.time_begin()
    op.transform_cuda_cube(d_rho, d_pot)
    direct_ms = ctx.time_end()
    direct = d_pot.download()
    d_tot.set_to_zero()
    bub.inject(grid, d_tot)
    inj = d_tot.download()
    inj_ref = O.bubbles_eval_grid(bset, ax.points, ax.points, ax.points[:8])
    out = {
        "ms_per_step": ms, "wall_s": wall, "n": n, "T": T, "kernels": kernels,
        "roofline": roofline,
        "bf16x6_nearfield": {"ms_per_step": round(ms_mixed, 3), "speedup": round(ms / ms_mixed, 3),
                             "max_rel_diff_vs_fp64": rel(pot_mixed, pot)},
        "gbfmm_vs_analytic": rel(pot_g, exact),
        "gbfmm_vs_direct_apply": rel(pot_g, direct),
        "direct_apply_ms": round(direct_ms, 3),
        "inject_vs_oracle_8_planes": rel(inj[:, :, :8], inj_ref),
        "inject_valu_fp64_tflops": round(FLOP_PER_POINT_BUBBLE * 4 * n ** 3 /
                                         (kernels["bubbles_inject_kernel"]["ms_per_launch"] * 1e-3)
                                         / 1e12, 2),
    }
    # CPU baseline: the oracle's near field of four interior boxes (181^3 ->
    # 61^3 each, all 50 t-points; the four boxes whose neighbourhoods start
    # at cell 0 or 10 along x and y) and its inject on 32 z-planes, scaled
    # by work to the whole step
    if args.config2_cpu and not args.no_cpu_baseline:
        s = 10
        sub_in = O.sub_axis(ax, 0, 3 * s)
        sub_out = O.sub_axis(ax, s, 2 * s)
        mats = O.coulomb3d_matrices([sub_in] * 3, [sub_out] * 3, gq.tp)
        m = sub_in.n
        h = 6 * s  # points per box
        t_boxes = []
        for ox, oy in ((0, 0), (h, 0), (0, h), (h, h)):
            nf = np.asfortranarray(rho[ox:ox + m, oy:oy + m, :m])
            c0 = time.perf_counter()
            O.transform_cube(*mats, O.coulomb_weights(gq.tw), 0.0, nf, add_coda=False)
            t_boxes.append(time.perf_counter() - c0)
        t_box = sum(t_boxes) / len(t_boxes)
        c0 = time.perf_counter()
        O.bubbles_eval_grid(bset, ax.points, ax.points, ax.points[:32])
        t_inj = (time.perf_counter() - c0) / 4

        def box_flops(nx, ny, nz, no=61):
            return 2 * T * (no * nx * ny * nz + no * no * ny * nz + no ** 3 * nz)
        nfs = [121, 181, 181, 121]
        total = sum(box_flops(nfs[x], nfs[y], nfs[z])
                    for x in range(4) for y in range(4) for z in range(4))
        est = t_box * total / box_flops(181, 181, 181) + t_inj * n / 8
        out["cpu_baseline"] = {
            "value": round(n ** 3 / est / 1e9, 6), "unit": "Gpts/s", "cores": 1, "kind": "port",
            "sample": f"oracle near field of four interior boxes (181^3 -> 61^3, T=50) "
                      f"{sum(t_boxes):.2f} s and oracle inject on 32 of 241 z-planes "
                      f"{4 * t_inj:.2f} s, scaled by FLOPs and planes to one step ({est:.1f} s); "
                      f"multipoles, translations and local expansions (under 3% of the step on "
                      f"the GPU) not included"}
    for obj in (op, fmm, bub, d_rho, d_tot, d_pot, grid):
        obj.close()
    return out


# ----------------------------------------------------------------- config 1
def bench_config1(ctx, args):
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
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, nlip)
    quad = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
    op = dage_amd.Coulomb3D(ctx, grid, grid, quad)
    d_in = dage_amd.DeviceCube(ctx, rho.shape, rho)
    d_out = dage_amd.DeviceCube(ctx, rho.shape)
    steps = max(args.steps, 10)
    for _ in range(args.warmup):
        op.transform_cuda_cube(d_in, d_out)
    ctx.profile(True)
    ctx.time_begin()
    for _ in range(steps):
        op.transform_cuda_cube(d_in, d_out)
    ms = ctx.time_end() / steps
    kernels = {}
    for name in ("xy_fused_kernel", "gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode",
                 "band_xmode", "band_ymode", "band_xymode", "band_zmode", "lowrank",
                 "splitk_reduce_kernel"):
        tot, k = ctx.profile_query(name)
        if k:
            kernels[name] = {"ms_per_launch": round(tot / k, 4), "launches_per_apply": k // steps}
    ctx.profile(False)
    V = d_out.download()
    mixed = {}
    d_mix = dage_amd.DeviceCube(ctx, rho.shape)
    for mode, label in ((dage_amd.DAGE_PREC_BF16X6, "bf16x6"), (dage_amd.DAGE_PREC_BF16X3, "bf16x3")):
        op.transform_cuda_cube(d_in, d_mix, precision=mode)
        ctx.time_begin()
        for _ in range(steps):
            op.transform_cuda_cube(d_in, d_mix, precision=mode)
        mms = ctx.time_end() / steps
        mixed[label] = (mms, d_mix.download())
    flops = 6.0 * T * float(N) ** 4
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, quad.tp)
    c0 = time.perf_counter()
    ref = O.transform_cube(fx, fy, fz, O.coulomb_weights(quad.tw), O.coda(500.0), rho)
    cpu_s = time.perf_counter() - c0
    Va = O.gauss_potential(q, alpha, centers, ax.points, ax.points, ax.points)
    out = {
        "workload": "config 1: plain Coulomb3D apply, FP64, 97^3 over [-8,8]^3, nlip 7, T=40, "
                    "4 Gaussians, coda on",
        "ms_per_apply": round(ms, 4),
        "Gpts_s": round(N ** 3 / (ms * 1e-3) / 1e9, 4),
        "tflops": round(flops / (ms * 1e-3) / 1e12, 3),
        "mfma_frac": round(flops / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4),
        "kernels": kernels,
        "traffic": {k: measured_traffic("config1", k) for k in kernels},
        "max_rel_err_vs_oracle": rel(V, ref),
        "max_rel_err_vs_analytic": rel(V, Va),
        "mixed": {label: {"ms_per_apply": round(mms, 4), "speedup_vs_fp64": round(ms / mms, 3),
                          "max_rel_err_vs_oracle": rel(vm, ref)}
                  for label, (mms, vm) in mixed.items()},
        "cpu_baseline": {"value": round(N ** 3 / cpu_s / 1e9, 6), "unit": "Gpts/s", "cores": 1,
                         "kind": "port",
                         "sample": f"full 97^3 T=40 apply, oracle C restatement of "
                                   f"Operator3D_transform_cube, 1 thread: {cpu_s:.2f} s"},
    }
    for obj in (d_in, d_out, d_mix, op, grid):
        obj.close()
    return out


# ----------------------------------------------------------------- config 3
def bench_config3(ctx, args):
    import numpy as np

    import dage_amd
    import oracle as O
    rng = np.random.default_rng(0xDA6E0003)
    nrhs = 32
    sizes = rng.choice([65, 97, 129, 161], nrhs)
    k2 = rng.uniform(0.1, 4.0, nrhs)
    energies = -0.5 * k2
    gq = dage_amd.GaussQuad(20, 20, tlin=2.0, tlog=500.0)
    T = gq.get_npoints()
    axes, grids = {}, {}
    for nn in sorted(set(int(s) for s in sizes)):
        ncell = (nn - 1) // 4
        half = ncell * 4 * 0.15 / 2
        axes[nn] = O.Axis(-half, np.full(ncell, 0.15), 5)
        grids[nn] = dage_amd.Grid3D(ctx, [-half] * 3, [axes[nn].steps] * 3, 5)
    rhos, gparams = [], []
    for s in sizes:
        ax = axes[int(s)]
        a = rng.uniform(0.5, 5.0, 4)
        c = rng.uniform(-2.0, 2.0, (4, 3))
        gparams.append((a, c))
        rhos.append(O.gauss_density(np.ones(4), a, c, ax.points, ax.points, ax.points))
    pts = float(sum(int(s) ** 3 for s in sizes))
    flops = float(sum(6.0 * T * int(s) ** 4 for s in sizes))
    gl = [grids[int(s)] for s in sizes]
    # plain: one FP64 apply per orbital with set_energy in between
    ops = {nn: dage_amd.Helmholtz3D(ctx, grids[nn], grids[nn], gq) for nn in grids}
    d_in = [dage_amd.DeviceCube(ctx, r.shape, r) for r in rhos]
    d_out = [dage_amd.DeviceCube(ctx, r.shape) for r in rhos]

    def plain():
        for i, s in enumerate(sizes):
            op = ops[int(s)]
            op.set_energy(energies[i])
            op.transform_cuda_cube(d_in[i], d_out[i])

    plain()
    reps = 2
    ctx.time_begin()
    for _ in range(reps):
        plain()
    plain_ms = ctx.time_end() / reps
    plain_out = [d.download() for d in d_out]
    res = {}
    for label, prec in (("fp64", dage_amd.DAGE_PREC_FP64), ("bf16x6", dage_amd.DAGE_PREC_BF16X6)):
        b = dage_amd.HelmholtzBatch(ctx, gl, energies, gq, precision=prec)
        t0 = time.perf_counter()
        tiles = b.tune(reps=1)
        tune_s = time.perf_counter() - t0
        for i, r in enumerate(rhos):
            b.upload(i, r)
        b.apply()
        ctx.time_begin()
        for _ in range(reps):
            b.apply()
        ms = ctx.time_end() / reps
        outs = [b.download(i) for i in range(nrhs)]
        res[label] = {
            "ms_per_batch": round(ms, 3),
            "Gpts_s": round(pts / (ms * 1e-3) / 1e9, 4),
            "dense_fp64_equiv_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
            "speedup_vs_plain_fp64": round(plain_ms / ms, 3),
            "max_rel_diff_vs_plain_fp64": max(rel(o, p) for o, p in zip(outs, plain_out)),
            "tuned_tiles": tiles.tolist(),
            "tune_s": round(tune_s, 2),
        }
        b.close()
    # closed-form check on the first two 65^3 orbitals
    checks = []
    for i, s in enumerate(sizes):
        if int(s) == 65 and len(checks) < 2:
            ax = axes[65]
            a, c = gparams[i]
            ex = O.yukawa_gauss_potential(np.ones(4), a, c, np.sqrt(k2[i]), ax.points, ax.points,
                                          ax.points)
            checks.append(rel(plain_out[i], ex))
    for obj in list(ops.values()) + d_in + d_out + list(grids.values()):
        obj.close()
    return {
        "workload": "config 3: 32 Helmholtz right-hand sides, k^2 U[0.1,4], cubes "
                    "U{65,97,129,161}^3 at h 0.15 (nlip 5), T=40, 4 Gaussians each",
        "sizes": [int(s) for s in sizes],
        "points": pts,
        "dense_fp64_flops": flops,
        "plain_fp64_ms": round(plain_ms, 3),
        "plain_fp64_Gpts_s": round(pts / (plain_ms * 1e-3) / 1e9, 4),
        "batched": res,
        "plain_vs_yukawa_closed_form": checks,
    }


# ----------------------------------------------------------------- config 4
def bench_config4(ctx, args):
    """Helmholtz SCF time to 1e-6 Ha, plain vs learned preconditioner.

    Both start from atom-centred Gaussians exp(-Z r^2 / 2) (s on every atom,
    p on the heaviest) rotated to the Ritz vectors of the nuclear Hamiltonian;
    the learned run evaluates the network until three of its proposals in a
    row lose to the plain step, then every 8th iteration until one wins.
    Both runs autotune their 12-orbital Helmholtz batch graph once first.

    129^3 cube over [-9.6, 9.6]^3 (nlip 5, h 0.15), 10 atoms (Z uniform in
    {1,6,7,8}, in a 6-bohr sphere, >= 1.8 bohr apart, nuclei smeared with
    beta 4), 12 doubly occupied orbitals, Hartree + Slater exchange.  The
    network is trained once, offline, on the plain iterations of another
    molecule from the same generator; the timed runs are the plain and the
    learned SCF of a molecule it has not seen, from the same initial guess,
    each to the convergence rule |dE| < 1e-6 Ha and all |d eps| < 1e-6."""
    import numpy as np

    import dage_amd
    import dage_amd.learned as L
    import dage_amd.scf as scf
    grid = dage_amd.Grid3D(ctx, [-9.6] * 3, [np.full(32, 0.15)] * 3, 5)
    train_mol = scf.random_molecule(np.random.default_rng(1), 10)
    test_mol = scf.random_molecule(np.random.default_rng(0), 10)
    nocc, maxit = 12, args.scf_max_iter

    class EveryK(L.SampleRecorder):
        def __init__(self, s, k):
            super().__init__(s)
            self.k, self.calls = k, 0

        def __call__(self, psi, delta, out):
            self.calls += 1
            if self.calls <= args.scf_record_iter and self.calls % self.k == 1:
                return super().__call__(psi, delta, out)
            for p_, d_, o_ in zip(psi, delta, out):
                o_.copy_from(p_)
                self.scf._axpby(self.scf.w, d_, 1.0, o_)

    t0 = time.perf_counter()
    tr = scf.HelmholtzSCF(ctx, grid, train_mol, nocc=nocc, beta=4.0)
    tr.initial_guess(exponents=train_mol.Z.copy())
    rec = EveryK(tr, 4)
    r_train = tr.run(max_iter=args.scf_train_iter, update=rec)
    model = L.ConvPreconditioner(ctx, tr.shape, seed=0, lr=2e-3)
    samples = rec.samples(model, tr.psi)
    rec.close()
    tr.close()
    c0 = time.perf_counter()
    losses = L.train(model, samples, epochs=2)
    ctx.synchronize()
    train_s = time.perf_counter() - c0
    progress(f"config4: trained on {len(samples)} pairs in {train_s:.1f} s")
    for xb, tb in samples:
        xb.close()
        tb.close()
    offline_s = time.perf_counter() - t0
    runs = {}
    for mode in ("plain", "learned"):
        S = scf.HelmholtzSCF(ctx, grid, test_mol, nocc=nocc, beta=4.0, tune=True)
        S.initial_guess(exponents=test_mol.Z.copy())
        upd = None if mode == "plain" else (lambda p_, d_, o_, S=S: model.propose(S, p_, d_, o_))
        r = S.run(max_iter=maxit, update=upd, backoff=(3, 8))
        runs[mode] = {"converged": r.converged, "iterations": r.iterations,
                      "wall_s": round(r.wall_s, 3),
                      "ms_per_iteration": round(1e3 * r.wall_s / r.iterations, 2),
                      "final_energy": r.energies[-1],
                      "last_dE": abs(r.energies[-1] - r.energies[-2]),
                      "learned_proposals_kept": int(sum(r.learned_used))}
        if not r.converged:
            runs[mode]["note"] = f"1e-6 Ha not reached within {maxit} iterations"
        progress(f"config4 {mode}: {r.iterations} iterations, {r.wall_s:.1f} s, "
                 f"converged {r.converged}")
        S.close()
    model.close()
    # the same molecule with bubble-carrying orbitals: closed-shell Hartree-Fock
    # (point nuclei -Z/r as bubbles, lmax-4 orbital bubbles on 199-point
    # radial grids, Taylor-series products, exact exchange), guess plus two
    # iterations timed
    hf = scf.BubbleHartreeFock(ctx, grid, test_mol, nocc)
    c0 = time.perf_counter()
    e_hf = [hf.initial_guess()]
    ctx.synchronize()
    guess_s = time.perf_counter() - c0
    it_s = []
    for _ in range(2):
        c0 = time.perf_counter()
        e_hf.append(hf.iterate())
        ctx.synchronize()
        it_s.append(time.perf_counter() - c0)
    bubble_hf = {
        "workload": "12 orbitals = cube + lmax-4 bubbles (199 radial points) on every atom, "
                    "point nuclei, RHF with exact exchange; guess + 2 iterations",
        "guess_s": round(guess_s, 2), "s_per_iteration": [round(x, 2) for x in it_s],
        "energies": e_hf, "eps_last": [round(float(x), 6) for x in hf.eps]}
    progress(f"config4 bubble RHF: guess {guess_s:.1f} s, iterations {it_s}")
    hf.close()
    grid.close()
    p_, l_ = runs["plain"], runs["learned"]
    return {
        "workload": "config 4: Helmholtz SCF, 129^3 over [-9.6,9.6]^3 (nlip 5, h 0.15), 10 atoms "
                    "Z~U{1,6,7,8} smeared (beta 4), 12 orbitals, Hartree + Slater exchange, FP64 "
                    "batched Helmholtz applies; time to |dE| < 1e-6 Ha and |d eps| < 1e-6",
        "Z": [float(z) for z in test_mol.Z],
        "plain": p_, "learned": l_,
        "time_to_1e-6_s": {"plain": p_["wall_s"] if p_["converged"] else None,
                           "learned": l_["wall_s"] if l_["converged"] else None},
        "wall_ratio_plain_over_learned": round(p_["wall_s"] / l_["wall_s"], 3),
        "bubble_rhf": bubble_hf,
        "energy_learned_minus_plain": l_["final_energy"] - p_["final_energy"],
        "training": {"molecule_Z": [float(z) for z in train_mol.Z],
                     "plain_iterations": r_train.iterations, "samples": len(samples),
                     "epochs": 2, "loss_per_epoch": [round(x, 4) for x in losses],
                     "train_s": round(train_s, 2), "offline_total_s": round(offline_s, 2)},
    }


# ----------------------------------------------------------------- config 5
def config5_setup(ctx):
    """The 385^3 combined workload's inputs, resident on the GPU."""
    import numpy as np

    import dage_amd
    import dage_amd.scf as scf
    import oracle as O
    rng = np.random.default_rng(0xDA6E0005)
    ax = O.Axis.ncell_fixed(-19.2, 19.2, 7, 64)
    n = ax.n
    grid = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    q = np.ones(8)
    alpha = np.exp(rng.uniform(np.log(0.3), np.log(30.0), 8))
    centers = rng.uniform(-6.0, 6.0, (8, 3))
    rho = O.gauss_density(q, alpha, centers, ax.points, ax.points, ax.points)
    alpha2 = np.exp(rng.uniform(np.log(0.3), np.log(30.0), 8))
    centers2 = rng.uniform(-6.0, 6.0, (8, 3))
    d_rho = dage_amd.DeviceCube(ctx, rho.shape, rho)
    del rho
    d_rho2 = dage_amd.DeviceCube(
        ctx, (n, n, n), O.gauss_density(q, alpha2, centers2, ax.points, ax.points, ax.points))
    # 24 atoms from the molecule generator (Z in {1,6,7,8}, >= 1.8 bohr apart
    # in a 6-bohr sphere), bubbles lmax 4 on 199-point radial grids
    mol = scf.random_molecule(rng, 24)
    lmax = 4
    bub = dage_amd.Bubbles(ctx, 24, lmax, 0, 7)
    steps, bfs = [], []
    for a in range(24):
        st = dage_amd.radial_grid_steps(mol.Z[a], 33, 7, 20.0)
        r = O.Axis(0.0, st, 7).points
        c = rng.uniform(-1.0, 1.0, (lmax + 1) ** 2)
        bf = np.zeros((len(r), (lmax + 1) ** 2), order="F")
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                bf[:, l * l + l + m] = c[l * l + l + m] * r ** l * np.exp(-mol.Z[a] * r)
        bub.set_bubble(a, mol.R[a], mol.Z[a], st, bf)
        steps.append(st)
        bfs.append(bf)
    bset = O.BubbleSet(mol.R, steps, bfs, lmax, 0, 7)
    nrhs = 64
    k2 = rng.uniform(0.05, 6.0, nrhs)
    coef = rng.uniform(-1.0, 1.0, nrhs)
    return dict(ax=ax, n=n, grid=grid, gauss=(q, alpha, centers), d_rho=d_rho, d_rho2=d_rho2,
                mol=mol, bub=bub, bset=bset, energies=-0.5 * k2, coef=coef, nrhs=nrhs)


def fill_orbitals(ctx, S, batch):
    """Orbital i's right-hand side = rho + coef_i rho2, built on the device."""
    from dage_amd._lib import call
    npts = S["n"] ** 3
    for i in range(S["nrhs"]):
        dst = ctypes.c_void_p(batch.input_ptr(i))
        call("dage_memcpy", ctx.handle, dst, ctypes.c_void_p(S["d_rho2"].ptr),
             ctypes.c_size_t(8 * npts), 2)
        call("dage_vec_axpby", ctx.handle, npts, 1.0, ctypes.c_void_p(S["d_rho"].ptr),
             float(S["coef"][i]), dst)


def bench_config5(ctx, args):
    """Headline: one step of the 385^3 combined workload = the resident
    density (8 Gaussians) plus 24 atoms' bubbles injected, its Coulomb
    potential by GBFMM (8x8x8 boxes, lmax 12, near field T = 50), and one
    Helmholtz sweep over 64 orbitals (k^2 uniform in [0.05, 6], T = 40) as
    one batch, all FP64 on the GPU; `value` = output points of the step
    (65 x 385^3) per second."""
    import numpy as np

    import dage_amd
    import oracle as O
    S = config5_setup(ctx)
    n, grid, nrhs = S["n"], S["grid"], S["nrhs"]
    gq = dage_amd.GaussQuad(25, 25)
    hq = dage_amd.GaussQuad(20, 20, tlin=2.0, tlog=500.0)
    c0 = time.perf_counter()
    fmm = dage_amd.GBFMMCoulomb3D(ctx, grid, maxlevel=3, lmax=12, gauss=gq)
    batch = dage_amd.HelmholtzBatch(ctx, [grid] * nrhs, S["energies"], hq)
    setup_s = time.perf_counter() - c0
    fill_orbitals(ctx, S, batch)
    d_tot = dage_amd.DeviceCube(ctx, (n, n, n))
    d_pot = dage_amd.DeviceCube(ctx, (n, n, n))
    progress(f"config5: setup {setup_s:.1f} s")

    def step(eager=False):
        d_tot.copy_from(S["d_rho"])
        S["bub"].inject(grid, d_tot)
        fmm.apply_device(d_tot, d_pot)
        if eager:
            batch.apply_eager()
        else:
            batch.apply()

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    # timed region: the library's eager launches with per-kernel HIP events
    # on the context stream (the kernels' own durations for the roofline)
    ctx.profile(True)
    wall0 = time.perf_counter()
    ctx.time_begin()
    for _ in range(args.steps):
        step(eager=True)
    ms = ctx.time_end() / args.steps
    wall = time.perf_counter() - wall0
    # batch_*: the sweep's passes (the batch's own event names); gemm_f64_*:
    # the GBFMM near field's
    names = ("batch_xy_fused_kernel", "batch_f64_xmode", "batch_f64_ymode", "batch_f64_zmode",
             "batch_band_xmode", "batch_band_ymode", "batch_band_xymode", "batch_band_zmode",
             "batch_lowrank",
             "batch_splitk_reduce",
             "xy_fused_kernel",
             "gemm_f64_xmode", "gemm_f64_ymode", "gemm_f64_zmode", "band_xmode", "band_ymode",
             "band_xymode", "band_zmode", "lowrank",
             "splitk_reduce_kernel",
             "bubbles_inject_kernel", "box_multipoles_kernel", "box_multipoles_gemm",
             "m2m_kernel", "m2l_gemm", "l2l_kernel", "local_expansion_kernel",
             "local_expansion_gemm", "cube_copy_kernel")
    kernels = {}
    for k in names:
        kms, kn = ctx.profile_query(k)
        if kn:
            kernels[k] = {"ms_per_launch": round(kms / kn, 4), "launches_per_step": kn // args.steps,
                          "ms_per_step": round(kms / args.steps, 4)}
    ctx.profile(False)
    # the same step with the batch as one graph launch
    ctx.time_begin()
    for _ in range(args.steps):
        step()
    ms_graph = ctx.time_end() / args.steps
    progress(f"config5: {ms:.0f} ms per step (graph {ms_graph:.0f} ms)")
    if args.headline_only:
        print(json.dumps({"mode": "headline_only", "ms_per_step": round(ms, 2),
                          "kernels": kernels}), flush=True)
        fmm.close()
        batch.close()
        return None
    T, Tn = hq.get_npoints(), gq.get_npoints()
    # dominant kernel of the step; a mode pass does 2 T n^4 dense FLOPs per
    # right-hand side (sweep: T = 40 for each of 64 orbitals; GBFMM near
    # field: T = 50 once), the fused x/y pass twice that, counted dense
    # although the all-zero K panels of the banded matrices (exact zeros and
    # the tails stored as zeros below 1e-20 of their row's largest entry) are
    # skipped: `frac` can exceed 1, `frac_executed` counts the executed FLOPs
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dom]
    mfma_kernels = ("batch_f64_", "gemm_f64_", "batch_xy_fused", "xy_fused")
    if dom.startswith(mfma_kernels):
        passes = 2.0 if "xy_fused" in dom else 1.0
        # the MFMA passes cover the t-points between the low-rank head (applied
        # through factors) and the banded tail (VALU passes)
        if "xy_fused" in dom:
            tsum = T * nrhs if dom.startswith("batch_") else Tn
        elif dom.startswith("batch_"):
            probe = dage_amd.Helmholtz3D(ctx, grid, grid, hq)
            til = probe.tiling()
            tsum = (til["band_t0"] - til["lowrank_t1"]) * nrhs
            probe.close()
        else:
            tsum = fmm.band_t0() - fmm.lowrank_t1()
        per_step = passes * 2.0 * float(n) ** 4 * tsum
        per_launch = per_step / dk["launches_per_step"]
        achieved = per_launch / (dk["ms_per_launch"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3),
                    "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4),
                    "traffic": measured_traffic("config5", dom),
                    "flops": "dense-equivalent (2 T n^4 per pass and right-hand side; the "
                             "fused x/y kernel does two passes); the skipped all-zero panels "
                             "count, so frac can exceed 1: frac_executed is the MFMA use",
                    "avg_launch_ms": dk["ms_per_launch"],
                    "launches_per_step": dk["launches_per_step"],
                    "algorithmic_flops_per_launch": per_launch}
        exe = measured_executed_fraction("config5", dom)
        if exe is not None:
            roofline["executed_fraction"] = exe
            roofline["frac_executed"] = round(exe * achieved / FP64_MFMA_PEAK_TFLOPS, 4)
    else:
        per_launch = 16.0 * n ** 3
        achieved = per_launch / (dk["ms_per_launch"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic("config5", dom),
                    "avg_launch_ms": dk["ms_per_launch"], "algorithmic_bytes_per_launch": per_launch}
    out = {"ms_per_step": ms, "ms_per_step_graph": round(ms_graph, 2), "wall_s": wall, "n": n,
           "T_sweep": T, "T_near": Tn, "points_per_step": (nrhs + 1) * n ** 3,
           "kernels": kernels, "roofline": roofline, "setup_s": round(setup_s, 2)}
    # parity on the GPU results of the step
    inj = S["bset"]
    d_tot.set_to_zero()
    S["bub"].inject(grid, d_tot)
    z0 = n // 2
    got = d_tot.download()[:, :, z0:z0 + 2]
    ax = S["ax"]
    out["inject_vs_oracle_2_planes"] = rel(got, O.bubbles_eval_grid(inj, ax.points, ax.points,
                                                                    ax.points[z0:z0 + 2]))
    fmm.apply_device(S["d_rho"], d_pot)
    q, alpha, centers = S["gauss"]
    zs = ax.points[z0 - 4:z0 + 4]
    pot_fp64 = d_pot.download()
    out["gbfmm_vs_analytic_8_planes"] = rel(pot_fp64[:, :, z0 - 4:z0 + 4],
                                            O.gauss_potential(q, alpha, centers, ax.points,
                                                              ax.points, zs))
    hop = dage_amd.Helmholtz3D(ctx, grid, grid, hq)
    d_mix = dage_amd.DeviceCube(ctx, (n, n, n))
    i_s = nrhs - 1
    d_mix.copy_from(S["d_rho2"])
    from dage_amd._lib import call
    call("dage_vec_axpby", ctx.handle, n ** 3, 1.0, ctypes.c_void_p(S["d_rho"].ptr),
         float(S["coef"][i_s]), ctypes.c_void_p(d_mix.ptr))
    hop.set_energy(S["energies"][i_s])
    hop.transform_cuda_cube(d_mix, d_pot)
    single = d_pot.download()
    sweep_fp64 = batch.download(i_s)
    out["sweep_vs_single_apply_orbital_63"] = rel(sweep_fp64, single)
    # plain per-orbital loop (the reference's order: set_energy + apply per
    # orbital) for the batching gain
    ctx.time_begin()
    for i in range(nrhs):
        hop.set_energy(S["energies"][i])
        hop.transform_cuda_cube(S["d_rho"], d_pot)
    plain_ms = ctx.time_end()
    batch.apply()
    ctx.time_begin()
    batch.apply()
    sweep_ms = ctx.time_end()
    hop.close()
    batch.close()
    progress("config5: FP64 step and parity done")
    # split-BF16 (bf16x6) near field and sweep, error against FP64
    fmm.apply_device(S["d_rho"], d_pot, dage_amd.DAGE_PREC_BF16X6)
    ctx.time_begin()
    fmm.apply_device(S["d_rho"], d_pot, dage_amd.DAGE_PREC_BF16X6)
    gbfmm_mixed_ms = ctx.time_end()
    gbfmm_mixed_err = rel(d_pot.download(), pot_fp64)
    ctx.time_begin()
    fmm.apply_device(S["d_rho"], d_pot)
    gbfmm_ms = ctx.time_end()
    del pot_fp64
    mb = dage_amd.HelmholtzBatch(ctx, [grid] * nrhs, S["energies"], hq,
                                 precision=dage_amd.DAGE_PREC_BF16X6)
    fill_orbitals(ctx, S, mb)
    mb.apply()
    ctx.time_begin()
    mb.apply()
    sweep_mixed_ms = ctx.time_end()
    sweep_mixed_err = rel(mb.download(i_s), sweep_fp64)
    mb.close()
    flops_sweep = 6.0 * T * float(n) ** 4 * nrhs
    out["gbfmm"] = {"fp64_ms": round(gbfmm_ms, 2), "bf16x6_ms": round(gbfmm_mixed_ms, 2),
                    "bf16x6_vs_fp64": gbfmm_mixed_err}
    out["sweep_64"] = {
        "plain_fp64_per_orbital_ms": round(plain_ms, 1),
        "batched_fp64_ms": round(sweep_ms, 1),
        "batched_bf16x6_ms": round(sweep_mixed_ms, 1),
        "batched_fp64_dense_equiv_tflops": round(flops_sweep / (sweep_ms * 1e-3) / 1e12, 2),
        "speedup_fp64_batch_vs_plain": round(plain_ms / sweep_ms, 3),
        "speedup_bf16x6_batch_vs_plain_fp64": round(plain_ms / sweep_mixed_ms, 3),
        "bf16x6_vs_fp64_orbital_63": sweep_mixed_err}
    progress("config5: split-BF16 legs done")
    # GBFMM-Helmholtz on the same grid (kernel exp(-kappa r)/r, k^2 = 2):
    # switching the operator (far-field matrices rebuilt), the apply, and
    # the direct whole-grid Helmholtz apply with the same t-quadrature (no
    # coda) beside it
    eh = -1.0
    c0 = time.perf_counter()
    fmm.set_energy(eh)
    switch_s = time.perf_counter() - c0
    fmm.apply_device(S["d_rho"], d_pot)
    ctx.time_begin()
    fmm.apply_device(S["d_rho"], d_pot)
    gh_ms = ctx.time_end()
    gh = d_pot.download()
    hd = dage_amd.Helmholtz3D(ctx, grid, grid, gq, coda=False)
    hd.set_energy(eh)
    hd.transform_cuda_cube(S["d_rho"], d_pot)
    ctx.time_begin()
    hd.transform_cuda_cube(S["d_rho"], d_pot)
    hdir_ms = ctx.time_end()
    out["gbfmm_helmholtz"] = {
        "energy": eh, "kappa": round(math.sqrt(-2 * eh), 4), "set_energy_s": round(switch_s, 3),
        "gbfmm_ms": round(gh_ms, 2), "direct_apply_ms": round(hdir_ms, 2),
        "gbfmm_vs_direct": rel(gh, d_pot.download())}
    del gh
    hd.close()
    fmm.set_energy(0.0)
    progress("config5: GBFMM-Helmholtz leg done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_config5(S, gq, hq)
    fmm.close()
    for obj in (d_tot, d_pot, d_mix, S["d_rho"], S["d_rho2"], S["bub"], grid):
        obj.close()
    return out


def cpu_threads():
    """Host threads for the all-core CPU baseline: OMP_NUM_THREADS when set
    (the GPU host's CPU share), else the cores this process may use."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_config5(S, gq, hq):
    """The oracle (C restatement of Operator3D_transform_cube, the GBFMM near
    field and Bubbles_eval_3dgrid) on bounded samples of one config-5 step,
    at 1 thread and at all threads (OpenMP over independent slices, same
    results), scaled by FLOPs / points to the whole step."""
    import numpy as np

    import oracle as O
    ax, n = S["ax"], S["n"]
    T, Tn = hq.get_npoints(), gq.get_npoints()
    # (a) one orbital's Helmholtz apply on 8 input z-planes and 10 of 40
    # t-points: 6 n^3 * 8 * 10 FLOPs of the 6 n^4 * 40 per orbital
    zs = slice(n // 2 - 4, n // 2 + 4)
    fx, fy, fz = O.coulomb3d_matrices([ax] * 3, [ax] * 3, hq.tp)
    rho = O.gauss_density(*S["gauss"], ax.points, ax.points, ax.points[zs])
    fz_s = np.asfortranarray(fz[zs])
    w = O.helmholtz_weights(hq.tp, hq.tw, S["energies"][0])
    # (b) the near field of one interior box (145^3 -> 49^3, all 50 t-points)
    s = 8
    sub_in = O.sub_axis(ax, 2 * s, 5 * s)
    sub_out = O.sub_axis(ax, 3 * s, 4 * s)
    nmats = O.coulomb3d_matrices([sub_in] * 3, [sub_out] * 3, gq.tp)
    m = sub_in.n
    o0 = 2 * s * 6
    nf = np.asfortranarray(O.gauss_density(*S["gauss"], ax.points[o0:o0 + m], ax.points[o0:o0 + m],
                                           ax.points[o0:o0 + m]))
    wn = O.coulomb_weights(gq.tw)

    def box_flops(nx, ny, nz, no=49):
        return 2 * Tn * (no * nx * ny * nz + no * no * ny * nz + no ** 3 * nz)
    nfs = [97] + [145] * 6 + [97]
    near_total = sum(box_flops(nfs[x], nfs[y], nfs[z])
                     for x in range(8) for y in range(8) for z in range(8))
    res = {}
    for threads in (1, cpu_threads()):
        O.set_threads(threads)
        c0 = time.perf_counter()
        O.transform_cube(fx, fy, fz_s, w, 0.0, rho, t0=0, t1=10, add_coda=False)
        t_apply = time.perf_counter() - c0
        c0 = time.perf_counter()
        O.transform_cube(*nmats, wn, 0.0, nf, add_coda=False)
        t_box = time.perf_counter() - c0
        c0 = time.perf_counter()
        O.bubbles_eval_grid(S["bset"], ax.points, ax.points, ax.points[zs][:2])
        t_inj = time.perf_counter() - c0
        est = (t_apply * (n / 8) * (T / 10) * S["nrhs"]
               + t_box * near_total / box_flops(145, 145, 145) + t_inj * n / 2)
        res[threads] = (est, t_apply + t_box + t_inj, t_apply, t_box, t_inj)
    O.set_threads(cpu_threads())
    pts = (S["nrhs"] + 1) * n ** 3
    nt = cpu_threads()
    est1, sample1 = res[1][0], res[1][1]
    estn, samplen = res[nt][0], res[nt][1]
    return {
        "value": round(pts / estn / 1e9, 8), "unit": "Gpts/s", "cores": nt, "kind": "port",
        "value_1_core": round(pts / est1 / 1e9, 8),
        "cpu_model": cpu_model(), "nproc": os.cpu_count(),
        "seconds_per_step": {"1": round(est1, 1), str(nt): round(estn, 1)},
        "sample": f"oracle C restatement on the GPU host: one orbital's Helmholtz apply on 8 of "
                  f"{n} input z-planes and 10 of {T} t-points (scaled x{n // 8 * T // 10} and x64 "
                  f"orbitals), the near field of 1 of 512 boxes (145^3 -> 49^3, T={Tn}, scaled by "
                  f"FLOPs), the 24-atom inject on 2 of {n} z-planes; multipoles, translations and "
                  f"local expansions (about 1 % of the GPU step) not included; sample "
                  f"{sample1:.1f} s at 1 thread, {samplen:.1f} s at {nt} threads"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; "
              "run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args)
        return
    import dage_amd
    ctx = dage_amd.Context(0)
    c5 = bench_config5(ctx, args)
    if c5 is None:  # --headline-only
        ctx.close()
        return
    progress(f"config5: {c5['ms_per_step']:.1f} ms per step")
    n = c5["n"]
    line = {
        "metric": METRIC,
        "value": round(c5["points_per_step"] / (c5["ms_per_step"] * 1e-3) / 1e9, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c5["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "config 5: 385^3 cube over [-19.2,19.2]^3 (nlip 7, 64 cells, 57.07 M "
                        "points); per step: 8 Gaussians + 24 atoms' bubbles (lmax 4, 199-point "
                        "radial grids) injected, their Coulomb potential by GBFMM (8x8x8 boxes, "
                        "lmax 12, near field T=50), and one Helmholtz sweep over 64 orbitals "
                        "(k^2 U[0.05,6], T=40, one batch); FP64; value = 65 x 385^3 output "
                        "points per second",
            "n": n, "orbitals": 64, "atoms": 24, "maxlevel": 3, "lmax": 12,
            "T_sweep": c5["T_sweep"], "T_near": c5["T_near"],
        },
        "roofline": c5["roofline"],
        "cpu_baseline": c5.get("cpu_baseline"),
        "kernels": c5["kernels"],
        "ms_per_step_graph": c5["ms_per_step_graph"],
        "inject_vs_oracle_2_planes": c5["inject_vs_oracle_2_planes"],
        "gbfmm_vs_analytic_8_planes": c5["gbfmm_vs_analytic_8_planes"],
        "sweep_vs_single_apply_orbital_63": c5["sweep_vs_single_apply_orbital_63"],
        "gbfmm": c5["gbfmm"],
        "gbfmm_helmholtz": c5["gbfmm_helmholtz"],
        "sweep_64": c5["sweep_64"],
        "setup_s": c5["setup_s"],
        "wall_s_timed": round(c5["wall_s"], 3),
    }
    if args.config2:
        line["config2"] = bench_config2(ctx, args)
        progress(f"config2: {line['config2']['ms_per_step']:.2f} ms per step")
    if args.config1:
        line["config1"] = bench_config1(ctx, args)
        progress("config1 done")
    if args.config3:
        line["config3"] = bench_config3(ctx, args)
        progress("config3 done")
    if args.config4:
        line["config4"] = bench_config4(ctx, args)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
