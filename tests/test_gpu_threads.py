"""The C-ABI called from several host threads at once, as the reference calls
its GPU path from three OpenMP threads (gbfmm_coulomb3d.F90:767-785).

Each thread owns a context (own stream), grid, operator and input and runs
its applies while the others run theirs (ctypes releases the GIL for the
duration of every library call); the results must equal, bit for bit, the
same applies run one thread at a time.  A thread that passes bad arguments
meanwhile, and each apply thread between its applies, must read its own
error message (the error buffer is thread-local).
"""
import threading

import numpy as np
import pytest

import dage_amd
import oracle as O
from _helpers import random_gaussians

pytestmark = pytest.mark.gpu


def _job(seed, reps, results, key, errors):
    try:
        ctx = dage_amd.Context(0)
        ax = O.Axis.ncell_fixed(-6.0, 6.0, 7, 8 + seed)  # 49^3 .. 61^3: distinct shapes
        g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
        q = dage_amd.GaussQuad(nlin=20, nlog=20, tlin=2.0, tlog=500.0)
        op = dage_amd.Helmholtz3D(ctx, g, g, q, energy=-0.3 - 0.2 * seed)
        rng = np.random.default_rng(300 + seed)
        rho = O.gauss_density(*random_gaussians(rng, 4), ax.points, ax.points, ax.points)
        a = dage_amd.DeviceCube(ctx, rho.shape, rho)
        b = dage_amd.DeviceCube(ctx, rho.shape)
        out = None
        for _ in range(reps):
            op.transform_cuda_cube(a, b)
            # an error of this thread's own between the applies
            try:
                dage_amd.Bubbles(ctx, 2, 13 + seed)
                errors.append(f"{key}: lmax {13 + seed} accepted")
            except dage_amd.DageError as e:
                if f"lmax {13 + seed}" not in str(e):
                    errors.append(f"{key}: read another thread's error: {e}")
            out = b.download()
        results[key] = out
        for obj in (a, b, op, g, ctx):
            obj.close()
    except Exception as e:  # reported by the main thread
        errors.append(f"{key}: {e!r}")


def _bad_args(messages, errors):
    try:
        for _ in range(20):
            with pytest.raises(dage_amd.DageError) as e:
                dage_amd.GaussQuad(nlin=4, nlog=4, tlin=2.0, tlog=1.0)  # tlog <= tlin
            messages.append(str(e.value))
    except Exception as e:
        errors.append(f"bad-args thread: {e!r}")


def test_three_host_threads_distinct_contexts():
    serial, errors = {}, []
    for s in range(3):
        _job(s, 1, serial, s, errors)
    assert not errors, errors
    conc, messages = {}, []
    threads = [threading.Thread(target=_job, args=(s, 4, conc, s, errors)) for s in range(3)]
    threads.append(threading.Thread(target=_bad_args, args=(messages, errors)))
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
        assert not t.is_alive(), "a host thread did not finish"
    assert not errors, errors
    for s in range(3):
        print(f"thread {s}: shape {conc[s].shape}, "
              f"{np.count_nonzero(conc[s] != serial[s])} points differ from the serial run")
        assert np.array_equal(conc[s], serial[s])
    assert len(messages) == 20
    # every error message read on the bad-args thread is its own
    assert all("tlog" in m or "interval" in m or "tlin" in m for m in messages), messages[:3]
