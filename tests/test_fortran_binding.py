"""The drop-in boundary from Fortran: dage_amd/fortran/dage_ops_mod.F90
(ISO_C_BINDING interfaces + an Operator3D-shaped wrapper) compiled with
amdflang against libdage_ops.so.  CPU: the host-side quadrature through the
Fortran binding equals the oracle's bit for bit.  GPU: the reference's
gauss_test.F90 TEST 1 self-energy rebuilt entirely from Fortran."""
import os
import re
import subprocess

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FCHECK = os.path.join(ROOT, "dage_amd", "fortran", "_build", "fortran_apply_check")

needs_binary = pytest.mark.skipif(not os.path.exists(FCHECK),
                                  reason="Fortran check not built (amdflang missing at build time)")


def _run(*args):
    out = subprocess.run([FCHECK, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


@needs_binary
def test_fortran_quadrature_through_c_abi():
    text = _run("quad")
    m = re.search(r"tp\(1\)=\s*(\S+)\s+tw\(18\)=\s*(\S+)", text)
    tp, tw = O.gauss_quad(10, 8, 0.0, 2.0, 500.0)
    assert float(m.group(1)) == tp[0] and float(m.group(2)) == tw[17]


@needs_binary
@pytest.mark.gpu
def test_fortran_drop_in_reproduces_reference_golden():
    text = _run()
    e = float(re.search(r"selfint=\s*(\S+)", text).group(1))
    ref = 0.7979283927253593  # tests/gauss_test.F90:89-96 of the reference
    assert abs(e - ref) < 1e-13 * ref, (e, ref)


@needs_binary
@pytest.mark.gpu
def test_fortran_gbfmm_drop_in(ctx):
    """GBFMMCoulomb3D from Fortran (dage_gbfmm_create / _apply / _box_count):
    the alpha = 1 Gaussian's self-interaction with the FMM potential on a 49^3
    grid, 4x4x4 boxes, lmax 6 -- equal to the same call through the Python
    mirror, and within the quadrature error of the analytic sqrt(2/pi)."""
    import math

    import numpy as np

    import dage_amd
    text = _run("gbfmm")
    m = re.search(r"gbfmm_selfint=\s*(\S+)\s+boxes=(\d+)", text)
    e_f, nbox = float(m.group(1)), int(m.group(2))
    assert nbox == 64
    ax = O.Axis(-6.0, np.full(8, 0.25), 7)
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 7)
    gq = dage_amd.GaussQuad(nlin=10, nlog=8, tlin=2.0, tlog=500.0)
    X, Y, Z = np.meshgrid(ax.points, ax.points, ax.points, indexing="ij")
    rho = np.asfortranarray(math.pi ** -1.5 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2)))
    fmm = dage_amd.GBFMMCoulomb3D(ctx, g, maxlevel=2, lmax=6, gauss=gq)
    pot = fmm.apply(rho)
    w = ax.ints()
    e_py = float(np.einsum("i,j,k,ijk->", w, w, w, rho * pot))
    print(f"Fortran GBFMM self-interaction {e_f:.15f}, Python {e_py:.15f}, "
          f"analytic {math.sqrt(2 / math.pi):.15f}")
    assert abs(e_f - e_py) < 1e-13 * abs(e_py)
    assert abs(e_f - math.sqrt(2 / math.pi)) < 1e-3
    fmm.close()
    g.close()
