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
