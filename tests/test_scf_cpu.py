"""The SCF convergence rule (SCFOptimizer_is_converged, scf.F90:99-136) and
the config-4 molecule generator, on the CPU."""
import numpy as np

from dage_amd import scf
from dage_amd.scf import is_converged, random_molecule


def test_convergence_rule():
    assert not is_converged([1.0], [[0.1]])  # iteration 1 never converges
    assert is_converged([1.0, 1.0 + 5e-7], [[0.1, 0.2], [0.1 + 1e-7, 0.2 - 9e-7]])
    assert not is_converged([1.0, 1.0 + 2e-6], [[0.1], [0.1]])
    assert not is_converged([1.0, 1.0], [[0.1, 0.2], [0.1, 0.2 + 2e-6]])


def test_molecule_generator():
    rng = np.random.default_rng(0xDA6E0004)
    m = random_molecule(rng)
    assert len(m.Z) == 10 and set(m.Z) <= {1.0, 6.0, 7.0, 8.0}
    assert np.all(np.linalg.norm(m.R, axis=1) <= 6.0)
    d = np.linalg.norm(m.R[:, None] - m.R[None], axis=-1)
    assert np.min(d + 1e9 * np.eye(10)) >= 1.8


def test_backoff_schedule():
    """Learned proposals: after 3 losses in a row, one evaluation every 8
    iterations until a win resets the count."""
    def run(outcomes, backoff=(3, 8)):
        rejected, marks = 0, []
        for win in outcomes:
            if scf.backoff_skips(rejected, backoff):
                rejected += 1
                marks.append("-")
            else:
                marks.append("W" if win else "L")
                rejected = 0 if win else rejected + 1
        return "".join(marks)

    assert run([False] * 20) == "LLLL-------L-------L"
    assert run([False] * 4 + [True] * 9 + [False] * 3) == "LLLL-------WWLLL"
    assert run([False] * 10, backoff=None) == "L" * 10
