"""scft! with Anderson mixing queues the next step's residuals! before it reads the step's
errors (src/simulate.jl:97-121 evaluates them after every step).  Once scft! has returned,
a residuals! of changed fields must be computed afresh, whether the loop ended by
convergence (the queued pass became the final residuals!) or by the step limit."""
import numpy as np
import pytest

from helpers import diblock_system

P = pytest.importorskip("polyfields_hip")
pytestmark = pytest.mark.gpu


def _system():
    return diblock_system(P, (32, 32, 1), 2, "hexagonal", (4.5,), N=100, f=0.3, chi=0.14, Ns=64,
                          precision="fp64")[0]


def _set(sys, fields):
    for m, f in fields.items():
        sys.fields[m] = f


@pytest.mark.parametrize("nsteps,ftol", [(200, 1e-6), (4, 1e-14)], ids=["converged", "step-limit"])
def test_residuals_after_scft_are_recomputed(nsteps, ftol):
    s = _system()
    P.fieldinit(s, seed=7, scale=0.1)
    r = P.scft(s, P.AndersonUpdater(nhist=5), opts=P.SCFTOptions(nsteps=nsteps, ftol=ftol))
    assert r.converged == (nsteps == 200)
    rng = np.random.default_rng(11)
    new = {m: 0.2 * rng.standard_normal(s.shape) for m in (1, 2)}
    _set(s, new)
    s.residuals_update()
    t = _system()
    _set(t, new)
    t.residuals_update()
    assert s.field_error() == t.field_error()
    assert s.free_energy() == t.free_energy()
    for m in (1, 2):
        np.testing.assert_array_equal(np.asarray(s.density[m]), np.asarray(t.density[m]))
