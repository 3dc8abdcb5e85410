"""The GPU batch on sweep instances the reference algorithm fails on (see
tests/test_oracle_sweep_failures.py): the first, pure-Euler phase of the sweep protocol
drops instances 94 and 235 as diverged (the propagator's Q turns negative, where the
reference's muphi! raises a DomainError), as the oracle fails on them, while a solvable
instance of the same batch converges to the oracle's state.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402
from test_oracle_sweep_failures import _instance, _phase1  # noqa: E402


def test_batch_drops_the_instances_the_reference_fails_on():
    cfg = bench.sweep_instances(256)
    idx = [94, 235, 56]
    systems = [bench.sweep_system(P, cfg[i], "fp64") for i in idx]
    res = P.Batch(systems).scft(**bench.PH1)
    for k in (0, 1):
        assert not res[k].converged and np.isnan(res[k].ferr), idx[k]
        with pytest.raises((RuntimeError, ValueError), match="Diverging|DomainError"):
            _phase1(_instance(idx[k]))
    # the solvable instance reaches the phase's tolerance on both, in the same steps
    ro = _phase1(_instance(56))
    assert res[2].converged and ro.converged
    assert res[2].nsteps == ro.nsteps
    assert res[2].fhelm == pytest.approx(ro.fhelm, rel=1e-10)


def test_plain_scft_raises_where_the_reference_raises():
    s = bench.sweep_system(P, bench.sweep_instances(256)[94], "fp64")
    p = bench.PH1
    with pytest.raises(P.PolyFieldsError, match="DomainError|Diverging"):
        P.scft(s, P.AndersonUpdater(nhist=p["nhist"], iter_euler=p["iter_euler"],
                                    lam_euler=p["lam_euler"]),
               opts=P.SCFTOptions(nsteps=p["nsteps"], ftol=p["ftol"]))
