"""Edge cases of the batched scft! loop (pf_batch_scft*, no counterpart in the reference
beyond scft! itself, src/simulate.jl:60-148): a step limit of zero, a batch whose every
system is masked, and a system that is converged before the first step, each against
what scft! does for the same system one at a time."""
import numpy as np
import pytest

from helpers import diblock_system

P = pytest.importorskip("polyfields_hip")
pytestmark = pytest.mark.gpu


def _lam(L, seed):
    s = diblock_system(P, (64, 1, 1), 1, "lamellar", (L,), N=100, f=0.5, chi=0.2, Ns=100,
                       precision="fp64")[0]
    P.fieldinit(s, seed=seed, scale=0.1)
    return s


UPDATERS = {
    "anderson": lambda: P.AndersonUpdater(nhist=5),
    "mom": lambda: P.MomentumUpdater(method="MOM", lam=0.05),
    "nag": lambda: P.MomentumUpdater(method="NAG", lam=0.05),
    "pece": lambda: P.EulerUpdater(method="PECE", lam=0.05),
}


@pytest.mark.parametrize("name", list(UPDATERS))
def test_zero_steps_matches_one_at_a_time(name):
    # nsteps = 0: no update; both report the first residuals' errors, and the fields are
    # those residuals! leaves (it subtracts their mean, src/system/system.jl:317-340)
    single = [_lam(L, 10 + k) for k, L in enumerate((9.0, 11.0))]
    ref = [P.scft(s, UPDATERS[name](), opts=P.SCFTOptions(nsteps=0, ftol=1e-5)) for s in single]
    batch_sys = [_lam(L, 10 + k) for k, L in enumerate((9.0, 11.0))]
    res = P.Batch(batch_sys).scft(updater=UPDATERS[name](), nsteps=0, ftol=1e-5)
    for k, (r, q) in enumerate(zip(res, ref)):
        assert r.nsteps == q.nsteps == 0
        assert r.converged == q.converged
        assert r.fhelm == pytest.approx(q.fhelm, rel=1e-12)
        assert r.ferr == pytest.approx(q.ferr, rel=1e-10)
        for m in (1, 2):
            np.testing.assert_allclose(np.asarray(batch_sys[k].fields[m]),
                                       np.asarray(single[k].fields[m]), rtol=0, atol=1e-14)


def test_all_masked_batch_leaves_every_system_alone():
    systems = [_lam(L, 20 + k) for k, L in enumerate((9.0, 10.0, 11.0))]
    before = [np.asarray(s.fields[1]).copy() for s in systems]   # no pass touches them
    b = P.Batch(systems)
    for upd in (None, UPDATERS["mom"](), UPDATERS["pece"]()):
        res = b.scft(updater=upd, nsteps=20, ftol=1e-5, mask=[0, 0, 0])
        for r in res:
            assert np.isnan(r.ferr) and not r.converged and r.nsteps == 0
    for s, w in zip(systems, before):
        np.testing.assert_array_equal(np.asarray(s.fields[1]), w)


@pytest.mark.parametrize("name", ["anderson", "mom", "pece"])
def test_already_converged_system_beside_a_running_one(name):
    # a system that meets ftol before the first step stops at nsteps = 0 and keeps its
    # (converged, mean-free) fields while the other system of the batch runs on
    def make():
        a = _lam(10.0, 30)
        P.scft(a, P.AndersonUpdater(nhist=10), opts=P.SCFTOptions(nsteps=400, ftol=1e-8))
        return [a, _lam(11.0, 31)]

    single = make()
    conv_fields = np.asarray(single[0].fields[1]).copy()
    ref = [P.scft(s, UPDATERS[name](), opts=P.SCFTOptions(nsteps=30, ftol=1e-6))
           for s in single]
    batch_sys = make()
    res = P.Batch(batch_sys).scft(updater=UPDATERS[name](), nsteps=30, ftol=1e-6)
    assert ref[0].converged and ref[0].nsteps == 0
    assert res[0].converged and res[0].nsteps == 0
    np.testing.assert_allclose(np.asarray(batch_sys[0].fields[1]), conv_fields, rtol=0,
                               atol=1e-12)
    assert res[1].nsteps == ref[1].nsteps
    assert res[1].fhelm == pytest.approx(ref[1].fhelm, rel=1e-12)
