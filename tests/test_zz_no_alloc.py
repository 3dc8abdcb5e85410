"""pf_scfstress_noalloc: the reference stress (src/system/system.jl:384-395,
src/species/multiblock.jl:233-295) with every device buffer carved from the library's
fixed 32 MiB scratch array instead of allocated, against pf_scfstress on the same states:
two systems in a row, the second smaller, then the first again; and the size check.
"""
import numpy as np
import pytest

from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def _noalloc(sys_):
    from polyfields_hip._lib import call, dptr
    out = np.zeros(len(sys_.cell.params))
    call("pf_scfstress_noalloc", sys_._h, dptr(out), len(out))
    return out


def _state(dims, dim, crystal, params, precision, Ns=40, seed=3):
    s, cell, _ = diblock_system(P, dims, dim, crystal, params, f=0.35, chi=0.2, Ns=Ns,
                                precision=precision)
    rng = np.random.default_rng(seed)
    P.fieldinit(s, fields={m: 0.5 * rng.standard_normal(s.shape) for m in (1, 2)})
    s.density_update()
    return s


def test_noalloc_stress_equals_the_allocating_entry_point():
    big = _state((48, 36, 32), 3, "orthorhombic", (6.0, 5.5, 5.0), "fp64")
    small = _state((16, 12, 1), 2, "rectangular", (5.0, 4.0), "fp64", Ns=30)
    big32 = _state((64, 64, 64), 3, "cubic", (8.0,), "fp32", Ns=20)
    for s in (big, small, big32, big):
        ref = np.array(s.scfstress())
        got = _noalloc(s)
        assert np.all(np.isfinite(got)) and np.abs(ref).max() > 1e-8
        np.testing.assert_array_equal(got, ref)


def test_noalloc_stress_after_a_cell_change():
    s = _state((32, 32, 32), 3, "cubic", (6.0,), "fp64")
    s.cell.params = [6.3]
    s.density_update()
    np.testing.assert_array_equal(_noalloc(s), np.array(s.scfstress()))


def test_noalloc_stress_refuses_what_does_not_fit():
    # 128^3 FP64: one row's two half spectra take 34 MB, beyond the 32 MiB scratch
    s = _state((128, 128, 128), 3, "cubic", (12.0,), "fp64", Ns=4)
    with pytest.raises(P.PolyFieldsError, match="bytes of scratch"):
        _noalloc(s)
    # a system whose propagator tables do not exist yet
    t, _, _ = diblock_system(P, (16, 16, 16), 3, "cubic", (5.0,), precision="fp64")
    with pytest.raises(P.PolyFieldsError, match="density!"):
        _noalloc(t)
