"""Field files written by scft! and cells read from field files, on the GPU.

* scft!(...; opts = SCFTOptions(nsave, savepath)) writes savefields(sys, savepath) every
  nsave steps (src/simulate.jl:124-128).  The library calls the host's save function at
  exactly those points (pf_scft_save) and once after the final residuals!.
* fieldinit!(sys, fpath; load_cell = true) installs the file's cell with add_cell!, crystal
  system included (src/fileio.jl:56-64, src/system/system.jl:102-106).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import diblock_system

pytestmark = pytest.mark.gpu

P = pytest.importorskip("polyfields_hip")


def _lam(mod, **kw):
    return diblock_system(mod, (64, 1, 1), 1, "lamellar", (15.0,), chi=0.2, **kw)


def test_scft_save_points_and_saved_fields():
    from polyfields_hip._lib import SaveFn, ScftOptions, ScftResults, call
    rng = np.random.default_rng(8)
    fields = {m: 0.3 * rng.standard_normal((1, 1, 64)) for m in (1, 2)}
    sg, _, _ = _lam(P, precision="fp64")
    P.fieldinit(sg, fields=fields)
    u = P.AndersonUpdater(nhist=10)
    seen, snap = [], {}

    def save(_user, step):
        seen.append(step)
        snap[step] = {m: np.array(sg.fields[m]) for m in (1, 2)}

    cb = SaveFn(save)
    o, r = ScftOptions(7, 0, 5, 1e-12, 1e-4), ScftResults()
    call("pf_scft_save", sg._h, u._h, None, C.byref(o), C.byref(r), cb, None)
    assert not r.converged and r.nsteps == 7
    assert seen == [0, 5, -1]
    # the step-5 save holds the fields as the update of step 5 left them: the oracle's
    # scft! stopped after that step, up to its final residuals!, which only subtracts
    # the fields' means (src/system/system.jl:332-337)
    so, _, _ = _lam(O)
    O.fieldinit(so, fields=fields)
    O.scft(so, O.AndersonUpdater(nhist=10), opts=O.SCFTOptions(nsteps=6, ftol=1e-12))
    for m in (1, 2):
        w = snap[5][m] - snap[5][m].mean()
        np.testing.assert_allclose(w, so.fields[m], rtol=0, atol=1e-9)
        np.testing.assert_array_equal(snap[-1][m], np.array(sg.fields[m]))


def test_scft_savepath_writes_the_final_fields(tmp_path):
    path = str(tmp_path / "fields.json")
    sg, _, _ = _lam(P, precision="fp64")
    P.fieldinit(sg, seed=3, scale=0.1)
    res = P.scft(sg, P.AndersonUpdater(nhist=10),
                 opts=P.SCFTOptions(nsteps=40, nsave=10, ftol=1e-6, savepath=path))
    d = P.loadfields(path)
    assert d["crystal"] == "lamellar" and d["grid"] == (64, 1, 1)
    assert d["cell_params"] == pytest.approx([15.0])
    for m in (1, 2):
        np.testing.assert_array_equal(d["fields"][m], np.array(sg.fields[m]))
    assert res.nsteps <= 40


def test_fieldinit_file_installs_the_files_cell(tmp_path):
    # a field file of a hexagonal cell read into a system built with a square cell
    rng = np.random.default_rng(12)
    fields = {m: 0.3 * rng.standard_normal((1, 16, 16)) for m in (1, 2)}
    src, _, _ = diblock_system(P, (16, 16, 1), 2, "hexagonal", (7.5,), chi=0.25, f=0.3,
                               precision="fp64")
    P.fieldinit(src, fields=fields)
    path = str(tmp_path / "hex.json")
    P.savefields(src, path)
    sg, _, _ = diblock_system(P, (16, 16, 1), 2, "square", (5.0,), chi=0.25, f=0.3,
                              precision="fp64")
    P.fieldinit(sg, path)
    assert sg.cell.crystal == "hexagonal" and sg.cell.params == pytest.approx([7.5])
    so, _, _ = diblock_system(O, (16, 16, 1), 2, "hexagonal", (7.5,), chi=0.25, f=0.3)
    O.fieldinit(so, fields=fields)
    assert sg.free_energy() == pytest.approx(so.free_energy(), rel=1e-12)
    for m in (1, 2):
        np.testing.assert_allclose(np.array(sg.density[m]), so.density[m], rtol=0, atol=1e-12)
    assert sg.scfstress() == pytest.approx(so.scfstress(), rel=1e-10)
    # a cell of another dimension is refused, as add_cell! refuses it (cell.jl:72-76)
    with pytest.raises(P.PolyFieldsError):
        sg.add_cell(P.UnitCell(1, "lamellar", 5.0))
