"""The C ABI without a GPU: the library loads, exports every symbol the header declares,
and its host-only geometry helpers agree with the oracle and the reference's KATs."""
import math
import os
import re

import numpy as np
import pytest

import oracle as O
from oracle.cell import NPARAMS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "polyfields_hip.h")


def _lib():
    import polyfields_hip
    try:
        return polyfields_hip.lib()
    except RuntimeError as e:
        pytest.skip(f"library not built: {e}")


def test_header_symbols_exported():
    import ctypes
    lib = _lib()
    text = open(HEADER).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(pf_\w+)\(", text, re.M))
    assert len(names) >= 40
    for n in sorted(names):
        assert hasattr(lib, n), n
    # every declared prototype is bound by the Python mirror
    from polyfields_hip._lib import SIGNATURES
    assert names == set(SIGNATURES)


@pytest.mark.parametrize("dim,crystal", sorted(NPARAMS))
def test_cell_geometry_matches_oracle(dim, crystal):
    import polyfields_hip as P
    _lib()
    params = {1: [7.0], 2: [6.0, 4.5, 1.2], 3: [6.0, 5.0, 4.0, 1.3, 1.4, 1.2]}[dim]
    if crystal == "monoclinic":
        params = [6.0, 5.0, 4.0, 1.3]
    params = params[:NPARAMS[(dim, crystal)]]
    c = P.UnitCell(dim, crystal, *params)
    o = O.UnitCell(dim, crystal, *params)
    assert c.volume == pytest.approx(o.volume, rel=1e-14)
    np.testing.assert_allclose(c.R, o.R, atol=1e-13)
    np.testing.assert_allclose(c.G, o.G, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(c.dGG, o.dGG, rtol=1e-12, atol=1e-13)


def test_cell_volume_kats():
    import polyfields_hip as P
    _lib()
    rtol = math.sqrt(np.finfo(float).eps)
    assert P.UnitCell(2, "hexagonal", 10.0).volume == pytest.approx(86.60254037844385, rel=rtol)
    assert P.UnitCell(3, "monoclinic", 10.0, 5.0, 2.0, math.pi / 4).volume == pytest.approx(
        70.71067811865474, rel=rtol)
    assert P.UnitCell(2, "oblique", 10.0, 5.0, math.pi / 4).volume == pytest.approx(
        35.35533905932737, rel=rtol)


@pytest.mark.parametrize("N_block,Ns", [([50, 50], 100), ([35, 65], 100), ([60, 80, 60], 200),
                                        ([1, 99], 10), ([30.0, 40.0, 30.0], 64)])
def test_chain_grid_matches_oracle(N_block, Ns):
    import polyfields_hip as P
    _lib()
    assert P.chain_grid(N_block, Ns) == pytest.approx(O.chain_grid(N_block, Ns))
    N, b, d = P.chain_grid(N_block, Ns)
    No, bo, do = O.chain_grid(N_block, Ns)
    assert b == bo and N == No


def test_errors_cross_the_abi_as_exceptions():
    import polyfields_hip as P
    _lib()
    with pytest.raises(P.PolyFieldsError, match="Unrecognized crystal"):
        P.UnitCell(2, "cubic", 4.0)
    with pytest.raises(ValueError):
        P.Diblock(P.Monomer(1), P.Monomer(2), 100, 0.29, 1.0, 1.0)


@pytest.mark.parametrize("seed", [0, 1, 12345, 2 ** 40 + 7, 2 ** 64 - 1])
def test_julia_stream_matches_oracle(seed):
    # pf_julia_randn (the library's host-side restatement of Julia 1.5's seeded randn,
    # behind fieldinit!(sys; seed, scale)) against the oracle's, bit for bit, over array
    # lengths on both sides of the 13-element bulk threshold and long enough to reach the
    # ziggurat's slow paths and a second dSFMT block
    import ctypes as C
    from oracle import julia_rng as J
    from polyfields_hip._lib import call
    _lib()
    counts = [1, 12, 13, 32, 1000, 5000, 7]
    arr = (C.c_int64 * len(counts))(*counts)
    out = np.zeros(sum(counts))
    call("pf_julia_randn", C.c_uint64(seed), arr, len(counts),
         out.ctypes.data_as(C.POINTER(C.c_double)))
    g = J.seeded(seed)
    ref = np.concatenate([np.array(J.randn_array(g, c)) for c in counts])
    assert np.array_equal(out, ref)
