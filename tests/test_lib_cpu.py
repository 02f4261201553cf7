"""Host-side checks of libravenhip (no GPU needed): the C ABI loads and exports
every declared symbol, and its host-side grid generation and LobattoCell
operators agree bit for bit with the oracle and the 256-bit fixtures."""
import json
import os
import re

import numpy as np
import pytest

import oracle as O
import problems as PB
import ravenhip as R
from ravenhip import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as fh:
                text = fh.read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(rv_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_symbols()
    assert "rv_rhs" in names and "rv_lsrk_stage" in names and "rv_create" in names
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_error_path_without_device_or_bad_descriptor():
    d = L.RvDesc()
    d.abi_version = 999
    h = L.H()
    rc = L.lib().rv_create(L.ctypes.byref(d), L.ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"ABI" in L.lib().rv_last_error(None)


@pytest.mark.parametrize("nq", range(2, 17))
def test_lobatto_matches_256bit_fixture(nq):
    with open(os.path.join(ROOT, "tests", "golden", "gll.json")) as fh:
        cell = json.load(fh)["cells"][str(nq)]
    c = R.LobattoCell(nq)
    assert np.array_equal(c.points_1d, [float.fromhex(v) for v in cell["points_f64"]])
    assert np.array_equal(c.weights_1d, [float.fromhex(v) for v in cell["weights_f64"]])
    fD = np.array([[float.fromhex(v) for v in row] for row in cell["D_f64"]])
    assert np.array_equal(c.derivatives_1d(), fD)


CASES = [
    (2, 5, [PB.julia_range(-1, 1, 6)] * 2, (True, True), 0),
    (2, 4, [PB.julia_range(0, 20, 4), PB.julia_range(-5, 5, 3)], (True, False), 0),
    (3, 3, [PB.julia_range(-1, 1, 4), PB.julia_range(0, 2, 3), PB.julia_range(0, 10, 5)],
     (True, False, True), 0),
    (3, 4, [PB.julia_range(0, 10, 3)] * 3, (True, True, True), 1),
    (3, 8, [PB.julia_range(0, 10, 5)] * 3, (True, True, True), 0),
    (1, 5, [PB.julia_range(-1, 1, 6)], (True,), 0),
    (2, 6, [PB.julia_range(0, 1, 4)] * 2, (False, False), 1),
]


@pytest.mark.parametrize("dim,nq,coords,per,level", CASES)
def test_brick_arrays_bitexact_against_oracle(dim, nq, coords, per, level):
    cell = R.LobattoCell(*([nq] * dim))
    g = R.generate(R.GridManager(cell, R.brick(coords, per), min_level=level))
    o = O.BrickGrid(nq, coords, per, level=level)
    assert g.K == o.K
    assert np.array_equal(g.points.reshape(-1), o.points)
    assert np.array_equal(g.metrics.reshape(-1), o.metrics)
    assert np.array_equal(g.surface.reshape(-1), o.surf)
    assert np.array_equal(g.vmapM.reshape(-1), o.vmapM)
    assert np.array_equal(g.vmapP.reshape(-1), o.vmapP)
    assert np.array_equal(g.boundarycodes.reshape(-1), o.bc)
    assert np.array_equal(g.dtoc.reshape(-1), o.dtoc)
    assert g.ncontinuous == o.nc


def test_nonuniform_brick_is_rejected():
    cell = R.LobattoCell(4, 4)
    with pytest.raises(NotImplementedError):
        R.generate(R.GridManager(cell, R.brick([[0.0, 0.3, 1.0], [0.0, 1.0]], (True, True))))


@pytest.mark.parametrize("scheme,nst,order", [(54, 5, 4), (144, 14, 4)])
def test_lsrk_coefficients_of_the_library(scheme, nst, order):
    """The library's tables equal the oracle's bit for bit, and a 2N-storage
    run of y' = y with them converges at the scheme's order."""
    from ravenhip.odesolvers import LSRK

    class _Dummy:
        def GridArray(self, fill):
            return None

    a, b, c = LSRK(_Dummy(), None, 0.1, 0.0, scheme).coefficients()
    oa, ob, oc = O.lsrk_coefficients(scheme)
    assert len(a) == nst
    assert np.array_equal(a, oa) and np.array_equal(b, ob) and np.array_equal(c, oc)

    def run(h):
        y, dy = 1.0, 0.0
        for _ in range(round(1 / h)):
            for s in range(nst):
                dy = a[s] * dy + y
                y += b[s] * h * dy
        return abs(y - np.e)

    e1, e2 = run(1 / 8), run(1 / 16)
    assert np.log2(e1 / e2) > order - 0.3


def test_flux_differencing_kernel_types():
    """FluxDifferencingForm(flux, kernel_type) takes Raven's three kernel types
    (dgsem.jl:8-17, 163-255) and rejects others."""
    import ravenhip as R
    for kt in ("naive", ":per_dir", "per_dir_symmetric"):
        assert R.FluxDifferencingForm(R.CentralFlux(), kt).kernel_type == kt.lstrip(":")
    assert R.FluxDifferencingForm(R.CentralFlux()).kernel_type == "per_dir"
    with pytest.raises(ValueError):
        R.FluxDifferencingForm(R.CentralFlux(), "per_cell")
