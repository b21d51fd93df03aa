"""CPU checks of the HIP library: it loads, exports every symbol the header
declares, and its host-side functions (bin segments, the k-table least-squares
map, the synthetic generator, argument errors) agree with the oracle.  No
device calls are made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    from heliosk import _lib
    return _lib.lib()


def test_exports_every_declared_symbol(L):
    from heliosk import _lib
    hdr = open(os.path.join(ROOT, "include", "heliosk.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hk_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.hk_version()


def test_bin_segments_match_oracle():
    from heliosk.engine import bin_segments, make_grid
    for numin, numax, dnu, nbins in ((0.0, 30000.0, 0.01, 300), (0.0, 30000.0, 0.001, 300),
                                     (24400.0, 24500.0, 0.001, 5), (0.3, 1000.3, 0.1, 7),
                                     (0.0, 40000.0, 0.01, 400), (0.0, 10000.0, 0.01, 100)):
        g = make_grid(numin, numax, dnu)
        np.testing.assert_array_equal(bin_segments(g, numax, nbins),
                                      O.bin_segments(numin, numax, dnu, g.nx, nbins))


def _chebyshev_moments(b, nC):
    Nxb = b.size
    x = np.arange(Nxb) * 2.0 / (Nxb - 1.0) - 1.0
    V = np.polynomial.chebyshev.chebvander(x, nC - 1)
    return V.T @ b


# the moment form is used for Nxb >= 25 nC (shorter bins take the literal
# Householder path on the GPU)
@pytest.mark.parametrize("Nxb,nC", [(500, 20), (10000, 20), (30000, 20), (175, 7), (800, 32)])
def test_ktable_plan_reproduces_reference_solve(Nxb, nC):
    from heliosk.engine import ktable_plan
    P = ktable_plan(Nxb, nC)
    rng = np.random.default_rng(Nxb)
    for trial in range(3):
        b = rng.normal(-30.0, 5.0, Nxb) + np.linspace(0, 20, Nxb) ** 1.5 * (trial + 1) / 20
        c_ref = O.lsq(b, nC)
        c = P[:, :nC] @ _chebyshev_moments(b, nC) + P[:, nC:] @ b[:nC]
        assert np.max(np.abs(c - c_ref)) <= 1e-11 * np.max(np.abs(c_ref))


def test_ktable_plan_on_p005_fixture(golden_dir):
    from heliosk.engine import ktable_plan
    Ks = np.loadtxt(os.path.join(golden_dir, "p005_Out_i_bin.dat"), usecols=(1,))
    ref = np.loadtxt(os.path.join(golden_dir, "p005_Out_r_cbin.dat"))
    Nxb, n = 500, 426
    k = np.arange(Nxb, dtype=np.float64)
    ii = (1.0 - n / (Nxb - 1.0)) * k + n
    ii = np.where(ii >= Nxb - 1, 0.999999 * (Nxb - 1), ii)
    il = (ii / Nxb * Nxb).astype(int)
    b = np.log(Ks[il] + (Ks[il + 1] - Ks[il]) * (ii - il))
    P = ktable_plan(Nxb, 20)
    c = P[:, :20] @ _chebyshev_moments(b, 20) + P[:, 20:] @ b[:20]
    np.testing.assert_allclose(c, ref[2:], rtol=1e-9, atol=1e-12)


def test_synth_lines_deterministic_and_distributed():
    from heliosk.engine import synth_lines
    a = synth_lines(1234, 0, 200_000, 0)
    b = synth_lines(1234, 0, 200_000, 0)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    assert 0.0 <= a["nu"].min() and a["nu"].max() < 30000.0
    lg = np.log10(a["S"])
    assert -30.0 <= lg.min() and lg.max() <= -19.0
    assert 0.01 <= a["gamma"].min() and a["gamma"].max() <= 0.1
    assert 0.3 <= a["n"].min() and a["n"].max() <= 0.8
    assert abs(np.mean(a["nu"]) - 15000.0) < 100.0
    c = synth_lines(1234, 1, 200_000, 0)
    assert not np.array_equal(a["nu"], c["nu"])          # species are independent streams
    h = synth_lines(99, 0, 200_000, 1)
    assert 1e-4 <= h["gamma"].min() and h["gamma"].max() <= 1.0
    # 5 % strong lines with log10 S in [-20, -17]: two thirds of them lie above -19
    strong = np.mean(np.log10(h["S"]) > -19.0)
    assert 0.030 < strong < 0.037


def test_argument_errors_are_reported(L):
    from heliosk import _lib
    st = _lib.LineState()
    rc = L.hk_line_state_init(ctypes.byref(st), 1000, ctypes.c_void_p(0x1000), 10)
    assert rc == 1 and b"bytes given" in L.hk_last_error()
    g = _lib.Grid(0.0, 1.0, 1001)
    rc = L.hk_ktable_prepare(ctypes.byref(g), 1001.0, 10, 20, ctypes.c_void_p(0x1000), 1 << 20)
    assert rc == 1 and b"not divisible" in L.hk_last_error()
    rc = L.hk_ktable_plan_host(10, 40, (ctypes.c_double * 800)())
    assert rc == 1
