"""hk_voigt_accumulate_noalloc (static device scratch, no allocation) against
hk_voigt_accumulate (caller workspace): both run the same kernels, so the
spectra must agree bit for bit -- on two calls in a row, the second on a
smaller grid and list, in both forms, and with the accumulate flag."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import random_lines, species_from

pytestmark = pytest.mark.gpu


def _noalloc(L, E, st, grid, mode, K, kmin=0.0, accumulate=False, tile_points=0):
    tune = E._tune(tile_points, False)
    return L.hk_voigt_accumulate_noalloc(ctypes.byref(st.st), ctypes.byref(grid), float(kmin),
                                         int(accumulate), int(mode), ctypes.byref(tune),
                                         E._ptr(K), E._stream(None))


def test_noalloc_matches_workspace_form():
    from heliosk import engine as E
    from heliosk._lib import HK_OK, lib
    L = lib()
    cases = [  # (lines, numin, numax, dnu, cut, tile_points): the second is smaller
        (6000, 1000.0, 1100.0, 0.002, 10.0, 0),
        (900, 1020.0, 1040.0, 0.004, 3.0, 2048),
    ]
    for mode in (E.HK_VOIGT_FP32C, E.HK_VOIGT_FP64):
        for i, (n, a, b, dnu, cut, tp) in enumerate(cases):
            d = random_lines(n, seed=11 + i, numin=a - 5.0, numax=b + 5.0)
            sp = species_from(d)
            st = E.line_params(sp, 1200.0 + 300 * i, 0.5, Q=list(d["iso_Q"]), cut_mode=0, cut=cut)
            grid = E.make_grid(a, b, dnu)
            Kref = E.voigt(st, grid, 1e-30, mode, tile_points=tp)
            K = torch.full((grid.nx,), float("nan"), dtype=torch.float64, device="cuda")
            rc = _noalloc(L, E, st, grid, mode, K, kmin=1e-30, tile_points=tp)
            assert rc == HK_OK, L.hk_last_error()
            torch.cuda.synchronize()
            assert torch.equal(K, Kref), (mode, i)
            # accumulate onto the result: twice the sum (minus one kmin)
            Kacc = E.voigt(st, grid, 0.0, mode, tile_points=tp, K=Kref.clone(), accumulate=True)
            rc = _noalloc(L, E, st, grid, mode, K, accumulate=True, tile_points=tp)
            assert rc == HK_OK, L.hk_last_error()
            torch.cuda.synchronize()
            assert torch.equal(K, Kacc), (mode, i)


def test_noalloc_rejects_grid_beyond_limit():
    from heliosk import engine as E
    from heliosk._lib import HK_ERR_ARG, Grid, lib
    L = lib()
    d = random_lines(10, seed=3)
    st = E.line_params(species_from(d), 1000.0, 1.0, Q=list(d["iso_Q"]), cut_mode=0, cut=5.0)
    grid = Grid(0.0, 1e-6, (1 << 25) + 1)
    K = torch.zeros(4, dtype=torch.float64, device="cuda")
    rc = _noalloc(L, E, st, grid, E.HK_VOIGT_FP32C, K)
    assert rc == HK_ERR_ARG
    assert b"static scratch limit" in L.hk_last_error()
