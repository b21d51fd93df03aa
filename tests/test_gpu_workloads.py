"""The batched workloads at full size on the GPU: every spectrum of the
batched uneven workload (6 species, 9.61e6 lines, x 64 (T, P), 3e6 points)
and of the mixed-width stress workload (5e7 lines x 16 (T, P), 4e6 points),
each computed in one graph replay per form, FP32C against FP64 to 1e-5, and
their k-tables (FP32C against FP64: identical Nxmin, coefficients to 1e-4)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from heliosk import engine
    return engine


def _replay(E, items, mode, tile):
    import torch
    b = E.VoigtBatch(items, mode=mode, tile_points=tile)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.launch()
    g.replay()
    torch.cuda.synchronize()
    return g, b


@pytest.mark.parametrize("wname,tile", [("c2", 1024), ("c4", 512)])
def test_full_batched_workload(E, wname, tile):
    import torch
    from heliosk import configs
    w = configs.WORKLOADS[wname]
    grid = E.make_grid(w.numin, w.numax, w.dnu)
    nspec = len(w.lines) * len(w.tp)
    print("%s: %d spectra of %d points" % (wname, nspec, grid.nx), flush=True)
    K = torch.empty(nspec, grid.nx, dtype=torch.float64, device="cuda")
    items = []
    keep = []
    for s in range(len(w.lines)):
        arr = configs.species_arrays(w, s)
        sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"],
                       iso_mass=arr["iso_mass"])
        keep.append(sp)
        for T, Pbar in w.tp:
            st = E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
            items.append((st, grid, 0.0, K[len(items)]))
    g32, b32 = _replay(E, items, E.HK_VOIGT_FP32C, tile)
    K32 = K.clone()
    del g32, b32
    g64, b64 = _replay(E, items, E.HK_VOIGT_FP64, tile)
    del g64, b64
    tau = 1e-12 * torch.amax(K, dim=1, keepdim=True)
    rel = torch.amax(torch.abs(K32 - K) / torch.maximum(torch.abs(K), tau), dim=1)
    worst = int(torch.argmax(rel))
    print("%s: FP32C vs FP64 over %d spectra: max %.3e (spectrum %d), median %.3e"
          % (wname, nspec, float(rel.max()), worst, float(rel.median())))
    assert float(rel.max()) <= 1e-5
    kt = E.KTable(grid, w.numax, w.nbins, w.nC, nspec=nspec)
    c32, n32 = kt(K32)
    c64, n64 = kt(K)
    assert torch.equal(n32, n64)
    # bins with Nxmin = Nxb (all K <= kmin = 0, e.g. above the last line's
    # reach) have no fit: the reference prints their coefficients as 0.0
    fitted = n64 < grid.nx // w.nbins
    d = float(torch.max(torch.abs(c32 - c64)[fitted]))
    print("%s: k-tables FP32C vs FP64 max |dc| %.3e over %d fitted bins of %d"
          % (wname, d, int(fitted.sum()), fitted.numel()))
    assert d <= 1e-4
