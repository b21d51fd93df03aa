"""GPU parity of the k-table stage (per-bin sort, findCut, rescale, fit,
reconstruction) against the reference's own outputs and the FP64 oracle."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from heliosk import engine
    return engine


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def test_p004_fit_sort_and_reconstruction(E, golden_dir):
    K = np.load(os.path.join(golden_dir, "p004_K.npy"))
    grid = E.make_grid(0.0, 30000.0, 1.0)
    coeff, nxmin, Ks = E.ktable(_dev(K), grid, 30000.0, 1, 20, kmin=0.0, want_sorted=True)
    ref = np.loadtxt(os.path.join(golden_dir, "p004_Out_r_cbin.dat"))
    assert int(nxmin[0, 0]) == 4533 and ref[1] == 4533 / 29999.0
    np.testing.assert_allclose(coeff[0, 0].cpu().numpy(), ref[2:], rtol=1e-10, atol=1e-12)
    sha = hashlib.sha256(Ks.cpu().numpy().tobytes()).hexdigest()
    assert sha == open(os.path.join(golden_dir, "p004_sorted.sha256")).read().split()[0]
    Ky = E.ktable_reconstruct(coeff, nxmin, 30000).cpu().numpy()
    Kref = np.load(os.path.join(golden_dir, "p004_Out_r_bin_K.npy"))
    np.testing.assert_allclose(Ky, Kref, rtol=1e-10)


def test_p005_fit(E, golden_dir):
    Ks = np.loadtxt(os.path.join(golden_dir, "p005_Out_i_bin.dat"), usecols=(1,))
    grid = E.make_grid(25000.0, 30000.0, 10.0)
    coeff, nxmin = E.ktable(_dev(Ks), grid, 30000.0, 1, 20, kmin=1e-7)
    ref = np.loadtxt(os.path.join(golden_dir, "p005_Out_r_cbin.dat"))
    assert int(nxmin[0, 0]) == 426
    np.testing.assert_allclose(coeff[0, 0].cpu().numpy(), ref[2:], rtol=1e-9, atol=1e-12)
    Ky = E.ktable_reconstruct(coeff, nxmin, 500, kmin=1e-7).cpu().numpy()
    Kref = np.loadtxt(os.path.join(golden_dir, "p005_Out_r_bin.dat"), usecols=(1,))
    np.testing.assert_allclose(Ky, Kref, rtol=1e-10)


@pytest.mark.parametrize("run", ["p1", "p3"])
def test_p009_sort_large_bins(E, golden_dir, run):
    # 5 bins of 20000 points: longer than one LDS sort, so the merge passes run
    K = np.load(os.path.join(golden_dir, "p009_%s_K.npz" % run))["K"]
    grid = E.make_grid(24400.0, 24500.0, 0.001)
    _, _, Ks = E.ktable(_dev(K), grid, 24500.0, 5, 20, kmin=0.0, want_sorted=True)
    sha = hashlib.sha256(Ks.cpu().numpy().tobytes()).hexdigest()
    assert sha == open(os.path.join(golden_dir, "p009_%s_sorted.sha256" % run)).read().split()[0]


@pytest.mark.parametrize("Nxb,nbins,nspec,kmin", [(10000, 6, 3, 0.0), (100000, 3, 2, 1e-9),
                                                  (777, 5, 1, 0.5), (300, 4, 2, 1e-9)])
def test_random_spectra_vs_oracle(E, Nxb, nbins, nspec, kmin):
    rng = np.random.default_rng(Nxb)
    nx = Nxb * nbins
    K = np.exp(rng.normal(-20.0, 4.0, (nspec, nx)))
    if kmin > 0:
        K[:, :Nxb // 3] = kmin * rng.uniform(0.1, 1.0, (nspec, Nxb // 3))   # cut bins
        K[0, nx - 5:] = kmin                                                   # last-bin rule
    grid = E.make_grid(0.0, float(nx), 1.0)
    coeff, nxmin, Ks = E.ktable(_dev(K.ravel()), grid, float(nx), nbins, 20, kmin=kmin,
                                want_sorted=True)
    Ks = Ks.cpu().numpy().reshape(nspec, nx)
    seg = O.bin_segments(0.0, float(nx), 1.0, nx, nbins)
    for s in range(nspec):
        so = O.sort_segments(K[s], seg[:nbins + 1])
        np.testing.assert_array_equal(Ks[s], so)
        co, no = O.ktable(so, Nxb, nbins, 20, kmin)
        np.testing.assert_array_equal(nxmin[s].cpu().numpy(), no)
        c = coeff[s].cpu().numpy()
        scale = np.max(np.abs(co), axis=1, keepdims=True)
        assert np.max(np.abs(c - co) / scale) <= 1e-11


def test_segment_rules_match_oracle(E):
    for numin, numax, dnu, nbins in ((0.0, 30000.0, 0.01, 300), (0.0, 30000.0, 0.001, 300),
                                     (24400.0, 24500.0, 0.001, 5), (0.3, 1000.3, 0.1, 7)):
        g = E.make_grid(numin, numax, dnu)
        np.testing.assert_array_equal(E.bin_segments(g, numax, nbins),
                                      O.bin_segments(numin, numax, dnu, g.nx, nbins))


@pytest.mark.parametrize("Nxb,nbins,nC,kmin", [(5000, 3, 40, 0.0), (2000, 2, 100, 1e-9),
                                               (30000, 2, 100, 0.0), (12, 4, 12, 0.0),
                                               (50, 3, 1, 0.0), (40000, 1, 2, 1e-9),
                                               (16384, 2, 33, 0.0)])
def test_nc_range_vs_oracle(E, Nxb, nbins, nC, kmin):
    """nC from 1 to the reference's maximum of 100 (def_NmaxSample), bins as
    short as nC points (the square QR of resample.h) and bins of both sort
    paths (one LDS sort up to 16384 points, runs + buckets above)."""
    rng = np.random.default_rng(nC * 7 + Nxb)
    nx = Nxb * nbins
    K = np.exp(rng.normal(-25.0, 3.0, nx))
    if kmin > 0:
        K[: Nxb // 4] = kmin
    grid = E.make_grid(0.0, float(nx), 1.0)
    coeff, nxmin, Ks = E.ktable(_dev(K), grid, float(nx), nbins, nC, kmin=kmin, want_sorted=True)
    seg = O.bin_segments(0.0, float(nx), 1.0, nx, nbins)
    so = O.sort_segments(K, seg[:nbins + 1])
    np.testing.assert_array_equal(Ks.cpu().numpy(), so)
    co, no = O.ktable(so, Nxb, nbins, nC, kmin)
    np.testing.assert_array_equal(nxmin[0].cpu().numpy(), no)
    c = coeff[0].cpu().numpy()
    scale = np.max(np.abs(co), axis=1, keepdims=True)
    err = np.max(np.abs(c - co) / scale)
    print("Nxb %d nC %d: coefficients vs oracle %.2e" % (Nxb, nC, err))
    assert err <= 1e-9


def test_many_spectra_beyond_grid_y_limit(E):
    """nspec * (nbins + 1) = 66220 segments (more than 65535 workgroups in one
    grid dimension): one call equals per-spectrum calls bit for bit."""
    import torch
    nspec, nbins, Nxb, nC = 220, 300, 20, 4
    nx = nbins * Nxb
    rng = np.random.default_rng(5)
    K = np.exp(rng.normal(-20.0, 2.0, (nspec, nx)))
    grid = E.make_grid(0.0, float(nx), 1.0)
    coeff, nxmin, Ks = E.ktable(_dev(K.ravel()), grid, float(nx), nbins, nC, want_sorted=True)
    kt1 = E.KTable(grid, float(nx), nbins, nC)
    for s in (0, 1, 117, nspec - 1):
        c1, n1, k1 = kt1(_dev(K[s]), want_sorted=True)
        assert torch.equal(coeff[s], c1[0]) and torch.equal(nxmin[s], n1[0])
        assert torch.equal(Ks.view(nspec, nx)[s], k1)


@pytest.mark.parametrize("Nxb,kind", [(300000, "ties"), (1200000, "kmin"), (70000, "ties")])
def test_long_bins_with_ties(E, Nxb, kind):
    """Bins longer than 16 LDS-sorted runs (merge passes before the bucket
    pass) and bins full of equal values (the bucket splitters repeat): the
    sorted bin equals numpy's sort and the fit equals the oracle's."""
    rng = np.random.default_rng(Nxb)
    if kind == "ties":
        K = rng.integers(1, 40, Nxb).astype(np.float64) * 1e-22
        kmin = 0.0
    else:
        K = np.where(rng.random(Nxb) < 0.8, 1e-12, np.exp(rng.normal(-20.0, 3.0, Nxb)))
        kmin = 1e-12
    grid = E.make_grid(0.0, float(Nxb), 1.0)
    coeff, nxmin, Ks = E.ktable(_dev(K), grid, float(Nxb), 1, 20, kmin=kmin, want_sorted=True)
    so = np.sort(K)
    np.testing.assert_array_equal(Ks.cpu().numpy(), so)
    co, no = O.ktable(so, Nxb, 1, 20, kmin)
    assert int(nxmin[0, 0]) == int(no[0])
    c = coeff[0, 0].cpu().numpy()
    assert np.max(np.abs(c - co[0])) / np.max(np.abs(co[0])) <= 1e-11


def test_sort_bins_mixed_lengths(E):
    """hk_sort_bins on segments of 0, 1, 17, 16384, 16385 and 250000 points in
    one call (the LDS and the run + bucket paths together)."""
    lens = [0, 1, 17, 16384, 16385, 250000, 3]
    seg = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    rng = np.random.default_rng(11)
    K = np.round(rng.normal(0.0, 1.0, int(seg[-1])), 3)      # signed values with ties
    Ks = E.sort_bins(_dev(K), seg).cpu().numpy()
    for b in range(len(lens)):
        np.testing.assert_array_equal(Ks[seg[b]:seg[b + 1]], np.sort(K[seg[b]:seg[b + 1]]))


def test_sort_bins_signed_zeros(E):
    """Bins holding negative zeros take the compare-and-select sort (min/max
    instructions may return either of -0 and +0).  Sorted whole in LDS (no
    segment of the call above 16384 values), every input bit pattern comes
    out once and the values are in order; such bins sit beside ordinary
    ones.  A call with a longer segment runs the bucket pass, which emits
    the copies of a splitter value from the splitter: there the values are
    compared."""
    rng = np.random.default_rng(5)
    bits = lambda a: np.sort(a.view(np.uint64))
    lens = [10000, 16384, 257, 5000]
    seg = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    K = np.round(rng.normal(0.0, 1.0, int(seg[-1])), 1)      # many +0 and -0
    K[seg[2]:seg[3]] = np.abs(K[seg[2]:seg[3]]) + 1.0         # an ordinary bin
    assert np.any(np.signbit(K[:seg[1]]) & (K[:seg[1]] == 0.0))
    Ks = E.sort_bins(_dev(K), seg).cpu().numpy()
    for b in range(len(lens)):
        src, out = K[seg[b]:seg[b + 1]], Ks[seg[b]:seg[b + 1]]
        np.testing.assert_array_equal(bits(out), bits(src))
        np.testing.assert_array_equal(out, np.sort(src))
    lens = [70000, 16384]
    seg = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    K = np.round(rng.normal(0.0, 1.0, int(seg[-1])), 1)
    Ks = E.sort_bins(_dev(K), seg).cpu().numpy()
    for b in range(len(lens)):
        np.testing.assert_array_equal(Ks[seg[b]:seg[b + 1]], np.sort(K[seg[b]:seg[b + 1]]))
