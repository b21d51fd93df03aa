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
