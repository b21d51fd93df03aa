"""GPU parity of the consumers of K: Planck / Rosseland mean sums and the
per-bin transmission function, against the reference's own example outputs
(plots/p009 means, plots/p006 transmission) and the FP64 oracle."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIGMA = (2 * 1.3806489e-16 ** 4 / (6.62606957e-27 ** 3 * 2.99792458e10 ** 2 * 15.0)) * math.pi ** 5


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


@pytest.mark.parametrize("run", ["p1", "p3"])
def test_means_p009(E, golden_dir, run):
    K = np.load(os.path.join(golden_dir, "p009_%s_K.npz" % run))["K"]
    grid = E.make_grid(24400.0, 24500.0, 0.001)
    assert grid.nx == K.size
    s = E.means(_dev(K), grid, 1500.0).cpu().numpy()
    so = O.means(K, 24400.0, 0.001, 1500.0)
    np.testing.assert_allclose(s, so, rtol=1e-13, atol=0)
    ref = np.loadtxt(os.path.join(golden_dir, "p009_%s_mean.dat" % run))
    got = [s[0] / s[2], s[3] / s[1], s[2] * 0.001, SIGMA * 1500.0 ** 4 / math.pi, s[3] * 0.001,
           4 * SIGMA * 1500.0 ** 3 / math.pi]
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


def test_means_random_with_zeros_and_x0(E):
    rng = np.random.default_rng(3)
    K = 10 ** rng.uniform(-30, -18, 3_000_001)
    K[rng.uniform(size=K.size) < 0.1] = 0.0
    grid = E.make_grid(0.0, 30000.0, 0.01)           # x_0 = 0: B = dB = 0 there
    grid.nx = K.size
    s = E.means(_dev(K), grid, 2200.0).cpu().numpy()
    so = O.means(K, 0.0, 0.01, 2200.0)
    np.testing.assert_allclose(s, so, rtol=1e-12, atol=0)


@pytest.mark.parametrize("b", [0, 4])
def test_transmission_p006(E, golden_dir, b):
    Ks = np.load(os.path.join(golden_dir, "p006_bin%d.npz" % b))["K"]
    ref = np.load(os.path.join(golden_dir, "p006_tr%d.npy" % b))
    Tr = E.transmission(_dev(Ks), 1, Ks.size, 1000, 0.05).cpu().numpy()[0]
    To = O.transmission(Ks, 1, Ks.size, 1000, 0.05)[0]
    np.testing.assert_allclose(Tr, To, rtol=2e-13, atol=0)
    np.testing.assert_allclose(Tr, ref[:, 1], rtol=2e-13, atol=0)


@pytest.mark.parametrize("nTr", [1, 7, 1000])
def test_transmission_bins_with_nxmin(E, nTr):
    rng = np.random.default_rng(nTr)
    nbins, Nxb = 37, 1003
    Ks = np.sort(10 ** rng.uniform(-4, 3, (nbins, Nxb)), axis=1).ravel()
    nxmin = rng.integers(0, Nxb + 1, nbins).astype(np.int32)
    Tr = E.transmission(_dev(Ks), nbins, Nxb, nTr, 0.03, kmin=1e-5, nxmin=nxmin).cpu().numpy()
    To = O.transmission(Ks, nbins, Nxb, nTr, 0.03, kmin=1e-5, nxmin=nxmin)
    np.testing.assert_allclose(Tr, To, rtol=1e-13, atol=1e-300)
