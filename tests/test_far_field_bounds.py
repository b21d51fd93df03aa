"""The FP32 form's far-field interpolation error, measured on the kernel's
own band constants: Chebyshev-Gauss interpolation on n nodes of the
narrowest line profile, 1/(u - d)^2 over a tile u in [-1, 1], with the
centre d at the band's inner edge (the worst case of the band; wider
Lorentz profiles have their poles off the real axis and interpolate
better).  The pointwise relative errors stay below the stated bounds, so a
later cut in node counts or band edges shows here."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _constants():
    src = open(os.path.join(ROOT, "helios-k_amd", "csrc", "hk_voigt.hip")).read()
    n = re.search(r"kFN1 = (\d+), kFN2 = (\d+), kFN3 = (\d+);", src)
    r = re.search(r"kFarRatio1 = ([\d.]+), kFarRatio2 = ([\d.]+), kFarRatio3 = ([\d.]+);", src)
    assert n and r, "far-field constants not found"
    return [int(v) for v in n.groups()], [float(v) for v in r.groups()]


def _interp_rel_err(f, n, npts=20001):
    k = np.arange(n)
    theta = (2 * k + 1) * np.pi / (2 * n)
    fx = f(np.cos(theta))
    c = np.array([2.0 / n * np.sum(fx * np.cos(j * theta)) for j in range(n)])
    c[0] /= 2
    u = np.linspace(-1.0, 1.0, npts)
    return float(np.max(np.abs(np.polynomial.chebyshev.chebval(u, c) - f(u)) / np.abs(f(u))))


def test_far_field_band_errors():
    nodes, edges = _constants()
    bounds = (2.5e-7, 2.0e-7, 1.0e-6)           # the bounds the kernel comment states
    for n, d, b in zip(nodes, edges, bounds):
        e_pole = _interp_rel_err(lambda u: 1.0 / (u - d) ** 2, n)
        e_lor = _interp_rel_err(lambda u: 0.02 / ((u - d) ** 2 + 0.02 ** 2), n)
        print("band n=%d d=%g: 1/(u-d)^2 %.3e, Lorentz %.3e" % (n, d, e_pole, e_lor))
        assert e_pole <= b and e_lor <= b
    # the bands' outer edges are covered by the next band's inner-edge bound
    assert edges[0] > edges[1] > edges[2] >= 1.5


def test_far_field_error_grows_toward_the_tile():
    """Sanity of the measure: nearer centres need more nodes (a plausible
    mistake, a node count one band too low, fails the bound)."""
    nodes, edges = _constants()
    assert _interp_rel_err(lambda u: 1.0 / (u - edges[2]) ** 2, nodes[1]) > 1e-6
    assert _interp_rel_err(lambda u: 1.0 / (u - edges[1]) ** 2, nodes[0]) > 2e-7
