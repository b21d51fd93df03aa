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


def test_cut_edge_expansion_error():
    """Cut-edge remainders (a band-1 line whose window ends inside the tile)
    are interpolated over one wave with the kernel's kEN nodes, for lines
    whose centre lies >= kEdgeMin wave half-widths from the wave's centre;
    nearer ones keep their records.  Band 1 alone would let such a line come
    within 11 half-widths (the line lies beyond the tile's far end by the
    band distance less a tile, and the wave's centre lies within a wave
    half-width of the window's end), where kEN nodes would not hold: the
    distance rule is what keeps the bound."""
    src = open(os.path.join(ROOT, "helios-k_amd", "csrc", "hk_voigt.hip")).read()
    n = int(re.search(r"constexpr int kEN = (\d+);", src).group(1))
    amin = float(re.search(r"constexpr double kEdgeMin = ([\d.]+);", src).group(1))
    _, edges = _constants()
    W = 1024
    hw = 0.5 * (W // 4 - 1)
    a_band = ((W / 2 + edges[0] * 0.5 * (W - 1)) - (W - 1) - hw) / hw
    assert 10.5 < a_band < 11.5
    e = _interp_rel_err(lambda u: 1.0 / (u - amin) ** 2, n)
    e_lor = _interp_rel_err(lambda u: 0.02 / ((u - amin) ** 2 + 0.02 ** 2), n)
    print("cut edge n=%d a=%.1f: 1/(u-a)^2 %.3e, Lorentz %.3e" % (n, amin, e, e_lor))
    assert e <= 4e-7 and e_lor <= 4e-7
    # the rule matters: at band 1's own nearest distance, or with one node
    # fewer, the bound fails
    assert _interp_rel_err(lambda u: 1.0 / (u - a_band) ** 2, n) > 1e-6
    assert _interp_rel_err(lambda u: 1.0 / (u - amin) ** 2, n - 1) > 4e-7
