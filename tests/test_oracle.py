"""Pins the FP64 host oracle (oracle/hk_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
wrong index, transposed operand) fails it: mpmath golden vectors, closed
forms, brute force on tiny inputs, and the reference's own example outputs
(plots/p004, p005, p009, copied into tests/golden).
"""
import hashlib
from fractions import Fraction
import math
import os

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

C2 = 6.62606957e-27 * 2.99792458e10 / 1.3806489e-16   # h c / kB, define.h:39-41
T0 = 296.0


# ---------------------------------------------------------------- Faddeeva
def test_faddeeva_vs_mpmath_golden(golden_dir):
    d = np.loadtxt(os.path.join(golden_dir, "faddeeva_mpmath.csv"), delimiter=",")
    x, y, re, im = d.T
    w = O.faddeeva(x, y)
    nz = re != 0.0            # exp(-x^2) below the double range reads as 0
    assert np.all(w.real[~nz] < 1e-300)
    relre = np.abs(w.real[nz] - re[nz]) / np.abs(re[nz])
    assert relre.max() < 1e-13, relre.max()
    # Im w: absolute accuracy relative to |w| (Im w -> 0 at x -> 0)
    absim = np.abs(w.imag - im) / np.abs(re + 1j * im)
    assert absim.max() < 1e-13


def test_faddeeva_closed_forms():
    ys = np.array([0.0, 1e-6, 0.1, 1.0, 4.9, 5.1, 30.0, 1e4])
    for y in ys:
        ref = float(mp.exp(mp.mpf(y) ** 2) * mp.erfc(mp.mpf(y)))
        assert abs(O.erfcx(y) - ref) <= 2e-15 * ref
        assert abs(O.faddeeva(0.0, y)[0].real - ref) <= 2e-15 * ref      # w(iy) = erfcx(y)
    xs = np.array([0.0, 0.3, 1.0, 3.0, 6.0, 9.0, 9.99, 10.01, 20.0])
    w = O.faddeeva(xs, 0.0)
    np.testing.assert_allclose(w.real, np.exp(-xs * xs), rtol=1e-14, atol=0)  # Re w(x,0)
    # Dawson: Im w(x,0) = 2/sqrt(pi) D(x)
    daw = np.array([float(2 / mp.sqrt(mp.pi) * mp.sqrt(mp.pi) / 2 * mp.exp(-mp.mpf(v) ** 2)
                          * mp.erfi(mp.mpf(v))) for v in xs])
    np.testing.assert_allclose(w.imag, daw, rtol=1e-13, atol=1e-300)
    # symmetry w(-x+iy) = conj w(x+iy)
    a = O.faddeeva(np.array([2.5, 50.0]), 0.7); b = O.faddeeva(np.array([-2.5, -50.0]), 0.7)
    np.testing.assert_allclose(b, np.conj(a), rtol=1e-15)


def test_faddeeva_normalisation():
    # integral of Re w(x, y) dx over the real line = sqrt(pi)
    for y in (0.05, 1.0, 20.0):
        L = 2.0e4
        x = np.linspace(-L, L, 4_000_001)
        f = O.faddeeva(x, y).real
        integral = np.trapezoid(f, x) + 2 * y / (math.sqrt(math.pi) * L)  # tails beyond +-L
        assert abs(integral - math.sqrt(math.pi)) < 1e-9 * math.sqrt(math.pi)


def test_faddeeva_methods_agree_on_overlap():
    # Algorithm 916 and the continued fraction are independent methods
    for x, y in ((8.0, 6.0), (10.0, 1e-3), (3.0, 12.0), (15.0, 0.2)):
        a = O.w916(x, y); b = O.wcf(x, y, 40)
        assert abs(a.real - b.real) <= 1e-13 * abs(b.real)


# ---------------------------------------------------------------- line parameters
def _lines(n, seed=1, delta=0.0, A=0.0):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(100.0, 3000.0, n))
    return dict(nu=nu, S=10 ** rng.uniform(-25, -19, n), EL=rng.uniform(0, 20000, n),
                A=np.full(n, A), delta=np.full(n, delta), vy=rng.uniform(0.01, 0.1, n),
                n=rng.uniform(0.3, 0.8, n), iso=None,
                iso_mass=np.array([18.010565 / 6.0221412927e23]), iso_ab=np.array([0.9973]),
                iso_sscale=np.array([1.0]), iso_Q=np.array([174.58]))


def test_line_params_closed_forms():
    L = _lines(200)
    m = L["iso_mass"][0]
    ia0 = 2.99792458e10 * math.sqrt(m / (2 * 1.3806489e-16 * T0))
    st = O.line_params(L, T0, 1.0, 0, 25.0)
    # at T0 the Boltzmann and stimulated-emission ratios are exactly 1
    S_expect = L["S"] / 174.58 * 0.9973 * (ia0 / L["nu"])
    np.testing.assert_allclose(st["S"], S_expect, rtol=1e-14)
    np.testing.assert_allclose(st["ialphaD"] * L["nu"], ia0, rtol=1e-15)
    np.testing.assert_allclose(st["y"], L["vy"] * st["ialphaD"], rtol=1e-15)
    assert np.all(st["hw"] == 25.0)
    # temperature scaling: alpha_D ~ sqrt(T); alpha_L ~ (T0/T)^n
    T = 1500.0
    s2 = O.line_params(L, T, 3.0, 0, 25.0)
    np.testing.assert_allclose(st["ialphaD"] / s2["ialphaD"], math.sqrt(T / T0), rtol=1e-14)
    aL = s2["y"] / s2["ialphaD"]
    np.testing.assert_allclose(aL, L["vy"] * 3.0 * (T0 / T) ** L["n"], rtol=1e-14)
    # S(T)/S(T0): Boltzmann and stimulated-emission factors
    ratio = (np.exp(-C2 * L["EL"] / T + C2 * L["EL"] / T0)
             * (1 - np.exp(-C2 * L["nu"] / T)) / (1 - np.exp(-C2 * L["nu"] / T0)))
    np.testing.assert_allclose(s2["S"] / st["S"], ratio * s2["ialphaD"] / st["ialphaD"], rtol=1e-12)
    # cut modes: 1 = Lorentz widths, 2 = Doppler widths
    s1 = O.line_params(L, T, 3.0, 1, 100.0)
    np.testing.assert_allclose(s1["hw"], 100.0 * aL, rtol=1e-15)
    s3 = O.line_params(L, T, 3.0, 2, 50.0)
    np.testing.assert_allclose(s3["hw"], 50.0 / s3["ialphaD"], rtol=1e-15)


def test_line_params_shift_and_natural_width():
    L = _lines(50, delta=-0.01, A=2.0e7)
    st = O.line_params(L, 800.0, 2.0, 0, 25.0)
    np.testing.assert_allclose(st["nu"], L["nu"] - 0.02, rtol=1e-15)
    aL = st["y"] / st["ialphaD"]
    expect = L["vy"] * 2.0 * (T0 / 800.0) ** L["n"] + 2.0e7 / (4 * math.pi * 2.99792458e10)
    np.testing.assert_allclose(aL, expect, rtol=1e-14)
    L["nu"][0] = 0.02       # shifted to exactly 0 -> zero line
    st = O.line_params(L, 800.0, 2.0, 0, 25.0)
    assert st["S"][0] == 0 and st["ialphaD"][0] == 0 and st["hw"][0] < 0


def test_line_params_exomol_and_superline_forms():
    """Sf_kernel (voigt.h:365-385) and SfSuper_kernel (voigt.h:397-410)
    against the HITRAN form: the ExoMol strength lacks the T0 normalisation
    (ratio exp(-c2 E/T0) (1 - exp(-c2 nu/T0))), widths scale with P/0.986923
    bar, centres are unshifted; super-lines carry no Q, no E'' and no A."""
    T, P = 1234.0, 2.5
    L = _lines(300, seed=7, delta=-0.003, A=3.0e7)
    L["iso_Q"] = np.array([321.0])
    hit = O.line_params(dict(L, delta=np.zeros(300), A=np.zeros(300)), T, P, 0, 25.0, form=0)
    exo0 = O.line_params(dict(L, A=np.zeros(300)), T, P, 0, 25.0, form=1)
    exo = O.line_params(L, T, P, 0, 25.0, form=1)
    sup = O.line_params(L, T, P, 0, 25.0, form=2)
    for st in (exo, sup):
        np.testing.assert_array_equal(st["nu"], L["nu"])       # no pressure shift
        np.testing.assert_allclose(st["ialphaD"], hit["ialphaD"], rtol=1e-15)
    np.testing.assert_allclose(exo["S"] / hit["S"],
                               np.exp(-C2 * L["EL"] / T0) * (1 - np.exp(-C2 * L["nu"] / T0)),
                               rtol=1e-12)
    np.testing.assert_allclose(exo0["y"] / hit["y"], 1.0 / 0.986923, rtol=1e-14)
    np.testing.assert_allclose((exo["y"] - exo0["y"]) / exo["ialphaD"],
                               3.0e7 / (4 * math.pi * 2.99792458e10), rtol=1e-9)
    # super-lines: S_file * Ab * Sscale / alpha_D, Lorentz width without A
    np.testing.assert_allclose(sup["S"], L["S"] * 0.9973 * sup["ialphaD"], rtol=1e-15)
    np.testing.assert_allclose(sup["y"], exo0["y"], rtol=1e-15)
    np.testing.assert_array_equal(sup["hw"], exo["hw"])


# ---------------------------------------------------------------- windows
def test_window_brute_force():
    rng = np.random.default_rng(3)
    numin, dnu, Nx = 10.0, 0.01, 5000
    j = np.arange(Nx, dtype=np.float64)
    x = numin + j * dnu   # fma equals mul+add here to within 1 ulp; check both sides
    for _ in range(300):
        nu = rng.uniform(0.0, 70.0)
        hw = rng.choice([rng.uniform(0.0, 5.0), 1e30, 0.0, dnu * rng.integers(1, 50)])
        lo, hi = O.window(nu, hw, numin, dnu, Nx)
        inside = np.abs(x - nu) < hw
        idx = np.nonzero(inside)[0]
        if idx.size == 0:
            assert hi == lo
        else:
            # a point exactly on the cut edge may flip with fma rounding: allow 1
            assert abs(lo - idx[0]) <= 1 and abs(hi - (idx[-1] + 1)) <= 1
            assert hi - lo >= idx.size - 2


# ---------------------------------------------------------------- Voigt sum
def _state(n=3, seed=0):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(1000.2, 1000.8, n))
    return dict(nu=nu, S=rng.uniform(0.5, 2.0, n), ialphaD=rng.uniform(5.0, 40.0, n),
                y=rng.uniform(1e-3, 3.0, n), hw=np.full(n, 0.3))


def test_voigt_points_vs_mpmath_brute_force():
    st = _state(3)
    numin, dnu, Nx = 1000.0, 0.02, 50
    K = O.voigt_points(st, numin, dnu, Nx, kmin=1e-9)
    mp.mp.dps = 30
    for j in range(0, Nx, 7):
        xj = mp.mpf(float(Fraction(j) * Fraction(dnu) + Fraction(numin)))  # fma(j, dnu, numin)
        tot = mp.mpf("1e-9")
        for i in range(3):
            d = xj - mp.mpf(st["nu"][i])
            if abs(d) < st["hw"][i]:
                z = mp.mpc(d * st["ialphaD"][i], st["y"][i])
                tot += st["S"][i] * (mp.exp(-z * z) * mp.erfc(-1j * z)).real / mp.sqrt(mp.pi)
        assert abs(K[j] - float(tot)) <= 1e-13 * float(tot)


def test_voigt_points_invariants():
    st = _state(5, seed=2)
    numin, dnu, Nx = 1000.0, 0.005, 400
    K1 = O.voigt_points(st, numin, dnu, Nx)
    st2 = dict(st); st2["S"] = st["S"] * 3.0
    np.testing.assert_allclose(O.voigt_points(st2, numin, dnu, Nx), 3.0 * K1, rtol=1e-15)  # linear
    # translation by whole grid steps
    m = 37
    st3 = dict(st); st3["nu"] = st["nu"] + m * dnu
    K3 = O.voigt_points(st3, numin, dnu, Nx + m)
    np.testing.assert_allclose(K3[m + 60:m + 340], K1[60:340], rtol=1e-9)
    # symmetry of one line centred on a grid point
    one = dict(nu=np.array([1000.0 + 200 * dnu]), S=np.array([1.0]), ialphaD=np.array([30.0]),
               y=np.array([0.4]), hw=np.array([1.0]))
    K = O.voigt_points(one, numin, dnu, Nx)
    np.testing.assert_allclose(K[200 - 150:200], K[200 + 150:200:-1], rtol=1e-12)
    # subset of points equals the full evaluation
    jj = np.array([0, 5, 199, 200, 399])
    np.testing.assert_array_equal(O.voigt_points(st, numin, dnu, Nx, jidx=jj), K1[jj])


def test_voigt_points_normalisation():
    # one uncut line on a wide grid: sum_j K_j dnu = S minus the Lorentz tails
    aD = 0.05
    for y in (0.02, 1.0):
        st = dict(nu=np.array([500.0]), S=np.array([1.0 / aD]), ialphaD=np.array([1 / aD]),
                  y=np.array([y]), hw=np.array([1e30]))
        numin, dnu, Nx = 300.0, 0.002, 200_000
        K = O.voigt_points(st, numin, dnu, Nx)
        aL = y * aD
        tails = aL / math.pi * (1 / 200.0 + 1 / 200.0)
        # state S carries 1/alpha_D, so S = 1/aD is a line of unit strength
        assert abs(K.sum() * dnu - (1 - tails)) < 1e-8


# ---------------------------------------------------------------- gradient
@pytest.mark.parametrize("form", [0, 1, 2])
def test_voigt_grad_vs_central_differences(form):
    L = _lines(6, seed=4, delta=-0.002 if form == 0 else 0.0, A=1.0e8)
    L["nu"] = np.sort(np.random.default_rng(5).uniform(1000.0, 1001.0, 6))
    L["S"] = np.full(6, 1e-20)
    L["EL"] = np.linspace(0.0, 300.0, 6)        # comparable line contributions
    numin, dnu, Nx = 999.5, 0.002, 1000
    T, P = 900.0, 0.7
    gK = np.random.default_rng(6).uniform(0.0, 1.0, Nx)

    def loss(LL, TT, PP):
        st = O.line_params(LL, TT, PP, 0, 0.8, form=form)
        return float(np.dot(gK, O.voigt_points(st, numin, dnu, Nx)))

    gT, gP, gS, gv = O.voigt_grad(L, T, P, 0, 0.8, numin, dnu, Nx, gK, form=form)
    h = 1e-5
    fdT = (loss(L, T * (1 + h), P) - loss(L, T * (1 - h), P)) / (2 * h * T)
    fdP = (loss(L, T, P * (1 + h)) - loss(L, T, P * (1 - h))) / (2 * h * P)
    assert abs(gT - fdT) <= 1e-6 * abs(fdT)
    assert abs(gP - fdP) <= 1e-6 * abs(fdP)
    for i in (0, 3, 5):
        for key, g in (("S", gS), ("vy", gv)):
            Lp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in L.items()}
            Lm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in L.items()}
            Lp[key][i] *= 1 + h; Lm[key][i] *= 1 - h
            fd = (loss(Lp, T, P) - loss(Lm, T, P)) / (2 * h * L[key][i])
            assert abs(g[i] - fd) <= 1e-6 * abs(fd), (key, i, g[i], fd)


def test_voigt_grad_far_field_lines():
    """Lines whose profile on the grid lies at |z|^2 >> 1e4 (centres near
    0 cm^-1: alpha_D -> 0, y and x of order 1e4-1e9).  The derivative there
    comes from the asymptotic series of w'; a single line is differentiated
    so the central difference is free of cancellation."""
    L = _lines(1, seed=7)
    numin, dnu, Nx = 0.0, 0.01, 5000
    gK = np.random.default_rng(8).uniform(0.0, 1.0, Nx)
    for nu, T, P in ((0.0314, 1200.0, 0.0987), (2.5, 300.0, 1.0), (0.5, 2500.0, 10.0)):
        L["nu"] = np.array([nu])

        def loss(LL, TT, PP):
            st = O.line_params(LL, TT, PP, 0, 50.0)
            return float(np.dot(gK, O.voigt_points(st, numin, dnu, Nx)))

        gT, gP, gS, gv = O.voigt_grad(L, T, P, 0, 50.0, numin, dnu, Nx, gK)
        st = O.line_params(L, T, P, 0, 50.0)
        assert st["y"][0] > 100.0              # every point has |z|^2 > 1e4
        h = 1e-5
        for key, g in (("vy", gv[0]), ("S", gS[0])):
            Lp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in L.items()}
            Lm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in L.items()}
            Lp[key][0] *= 1 + h; Lm[key][0] *= 1 - h
            fd = (loss(Lp, T, P) - loss(Lm, T, P)) / (2 * h * L[key][0])
            assert abs(g - fd) <= 1e-7 * abs(fd), (nu, key, g, fd)
        fdT = (loss(L, T * (1 + h), P) - loss(L, T * (1 - h), P)) / (2 * h * T)
        assert abs(gT - fdT) <= 1e-6 * abs(fdT), (nu, gT, fdT)


# ---------------------------------------------------------------- k-table
def _cbin(path):
    return np.loadtxt(path)


def test_sort_matches_p004_and_p009(golden_dir):
    K = np.load(os.path.join(golden_dir, "p004_K.npy"))
    seg = O.bin_segments(0.0, 30000.0, 1.0, 30000, 1)
    assert list(seg) == [0, 30000, 30000]
    s = O.sort_segments(K, seg[:2])
    ref = open(os.path.join(golden_dir, "p004_sorted.sha256")).read().split()[0]
    assert hashlib.sha256(s.tobytes()).hexdigest() == ref
    for run in ("p1", "p3"):
        K = np.load(os.path.join(golden_dir, "p009_%s_K.npz" % run))["K"]
        seg = O.bin_segments(24400.0, 24500.0, 0.001, 100000, 5)
        assert list(seg) == [0, 20000, 40000, 60000, 80000, 100000, 100000]
        s = O.sort_segments(K, seg[:6])
        ref = open(os.path.join(golden_dir, "p009_%s_sorted.sha256" % run)).read().split()[0]
        assert hashlib.sha256(s.tobytes()).hexdigest() == ref


def test_ktable_matches_p004(golden_dir):
    K = np.load(os.path.join(golden_dir, "p004_K.npy"))
    Ks = np.sort(K)
    coeff, nxmin = O.ktable(Ks, 30000, 1, 20, 0.0)
    ref = _cbin(os.path.join(golden_dir, "p004_Out_r_cbin.dat"))
    assert ref[0] == 0.0 and nxmin[0] == 4533 and ref[1] == min(nxmin[0] / 29999.0, 1.0)
    np.testing.assert_allclose(coeff[0], ref[2:], rtol=2e-12, atol=1e-13)
    Ky = O.reconstruct(coeff, nxmin, 30000, 0.0)
    Kref = np.load(os.path.join(golden_dir, "p004_Out_r_bin_K.npy"))
    np.testing.assert_allclose(Ky, Kref, rtol=1e-12)


def test_ktable_matches_p005(golden_dir):
    Ks = np.loadtxt(os.path.join(golden_dir, "p005_Out_i_bin.dat"), usecols=(1,))
    coeff, nxmin = O.ktable(Ks, 500, 1, 20, 1e-7)
    ref = _cbin(os.path.join(golden_dir, "p005_Out_r_cbin.dat"))
    assert nxmin[0] == 426 and ref[1] == 426 / 499.0
    np.testing.assert_allclose(coeff[0], ref[2:], rtol=1e-11, atol=1e-13)
    Ky = O.reconstruct(coeff, nxmin, 500, 1e-7)
    Kref = np.loadtxt(os.path.join(golden_dir, "p005_Out_r_bin.dat"), usecols=(1,))
    np.testing.assert_allclose(Ky, Kref, rtol=1e-12)


def test_ktable_quirk_is_pinned(golden_dir):
    # Planting the textbook NC reflections (instead of NC-1) must break p005.
    Ks = np.loadtxt(os.path.join(golden_dir, "p005_Out_i_bin.dat"), usecols=(1,))
    ref = _cbin(os.path.join(golden_dir, "p005_Out_r_cbin.dat"))
    n = 426; Nxb = 500
    s = 1.0 - n / (Nxb - 1.0)
    ii = np.minimum(n + s * np.arange(Nxb), 0.999999 * (Nxb - 1))
    ii = np.where(n + s * np.arange(Nxb) >= Nxb - 1, 0.999999 * (Nxb - 1), n + s * np.arange(Nxb))
    il = (ii / Nxb * Nxb).astype(int)
    b = np.log(Ks[il] + (Ks[il + 1] - Ks[il]) * (ii - il))
    x = np.arange(Nxb) * 2.0 / (Nxb - 1.0) - 1.0
    V = np.polynomial.chebyshev.chebvander(x, 19)
    c_textbook = np.linalg.lstsq(V, b, rcond=None)[0]
    assert np.max(np.abs(c_textbook - ref[2:]) / np.abs(ref[2:])) > 1e-3


def test_findcut_rules():
    # non-last bin: Nxmin = count(K <= kmin); last bin: K[Nx-3] <= kmin -> Nxb
    Nxb, nbins = 10, 3
    K = np.concatenate([np.r_[np.zeros(4), np.linspace(1, 2, 6)],
                        np.zeros(10),
                        np.r_[np.zeros(8), 1.0, 2.0]])
    _, nxmin = O.ktable(K.copy(), Nxb, nbins, 4, 0.0)
    assert list(nxmin) == [4, 10, 10]
    K[20:] = np.r_[np.zeros(5), np.linspace(1, 2, 5)]
    _, nxmin = O.ktable(K.copy(), Nxb, nbins, 4, 0.0)
    assert list(nxmin) == [4, 10, 5]
    K[:10] = np.linspace(1, 2, 10)
    _, nxmin = O.ktable(K.copy(), Nxb, nbins, 4, 0.0)
    assert nxmin[0] == 0


# ---------------------------------------------------------------- means and transmission
SIGMA = (2 * 1.3806489e-16 ** 4 / (6.62606957e-27 ** 3 * 2.99792458e10 ** 2 * 15.0)) * math.pi ** 5


def mean_lines(sums, T, dnu):
    """The six numbers of Out_<name>_mean.dat for one pressure
    (heliosk.cu:1976-1999)."""
    return [sums[0] / sums[2], sums[3] / sums[1], sums[2] * dnu, SIGMA * T ** 4 / math.pi,
            sums[3] * dnu, 4.0 * SIGMA * T ** 3 / math.pi]


@pytest.mark.parametrize("run", ["p1", "p3"])
def test_means_match_p009(golden_dir, run):
    K = np.load(os.path.join(golden_dir, "p009_%s_K.npz" % run))["K"]
    ref = np.loadtxt(os.path.join(golden_dir, "p009_%s_mean.dat" % run))
    got = mean_lines(O.means(K, 24400.0, 0.001, 1500.0), 1500.0, 0.001)
    np.testing.assert_allclose(got, ref, rtol=2e-14, atol=0)


def test_means_closed_forms():
    T = 800.0
    K = np.full(600000, 3.5e-21)
    s = O.means(K, 0.0, 0.05, T)            # x from 0 to 3e4 cm^-1: the whole Planck curve
    assert abs(s[0] / s[2] - 3.5e-21) <= 1e-12 * 3.5e-21     # constant K: both means are K
    assert abs(s[3] / s[1] - 3.5e-21) <= 1e-12 * 3.5e-21
    assert abs(s[2] * 0.05 / (SIGMA * T ** 4 / math.pi) - 1) <= 1e-9       # int B = sigma T^4/pi
    assert abs(s[3] * 0.05 / (4 * SIGMA * T ** 3 / math.pi) - 1) <= 1e-9   # int dB/dT
    K[::7] = 0.0                             # zeros drop out of the Rosseland sum only
    s2 = O.means(K, 0.0, 0.05, T)
    assert s2[1] < s[1] and s2[2] == s[2] and s2[3] == s[3]


@pytest.mark.parametrize("b", [0, 4])
def test_transmission_matches_p006(golden_dir, b):
    Ks = np.load(os.path.join(golden_dir, "p006_bin%d.npz" % b))["K"]
    ref = np.load(os.path.join(golden_dir, "p006_tr%d.npy" % b))
    Tr = O.transmission(Ks, 1, Ks.size, 1000, 0.05)[0]
    np.testing.assert_allclose(np.exp((np.arange(1000) - 500) * 0.05), ref[:, 0], rtol=1e-15)
    np.testing.assert_allclose(Tr, ref[:, 1], rtol=2e-13, atol=0)


def test_transmission_closed_forms():
    nTr, dTr, N = 40, 0.2, 50
    m = np.exp((np.arange(nTr) - nTr // 2) * dTr)
    Ks = np.repeat([0.3, 2.0, 7.0], N)
    Tr = O.transmission(Ks, 3, N, nTr, dTr)
    # a constant bin: the end weights sum to -1, so Tr = exp(-K m) exactly
    np.testing.assert_allclose(Tr, np.exp(-np.outer([0.3, 2.0, 7.0], m)), rtol=1e-14)
    nmin = np.array([0, 10, 50], dtype=np.int32)
    Tr = O.transmission(Ks, 3, N, nTr, dTr, kmin=0.1, nxmin=nmin)
    expect = (np.exp(-np.outer([0.3, 2.0, 7.0], m)) * ((N - nmin) / N)[:, None]
              + np.exp(-0.1 * m)[None, :] * (nmin / N)[:, None])
    np.testing.assert_allclose(Tr, expect, rtol=1e-14)
    # the end weights, point by point: the first point carries -5/8, the last
    # -1/24 (weights w_i at i and w_{2-i} at N-1-i, resample.h:591-593)
    f = np.exp(-m)
    K1 = np.full(N, 1.0); K1[0] = 0.0
    t1 = O.transmission(K1, 1, N, nTr, dTr)[0]
    np.testing.assert_allclose(t1, (3 / 8 + (N - 1 - 3 / 8) * f) / (N - 1), rtol=1e-13)
    K2 = np.full(N, 1.0); K2[-1] = 0.0
    t2 = O.transmission(K2, 1, N, nTr, dTr)[0]
    np.testing.assert_allclose(t2, (23 / 24 + (N - 1 - 1 / 2 + 1 / 6 - 5 / 8) * f) / (N - 1),
                               rtol=1e-13)


# ---------------------------------------------------------------- profiles 2-4, plinth
def _one_line_state(nu, S, ia, y, hw):
    return dict(nu=np.array([nu]), S=np.array([S]), ialphaD=np.array([ia]), y=np.array([y]),
                hw=np.array([hw]))


def test_profiles_closed_forms():
    from scipy.special import erf
    numin, dnu, Nx = 990.0, 0.001, 20001
    nu, Strue, aD, aL, hw = 1000.0123, 2.5e-20, 0.02, 0.013, 3.0
    ia = 1.0 / aD
    st = _one_line_state(nu, Strue * ia, ia, aL * ia, hw)
    x = (numin + dnu * np.arange(Nx) - nu) * ia
    inwin = np.abs(numin + dnu * np.arange(Nx) - nu) < hw
    # binned Gaussian: the erf differences telescope over the window
    K4 = O.profile_points(st, numin, dnu, Nx, 4)
    vb = dnu * ia
    lo, hi = np.nonzero(inwin)[0][[0, -1]]
    exact = Strue * ia / vb * 0.5 * (erf(x[hi] + vb / 2) - erf(x[lo] - vb / 2))
    assert abs(K4.sum() / exact - 1) < 1e-13
    assert abs(K4.sum() * dnu / Strue - 1) < 1e-13          # the window holds the whole Gaussian
    # Doppler: a sampled Gaussian, normalised
    K3 = O.profile_points(st, numin, dnu, Nx, 3)
    assert abs(K3.sum() * dnu / Strue - 1) < 1e-12
    np.testing.assert_allclose(K3[inwin], Strue * ia / math.sqrt(math.pi) * np.exp(-x[inwin] ** 2),
                               rtol=1e-15, atol=0)
    # Lorentz: the integral over the window, (2/pi) atan(hw / alpha_L), sampled
    K2 = O.profile_points(st, numin, dnu, Nx, 2)
    expect = Strue * (2 / math.pi) * math.atan(hw / aL)
    assert abs(K2.sum() * dnu / expect - 1) < 1e-6
    np.testing.assert_allclose(K2[inwin], Strue / math.pi * aL / ((x[inwin] * aD) ** 2 + aL ** 2),
                               rtol=1e-13)
    assert np.all(K2[~inwin] == 0) and np.all(K3[~inwin] == 0) and np.all(K4[~inwin] == 0)
    # Voigt through the profile path = the Voigt sum
    K1 = O.profile_points(st, numin, dnu, Nx, 1)
    np.testing.assert_array_equal(K1, O.voigt_points(st, numin, dnu, Nx))


@pytest.mark.parametrize("profile", [1, 2, 3])
def test_plinth_removal(profile):
    numin, dnu, Nx = 990.0, 0.002, 10001
    rng = np.random.default_rng(profile)
    n = 40
    nu = np.sort(rng.uniform(995.0, 1005.0, n))
    ia = 1.0 / rng.uniform(0.005, 0.05, n)
    st = dict(nu=nu, S=10 ** rng.uniform(-22, -19, n) * ia, ialphaD=ia,
              y=rng.uniform(0.01, 3.0, n) , hw=rng.uniform(0.05, 2.0, n))
    K = O.profile_points(st, numin, dnu, Nx, profile)
    Kp = O.profile_points(st, numin, dnu, Nx, profile, remove_plinth=True)
    xj = numin + dnu * np.arange(Nx)
    sub = np.zeros(Nx)
    for i in range(n):
        pl = O.profile_value(profile, st["S"][i], st["hw"][i] * ia[i], st["y"][i], dnu * ia[i])
        cover = np.abs(xj - nu[i]) < st["hw"][i]
        sub[cover] += pl
        # the plinth is the profile at the cut: a positive constant below the peak
        assert 0 <= pl < O.profile_value(profile, st["S"][i], 0.0, st["y"][i], dnu * ia[i])
    np.testing.assert_allclose(Kp, K - sub, rtol=1e-12, atol=1e-12 * np.max(K))
    # one line: the plinth-removed profile falls to ~0 at the cut
    one = {k: v[:1] for k, v in st.items()}
    K1 = O.profile_points(one, numin, dnu, Nx, profile, remove_plinth=True)
    cover = np.nonzero(np.abs(xj - nu[0]) < st["hw"][0])[0]
    edge = K1[cover[[0, -1]]]
    assert np.all(edge >= -1e-9 * K1.max()) and np.all(edge < 0.05 * K1.max())


def test_means_x_trapezoid_closed_forms():
    T = 1200.0
    # per-bin grids of 3 bins with different spacings, covering the Planck curve
    edges = [0.0, 2000.0, 9000.0, 40000.0]
    Nxb = 200000
    x = np.concatenate([edges[b] + np.arange(Nxb) * ((edges[b + 1] - edges[b]) / Nxb)
                        for b in range(3)])
    K = np.full(x.size, 2.0e-22)
    s = O.means_x(K, x, T)
    assert abs(s[0] / s[2] - 2.0e-22) <= 1e-12 * 2.0e-22
    assert abs(s[3] / s[1] - 2.0e-22) <= 1e-12 * 2.0e-22
    assert abs(s[2] / (SIGMA * T ** 4 / math.pi) - 1) <= 1e-6      # trapezoid of B, O(dx^2)
    assert abs(s[3] / (4 * SIGMA * T ** 3 / math.pi) - 1) <= 1e-6
    # on a uniform grid the trapezoid sums are the plain sums minus half the end terms, times dx
    xu = 1000.0 + 0.01 * np.arange(5001)
    Ku = 10 ** np.random.default_rng(1).uniform(-25, -20, xu.size)
    su, sp = O.means_x(Ku, xu, T), O.means(Ku, 1000.0, 0.01, T)
    B = O.means(np.ones(1), 1000.0, 1.0, T)[2], O.means(np.ones(1), xu[-1], 1.0, T)[2]
    assert abs(su[2] - (sp[2] - 0.5 * (B[0] + B[1])) * 0.01) <= 1e-12 * su[2]


# ---------------------------------------------------------------- CIA tables
def test_cia_bilinear_closed_form(tmp_path):
    """A table sampled from c(nu, T) = a + b nu + c T: linear interpolation in
    nu within a set and in T between sets reproduces it exactly.  As in the
    reference, the sweep over the points stops once a set's last table point
    is loaded, so only the first point at or above the next-to-last table
    point is filled from the last interval."""
    from heliosk import files as F
    a, b, c = 1e-46, 3e-49, 2e-48
    nu = np.concatenate([[20.0], np.sort(np.random.default_rng(7).uniform(20, 980, 40)), [980.0]])
    nu = np.round(nu, 4)
    sets = [(T, nu, a + b * nu + c * T) for T in (200.0, 300.0, 500.0)]
    p = tmp_path / "t.cia"
    F.write_cia(str(p), sets)
    x = np.linspace(0.0, 1000.0, 2001)
    rho = lambda T, P: P * 273.15 / T * 2.6867774e19
    NA, m1 = 6.0221412927e23, 4.002602
    xstar = x[np.searchsorted(x, nu[-2])]
    inside = (x >= 20.0) & (x <= xstar)
    for T, P in ((250.0, 1.0), (300.0, 2.0), (420.0, 0.5)):
        K = O.cia_K(str(p), 3, x, T, P, m1, kmin=1e-30)
        exact = np.where(inside, (a + b * x + c * T) * rho(T, P) * (NA / m1), 0.0) + 1e-30
        np.testing.assert_allclose(K, exact, rtol=1e-11)
    # below the first set: nothing is set; above the last: kmin
    assert np.all(O.cia_K(str(p), 3, x, 150.0, 1.0, m1, kmin=1e-30) == 1e-30)
    assert np.all(O.cia_K(str(p), 3, x, 600.0, 1.0, m1, kmin=1e-30) == 1e-30)


# ---------------------------------------------------------------- sub-Lorentzian wings
CHI_CO2 = [3.0, 30.0, 120.0]    # sigma1..3 of chiPH89__CO2-CO2.dat


def _chi_params(T, lo=2100.0, hi=2500.0):
    """B_i = alpha_i + beta_i exp(-epsilon_i T) with the CO2-CO2 coefficients."""
    ab = [(0.0888, -0.160, 0.00410), (0.0, 0.0526, 0.00152), (0.0232, 0.0, 0.0)]
    B = [a + b * math.exp(-e * T) for a, b, e in ab]
    return CHI_CO2 + B + [lo, hi]


def test_chi_properties():
    chi = _chi_params(296.0)
    s1, s2, s3, B1, B2, B3 = chi[:6]
    assert all(O.chi_factor(chi, d) == 1.0 for d in (0.0, 1.0, 2.999))
    eps = 1e-9
    for s in (s1, s2, s3):                         # continuous at each break
        a, b = O.chi_factor(chi, s - eps), O.chi_factor(chi, s + eps)
        assert abs(a - b) <= 1e-8 * a
    for lo, hi, B in ((s1, s2, B1), (s2, s3, B2), (s3, 400.0, B3)):   # log-slope -B_i
        d0, d1 = lo + 0.3 * (hi - lo), lo + 0.7 * (hi - lo)
        slope = (math.log(O.chi_factor(chi, d1)) - math.log(O.chi_factor(chi, d0))) / (d1 - d0)
        assert abs(slope + B) <= 1e-12 * max(B, 1e-3)


def test_sublorentzian_profile_factor():
    """One Lorentz line in the range: K_chi = K chi(|x - nu|) x / nu; outside
    the range, or for |x - nu| < sigma1, K is unchanged."""
    chi = _chi_params(500.0, lo=2200.0, hi=2300.0)
    numin, dnu, Nx = 2150.0, 0.05, 4001
    st = dict(nu=np.array([2250.0]), S=np.array([1e-20 * 30.0]), ialphaD=np.array([30.0]),
              y=np.array([3.0]), hw=np.array([90.0]))
    K = O.profile_points(st, numin, dnu, Nx, 2)
    Kc = O.profile_points(st, numin, dnu, Nx, 2, chi=chi)
    x = numin + dnu * np.arange(Nx)
    d = np.abs(x - 2250.0)
    inr = (x > 2200.0) & (x < 2300.0) & (d < 90.0)
    np.testing.assert_array_equal(Kc[~inr], K[~inr])
    near = inr & (d < 3.0)
    np.testing.assert_allclose(Kc[near], K[near] * x[near] / 2250.0, rtol=1e-15)
    # two points at the same distance on either side: the ratio is x1/x2
    j1, j2 = np.searchsorted(x, 2250.0 - 40.0), np.searchsorted(x, 2250.0 + 40.0)
    r1, r2 = Kc[j1] / K[j1], Kc[j2] / K[j2]
    assert abs((r1 / r2) / (x[j1] / x[j2]) - 1) <= 1e-13
    assert r2 < 1.0                               # wings are sub-Lorentzian


def test_compat_region_formulas_against_exact():
    """The restatement of the reference's FP32 region formulas
    (Line6fA/B/C_kernel, voigt.h:1177-1706) against the exact profile
    Re w/sqrt(pi): region B's rational approximation is off by up to 1.39e-2
    (near x = 10, y -> 0), region A's one-term wing by ~1.6e-6 and region C's
    FP32 Algorithm 916 by < 1e-6 -- the reference's own error levels, so a
    dropped or mistyped term in any region fails here; x = 0 gives
    erfcx(y)/sqrt(pi)."""
    xs = np.concatenate([np.linspace(0.0, 12.0, 97), np.geomspace(12.0, 1000.0, 120)])
    ys = np.concatenate([[1e-4, 1e-2], np.geomspace(0.05, 99.9, 40)])
    worst = {"A": 0.0, "B": 0.0, "C": 0.0}
    for y in ys:
        c = O.compat_profile(xs, np.full(xs.size, y))
        for x, v in zip(xs, c):
            xf, yf = np.float32(x), np.float32(y)
            r2 = float(xf * xf + yf * yf)
            e = O.faddeeva(float(xf), float(yf))[0].real / np.sqrt(np.pi)
            k = "A" if r2 >= 1e6 else ("B" if r2 >= 100.0 else "C")
            worst[k] = max(worst[k], abs(v - e) / e)
    print("compat worst relative error by region:", worst)
    assert 1.35e-2 <= worst["B"] <= 1.40e-2
    assert 1e-6 <= worst["A"] <= 2e-6
    assert worst["C"] <= 1e-6
    for y in (0.3, 2.0, 9.0):
        assert O.compat_profile([0.0], [y])[0] == pytest.approx(O.erfcx(y) / np.sqrt(np.pi),
                                                                rel=1e-6)
