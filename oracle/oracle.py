"""ctypes binding of the FP64 host oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
(as the checker) and by the cpu_baseline leg of bench.py.  The product path
(the HIP library and the heliosk driver) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_INT = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _bind(L):
    c_d, c_i64, c_int = ctypes.c_double, ctypes.c_int64, ctypes.c_int
    L.hko_erfcx.restype = c_d
    L.hko_erfcx.argtypes = [c_d]
    L.hko_faddeeva_batch.argtypes = [c_i64, _D, _D, _D, _D]
    L.hko_w916.argtypes = [c_d, c_d, c_d, _D, _D]
    L.hko_wcf.argtypes = [c_d, c_d, c_int, _D, _D]
    L.hko_line_params.argtypes = ([c_i64] + [_D] * 7 + [_I32] + [_D] * 4
                                  + [c_d, c_d, c_int, c_d, c_int] + [_D] * 5)
    L.hko_window.argtypes = [c_d, c_d, c_d, c_d, c_i64, _I64, _I64]
    L.hko_voigt_points.argtypes = ([c_i64] + [_D] * 5 + [c_d, c_d, c_d, c_i64, c_d,
                                                        c_i64, _I64, _D, c_int])
    L.hko_voigt_compat_points.argtypes = L.hko_voigt_points.argtypes
    L.hko_compat_profile.restype = ctypes.c_float
    L.hko_compat_profile.argtypes = [ctypes.c_float, ctypes.c_float]
    L.hko_voigt_grad.argtypes = ([c_i64] + [_D] * 7 + [_I32] + [_D] * 4
                                 + [c_d, c_d, c_int, c_d, c_int, c_d, c_d, c_i64, _D,
                                    _D, _D, _D, c_int])
    L.hko_bin_segments.argtypes = [c_d, c_d, c_d, c_i64, c_int, _I64]
    L.hko_sort_segments.argtypes = [_D, c_int, _I64]
    L.hko_qr.argtypes = [c_int, c_int, _D, _D, _D]
    L.hko_ktable.argtypes = [_D, c_i64, c_int, c_int, c_int, c_d, c_d, _D, _INT]
    L.hko_lsq.argtypes = [c_int, c_int, _D, _D]
    L.hko_reconstruct.argtypes = [_D, _INT, c_int, c_int, c_int, c_d, c_d, _D]
    L.hko_means.argtypes = [_D, c_d, c_d, c_i64, c_d, _D]
    L.hko_profile_value.restype = c_d
    L.hko_profile_value.argtypes = [c_int, c_d, c_d, c_d, c_d]
    L.hko_profile_points.argtypes = ([c_i64] + [_D] * 5 + [c_d, c_d, c_d, c_i64, c_d, c_int,
                                                          c_int, _D, _D, c_int])
    L.hko_chi.restype = c_d
    L.hko_chi.argtypes = [_D, c_d]
    L.hko_means_x.argtypes = [_D, _D, c_i64, c_d, _D]
    L.hko_transmission.argtypes = [_D, c_int, c_int, _I32, c_d, c_int, c_d, _D]
    return L


def lib():
    """The checker: built with -march=x86-64-v3 -ffp-contract=off."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


_native = None


def lib_native():
    """The same source built for the CPU this runs on (-O3 -march=native
    -fopenmp, the compiler's default contraction), compiled on first use
    into a directory keyed by the host's CPU model: the timed CPU baseline
    only.  Returns (library, build description)."""
    global _native
    if _native is None:
        import hashlib
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith(("model name", "flags")):
                        model += line
                        if line.startswith("flags"):
                            break
        except OSError:
            pass
        tag = hashlib.sha1(model.encode()).hexdigest()[:12]
        out = os.path.join(_HERE, "_build", "native-" + tag)
        path = os.path.join(out, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE, "native", "NATIVE_OUT=" + out], check=True)
        _native = (_bind(ctypes.CDLL(path)),
                   "gcc -O3 -march=native -fopenmp (built on this host, %s)" % tag)
    return _native


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def erfcx(y):
    return lib().hko_erfcx(float(y))


def faddeeva(x, y):
    """Complex w(x + iy) for y >= 0 (arrays)."""
    x = _f64(np.atleast_1d(x)); y = _f64(np.broadcast_to(y, x.shape))
    re = np.empty_like(x); im = np.empty_like(x)
    lib().hko_faddeeva_batch(x.size, _p(x), _p(y), _p(re), _p(im))
    return re + 1j * im


def w916(x, y, tol=1e-17):
    re = ctypes.c_double(); im = ctypes.c_double()
    lib().hko_w916(float(x), float(y), tol, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def wcf(x, y, depth):
    re = ctypes.c_double(); im = ctypes.c_double()
    lib().hko_wcf(float(x), float(y), int(depth), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def _iso_arrays(lines):
    iso = lines.get("iso")
    iso = None if iso is None else np.ascontiguousarray(iso, dtype=np.int32)
    tabs = [_f64(lines[k]) for k in ("iso_mass", "iso_ab", "iso_sscale", "iso_Q")]
    return iso, tabs


def line_params(lines, T, P, cut_mode, cut, form=0, L=None):
    """Line state at (T, P[atm]): dict nu, S, ialphaD, y, hw.  form: 0 HITRAN
    (S2_kernel), 1 ExoMol family (Sf_kernel), 2 super-lines (SfSuper_kernel)."""
    nl = len(lines["nu"])
    cols = [_f64(lines[k]) for k in ("nu", "S", "EL", "A", "delta", "vy", "n")]
    iso, tabs = _iso_arrays(lines)
    out = {k: np.empty(nl) for k in ("nu", "S", "ialphaD", "y", "hw")}
    (L or lib()).hko_line_params(nl, *[_p(c) for c in cols],
                                 None if iso is None else _p(iso, _I32),
                                 *[_p(t) for t in tabs], float(T), float(P), int(cut_mode),
                                 float(cut), int(form),
                                 *[_p(out[k]) for k in ("nu", "S", "ialphaD", "y", "hw")])
    return out


def profile_points(state, numin, dnu, Nx, profile, remove_plinth=False, kmin=0.0, nthreads=0,
                   chi=None):
    """K on all Nx points for profile 1 (Voigt), 2 (Lorentz), 3 (Doppler) or
    4 (binned Gaussian), optionally with each line's plinth removed and the
    sub-Lorentzian factor chi = (sigma1..3, B1..3, range_lo, range_hi)."""
    cols = [_f64(state[k]) for k in ("nu", "S", "ialphaD", "y", "hw")]
    if cols[0].size > 1 and np.any(np.diff(cols[0]) < 0):
        o = np.argsort(cols[0], kind="stable")
        cols = [np.ascontiguousarray(c[o]) for c in cols]
    nl = cols[0].size
    hw = cols[4]
    R = float(hw[np.isfinite(hw)].max()) if nl else 0.0
    K = np.empty(int(Nx))
    lib().hko_profile_points(nl, *[_p(c) for c in cols], R, float(numin), float(dnu), int(Nx),
                             float(kmin), int(profile), int(bool(remove_plinth)),
                             None if chi is None else _p(_f64(chi)), _p(K), int(nthreads))
    return K


def chi_factor(chi, d):
    return lib().hko_chi(_p(_f64(chi)), float(d))


def profile_value(profile, S, x, y, vb):
    return lib().hko_profile_value(int(profile), float(S), float(x), float(y), float(vb))


def window(nu, hw, numin, dnu, Nx):
    lo = ctypes.c_int64(); hi = ctypes.c_int64()
    lib().hko_window(float(nu), float(hw), float(numin), float(dnu), int(Nx),
                     ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def voigt_points(state, numin, dnu, Nx, kmin=0.0, jidx=None, search_radius=None,
                 nthreads=0, L=None):
    """K at grid points jidx (all points if None).  The C routine searches
    the lines by centre, so a state whose (pressure-shifted) centres are not
    ascending is put in order first."""
    cols = [_f64(state[k]) for k in ("nu", "S", "ialphaD", "y", "hw")]
    if cols[0].size > 1 and np.any(np.diff(cols[0]) < 0):
        o = np.argsort(cols[0], kind="stable")
        cols = [np.ascontiguousarray(c[o]) for c in cols]
    nl = cols[0].size
    if search_radius is None:
        hw = cols[4]
        search_radius = float(hw[np.isfinite(hw)].max()) if nl else 0.0
    if jidx is None:
        npts = int(Nx); jp = None
    else:
        jidx = np.ascontiguousarray(jidx, dtype=np.int64); npts = jidx.size
        jp = _p(jidx, _I64)
    K = np.empty(npts)
    (L or lib()).hko_voigt_points(nl, *[_p(c) for c in cols], float(search_radius),
                                  float(numin), float(dnu), int(Nx), float(kmin), npts, jp,
                                  _p(K), int(nthreads))
    return K


def voigt_compat_points(state, numin, dnu, Nx, kmin=0.0, jidx=None, nthreads=0, L=None):
    """K at grid points jidx (all if None) in the reference's own FP32
    arithmetic (hko_voigt_compat_points), lines in centre order."""
    cols = [_f64(state[k]) for k in ("nu", "S", "ialphaD", "y", "hw")]
    if cols[0].size > 1 and np.any(np.diff(cols[0]) < 0):
        o = np.argsort(cols[0], kind="stable")
        cols = [np.ascontiguousarray(c[o]) for c in cols]
    nl = cols[0].size
    hw = cols[4]
    R = float(hw[np.isfinite(hw)].max()) if nl else 0.0
    if jidx is None:
        npts = int(Nx); jp = None
    else:
        jidx = np.ascontiguousarray(jidx, dtype=np.int64); npts = jidx.size
        jp = _p(jidx, _I64)
    K = np.empty(npts)
    (L or lib()).hko_voigt_compat_points(nl, *[_p(c) for c in cols], R, float(numin),
                                         float(dnu), int(Nx), float(kmin), npts, jp, _p(K),
                                         int(nthreads))
    return K


def compat_profile(x, y):
    """The reference's FP32 region formulas for a unit line (no cut), per
    (x, y) pair: Re w(x, y)/sqrt(pi) as its kernels approximate it."""
    x = np.asarray(x, dtype=np.float32).ravel(); y = np.asarray(y, dtype=np.float32).ravel()
    f = lib().hko_compat_profile
    return np.array([f(float(a), float(b)) for a, b in zip(x, y)], dtype=np.float64)


def voigt_grad(lines, T, P, cut_mode, cut, numin, dnu, Nx, gK, nthreads=0, form=0):
    """(dL/dT, dL/dP, dL/dS_i, dL/dvy_i) for L = sum_j gK_j K_j."""
    nl = len(lines["nu"])
    cols = [_f64(lines[k]) for k in ("nu", "S", "EL", "A", "delta", "vy", "n")]
    iso, tabs = _iso_arrays(lines)
    gK = _f64(gK)
    gS = np.empty(nl); gv = np.empty(nl); gTP = np.empty(2)
    lib().hko_voigt_grad(nl, *[_p(c) for c in cols],
                         None if iso is None else _p(iso, _I32),
                         *[_p(t) for t in tabs], float(T), float(P), int(cut_mode),
                         float(cut), int(form), float(numin), float(dnu), int(Nx), _p(gK),
                         _p(gS), _p(gv), _p(gTP), int(nthreads))
    return gTP[0], gTP[1], gS, gv


def bin_segments(numin, numax, dnu, Nx, nbins):
    seg = np.empty(nbins + 2, dtype=np.int64)
    lib().hko_bin_segments(float(numin), float(numax), float(dnu), int(Nx), int(nbins),
                           _p(seg, _I64))
    return seg


def sort_segments(K, seg, L=None):
    K = _f64(K).copy()
    seg = np.ascontiguousarray(seg, dtype=np.int64)
    (L or lib()).hko_sort_segments(_p(K), int(seg.size - 1), _p(seg, _I64))
    return K


def qr(Nxb, nC):
    V = np.empty(Nxb * nC); C = np.empty(nC); D = np.empty(nC)
    lib().hko_qr(int(Nxb), int(nC), _p(V), _p(C), _p(D))
    return V.reshape(nC, Nxb), C, D


def ktable(Ksorted, Nxb, nbins, nC, kmin, unit_scale=1.0, L=None):
    """Coefficients [nbins, nC] and Nxmin [nbins] from a per-bin sorted K."""
    K = _f64(Ksorted).copy()
    coeff = np.empty(nbins * nC); nxmin = np.empty(nbins, dtype=np.intc)
    (L or lib()).hko_ktable(_p(K), K.size, int(Nxb), int(nbins), int(nC), float(kmin),
                            float(unit_scale), _p(coeff), _p(nxmin, _INT))
    return coeff.reshape(nbins, nC), nxmin.astype(np.int64)


def lsq(b, nC):
    """leastSquare_kernel alone on an arbitrary right-hand side b."""
    b = _f64(b); c = np.empty(nC)
    lib().hko_lsq(int(b.size), int(nC), _p(b), _p(c))
    return c


def reconstruct(coeff, nxmin, Nxb, kmin, unit_scale=1.0):
    coeff = _f64(coeff); nbins, nC = coeff.shape
    nx = np.ascontiguousarray(nxmin, dtype=np.intc)
    Ky = np.empty(nbins * Nxb)
    lib().hko_reconstruct(_p(coeff), _p(nx, _INT), int(Nxb), int(nbins), int(nC),
                          float(kmin), float(unit_scale), _p(Ky))
    return Ky


def means(K, numin, dnu, T):
    """[sum B K, sum_{K!=0} dB/K, sum B, sum dB] on x_j = fma(j, dnu, numin)."""
    K = _f64(K)
    out = np.empty(4)
    lib().hko_means(_p(K), float(numin), float(dnu), int(K.size), float(T), _p(out))
    return out


def transmission(Ks, nbins, Nxb, nTr, dTr, kmin=0.0, nxmin=None):
    """Tr[b, t] of nbins sorted bins of Nxb values."""
    Ks = _f64(Ks)
    assert Ks.size >= nbins * Nxb
    nx = None if nxmin is None else np.ascontiguousarray(nxmin, dtype=np.int32)
    Tr = np.empty(nbins * nTr)
    lib().hko_transmission(_p(Ks), int(nbins), int(Nxb), None if nx is None else _p(nx, _I32),
                           float(kmin), int(nTr), float(dTr), _p(Tr))
    return Tr.reshape(nbins, nTr)


def means_x(K, x, T):
    """Trapezoid mean sums on explicit points x."""
    K = _f64(K); x = _f64(x)
    assert K.size == x.size
    out = np.empty(4)
    lib().hko_means_x(_p(K), _p(x), int(K.size), float(T), _p(out))
    return out


def cia_K(path, nsets, x, T, P, mass1, kmin=0.0):
    """K of a HITRAN CIA table on the points x at (T, P): readCiaFile
    (host.h:670-789) restated -- sets in file order, linear in nu within a
    set, replaced by a set with T_set <= T or blended linearly in T from the
    previous set when T < T_set, stopping at the first set with T <= T_set;
    then K rho N_A / mass1 + kmin, rho = P 273.15 / T amagat, or kmin above
    every set.  Pure Python: for small tables."""
    toks = []
    with open(path) as f:
        lines = f.read().split("\n")
    K = [0.0] * len(x)
    li = 0
    Tc1 = 0.0
    for _ in range(nsets):
        Tc0 = Tc1
        h = lines[li]; li += 1
        ncia = int(h[40:47]); Tc1 = float(h[47:54]); numax = float(h[30:40])
        pts = []
        for _k in range(ncia):
            a, b = lines[li].split(); li += 1
            pts.append((float(a), float(b)))
        nu0, a0 = pts[0]; nu1, a1 = pts[1]
        nc = 2
        for i, xi in enumerate(x):
            while xi >= nu1:
                nu0, a0 = nu1, a1
                nu1, a1 = pts[nc]
                nc += 1
                if nc == ncia:
                    break
            if nu0 <= xi < nu1:
                k1 = a0 + (a1 - a0) * (xi - nu0) / (nu1 - nu0)
                if Tc1 <= T:
                    K[i] = k1
                if T < Tc1 and Tc0 > 0.0:
                    K[i] = K[i] + (k1 - K[i]) * (T - Tc0) / (Tc1 - Tc0)
            if nu1 == numax:
                break
        if T <= Tc1:
            break
    rho = P * 273.15 / T * 2.6867774e19
    NA = 6.0221412927e23
    return np.array([kmin if Tc1 < T else k * rho * (NA / mass1) + kmin for k in K])
