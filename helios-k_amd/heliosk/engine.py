"""Python host side over the C ABI, with torch tensors as device memory.

torch provides the device allocations and the stream; all computation runs in
the HIP kernels of libheliosk.so.  The functions mirror the reference's
stages: line parameters (S2_kernel), the Voigt sum (Line6A/B/C_Call), the
k-table (heliosk.cu:2046-2193) and the adjoint.
"""
import copy
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import (HK_FORM_EXOMOL, HK_FORM_HITRAN, HK_FORM_SUPERLINE, HK_PROFILE_BINNED_GAUSS,
                   HK_PROFILE_DOPPLER, HK_PROFILE_LORENTZ, HK_PROFILE_VOIGT, HK_VOIGT_FP32C,
                   HK_VOIGT_FP64, HK_VOIGT_COMPAT, HK_MAX_ISO, Cut, Grid, LineState, Lines, Spectrum, TP, Tune,
                   check, lib)

NA = 6.0221412927e23          # define.h:42

__all__ = ["HK_FORM_HITRAN", "HK_FORM_EXOMOL", "HK_FORM_SUPERLINE", "Species", "sort_lines",
           "State", "make_grid", "line_params", "voigt", "voigt_batched", "VoigtBatch",
           "autotune", "ktable", "KTable", "KTableFit", "ktable_fit", "VoigtOpacity", "opacity",
           "binned_log_mean",
           "ktable_reconstruct", "adjoint", "synth_lines", "means", "means_x", "sort_bins", "transmission", "profile", "HK_VOIGT_FP64", "HK_VOIGT_FP32C",
           "HK_VOIGT_COMPAT"]


def _dev():
    if not torch.cuda.is_available():
        raise _lib.HKError("no GPU visible: the heliosk engine runs only on the HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _keep(stream, *tensors):
    """Tensors that work queued on `stream` reads or writes: when that stream
    is not torch's current stream, the caching allocator is told to wait for
    it before handing their memory to new work (a transient workspace freed
    on return would otherwise be reused while the kernels still run)."""
    if stream is None or stream == torch.cuda.current_stream():
        return
    for t in tensors:
        if t is not None:
            t.record_stream(stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def make_grid(numin, numax, dnu):
    """Grid of the reference (heliosk.cu:448): Nx = (int)((numax-numin)/dnu + 0.5)."""
    nx = int((numax - numin) / dnu + 0.5)
    return Grid(float(numin), float(dnu), nx)


class Species:
    """Raw line list of one species on the device, sorted by nu (stable).
    form: HK_FORM_HITRAN, HK_FORM_EXOMOL or HK_FORM_SUPERLINE (how the
    strengths and widths depend on T and P)."""

    def __init__(self, nu, S, EL, vy, n, A=None, delta=None, iso=None, iso_mass=(18.0 / NA,),
                 iso_ab=(1.0,), iso_sscale=(1.0,), name="species", form=HK_FORM_HITRAN):
        nu = np.asarray(nu, dtype=np.float64)
        order = np.argsort(nu, kind="stable")
        dev = _dev()
        nl = nu.size

        def up(a, fill=0.0):
            a = np.full(nl, fill) if a is None else np.asarray(a, dtype=np.float64)
            return torch.from_numpy(np.ascontiguousarray(a[order])).to(dev)

        self.name = name
        self.form = int(form)
        self.order = order
        self.nu = up(nu); self.S = up(S); self.EL = up(EL); self.vy = up(vy); self.n = up(n)
        self.A = up(A); self.delta = up(delta)
        self.iso = None if iso is None else torch.from_numpy(
            np.ascontiguousarray(np.asarray(iso, dtype=np.int32)[order])).to(dev)
        self.iso_mass = list(iso_mass); self.iso_ab = list(iso_ab)
        self.iso_sscale = list(iso_sscale)
        self.max_abs_delta = float(np.max(np.abs(delta))) if delta is not None and nl else 0.0
        self.nlines = nl

    def with_fields(self, **fields):
        """A shallow copy whose per-line device arrays (S, vy, ...) are
        replaced by the given tensors (float64, on the device, sorted order)."""
        c = copy.copy(self)
        for k, v in fields.items():
            if k not in ("nu", "S", "EL", "vy", "n", "A", "delta"):
                raise KeyError(k)
            v = v.detach().to(dtype=torch.float64).contiguous()
            if v.numel() != self.nlines or v.device != self.nu.device:
                raise ValueError("%s: %d values on %s, species has %d on %s"
                                 % (k, v.numel(), v.device, self.nlines, self.nu.device))
            setattr(c, k, v)
        return c

    def lines(self):
        L = Lines()
        L.nlines = self.nlines
        L.nu = self.nu.data_ptr(); L.S = self.S.data_ptr(); L.EL = self.EL.data_ptr()
        L.A = self.A.data_ptr(); L.delta = self.delta.data_ptr(); L.vy = self.vy.data_ptr()
        L.n = self.n.data_ptr()
        L.iso = self.iso.data_ptr() if self.iso is not None else None
        L.niso = len(self.iso_mass)
        L.form = self.form
        for k in range(L.niso):
            L.iso_mass[k] = self.iso_mass[k]; L.iso_ab[k] = self.iso_ab[k]
            L.iso_sscale[k] = self.iso_sscale[k]
        L.max_abs_delta = self.max_abs_delta
        return L


def sort_lines(nu, S, EL, vy, n, A=None, delta=None, iso=None, iso_mass=(18.0 / NA,),
               iso_ab=(1.0,), iso_sscale=(1.0,), name="species", form=HK_FORM_HITRAN,
               stream=None):
    """Species from an unsorted line list already on the device, ordered by
    the library's stable radix sort (hk_lines_sort; the thrust sort and
    gathers of heliosk.cu:1147-1171).  Per-line arguments are device
    tensors (float64; iso int32) or array-likes, which are uploaded."""
    dev = _dev()

    def on_dev(a, dtype=torch.float64):
        t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
        return t.to(device=dev, dtype=dtype).contiguous()

    raw = {k: on_dev(v) for k, v in (("nu", nu), ("S", S), ("EL", EL), ("vy", vy), ("n", n))}
    nl = raw["nu"].numel()
    raw["A"] = on_dev(A) if A is not None else torch.zeros(nl, dtype=torch.float64, device=dev)
    raw["delta"] = (on_dev(delta) if delta is not None
                    else torch.zeros(nl, dtype=torch.float64, device=dev))
    raw_iso = on_dev(iso, torch.int32) if iso is not None else None
    out = {k: torch.empty_like(v) for k, v in raw.items()}
    out_iso = torch.empty_like(raw_iso) if raw_iso is not None else None
    sp = object.__new__(Species)
    sp.name = name; sp.form = int(form); sp.order = None
    sp.iso_mass = list(iso_mass); sp.iso_ab = list(iso_ab); sp.iso_sscale = list(iso_sscale)
    sp.nlines = nl
    sp.max_abs_delta = float(raw["delta"].abs().max()) if nl else 0.0
    for k, v in raw.items():
        setattr(sp, k, v)
    sp.iso = raw_iso
    Lin = sp.lines()
    for k, v in out.items():
        setattr(sp, k, v)
    sp.iso = out_iso
    Lout = sp.lines()
    nbytes = lib().hk_lines_sort_workspace_bytes(nl)
    if nbytes == 0:
        raise _lib.HKError("hk_lines_sort_workspace_bytes(%d) failed" % nl)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    check(lib().hk_lines_sort(ctypes.byref(Lin), ctypes.byref(Lout), _ptr(ws), nbytes,
                              _stream(stream)))
    _keep(stream, ws, raw_iso, out_iso, *raw.values(), *out.values())
    return sp


class State:
    """Line state at one (T, P): one device allocation laid out by the library."""

    def __init__(self, nlines):
        nbytes = lib().hk_line_state_bytes(nlines)
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
        self.st = LineState()
        check(lib().hk_line_state_init(ctypes.byref(self.st), nlines, _ptr(self.buf), nbytes))
        self.nlines = nlines

    def array(self, field):
        """Device view of one per-line field as a torch tensor (copy)."""
        addr = getattr(self.st, field)
        off = addr - self.buf.data_ptr()
        return self.buf[off:off + 8 * self.nlines].view(torch.float64).clone()


def _tp(T, P, Q):
    tp = TP()
    tp.T = float(T); tp.P = float(P)
    Q = [1.0] if Q is None else list(Q)
    for k in range(HK_MAX_ISO):
        tp.Q[k] = Q[k] if k < len(Q) else 1.0
    return tp


def line_params(species, T, P, Q=None, cut_mode=0, cut=0.0, state=None, stream=None):
    """S2_kernel semantics at (T [K], P [atm]); returns a State."""
    st = state or State(species.nlines)
    L = species.lines()
    tp = _tp(T, P, Q)
    c = Cut(int(cut_mode), float(cut))
    check(lib().hk_line_params(ctypes.byref(L), ctypes.byref(tp), ctypes.byref(c),
                               ctypes.byref(st.st), _stream(stream)))
    st.species = species        # st.st.key points into species.nu: keep it alive
    return st


def _tune(tile_points=0, exact_pairs=False, stage_lines=0):
    return Tune(int(tile_points), int(bool(exact_pairs)), int(stage_lines))


def voigt(state, grid, kmin=0.0, mode=HK_VOIGT_FP32C, tile_points=0, K=None, accumulate=False,
          stream=None, workspace=None, exact_pairs=False, stage_lines=0):
    """K_j = kmin + sum_i S_i Re w / sqrt(pi) on the grid (or K += sum).
    exact_pairs: FP32C evaluates every pair instead of interpolating the far
    field of distant lines; stage_lines: lines per LDS stage (256 or 512)."""
    tune = _tune(tile_points, exact_pairs, stage_lines)
    if K is None:
        K = torch.empty(grid.nx, dtype=torch.float64, device=_dev())
    nbytes = lib().hk_voigt_workspace_bytes(ctypes.byref(grid), ctypes.byref(tune))
    ws = workspace if workspace is not None else torch.empty(nbytes, dtype=torch.uint8, device=_dev())
    check(lib().hk_voigt_accumulate(ctypes.byref(state.st), ctypes.byref(grid), float(kmin),
                                    int(bool(accumulate)), int(mode), ctypes.byref(tune),
                                    _ptr(K), _ptr(ws), ws.numel(), _stream(stream)))
    _keep(stream, K, ws)
    return K


def profile(state, grid, profile, remove_plinth=False, kmin=0.0, voigt_mode=HK_VOIGT_FP64, K=None,
            accumulate=False, chi=None, stream=None):
    """K_j = kmin + sum_i profile_i(x_j) (minus each line's plinth) for
    profile 1 (Voigt), 2 (Lorentz), 3 (Doppler) or 4 (binned Gaussian);
    chi = (sigma1..3, B1..3, range_lo, range_hi) applies sub-Lorentzian wings."""
    c = None
    if chi is not None:
        c = _lib.Chi()
        for k in range(3):
            c.sigma[k] = float(chi[k]); c.B[k] = float(chi[3 + k])
        c.range[0] = float(chi[6]); c.range[1] = float(chi[7])
    if K is None:
        K = torch.empty(grid.nx, dtype=torch.float64, device=_dev())
    nbytes = lib().hk_profile_workspace_bytes(ctypes.byref(grid), state.nlines)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
    check(lib().hk_profile_accumulate(ctypes.byref(state.st), ctypes.byref(grid), float(kmin),
                                      int(bool(accumulate)), int(profile), int(bool(remove_plinth)),
                                      int(voigt_mode), None if c is None else ctypes.byref(c),
                                      _ptr(K), _ptr(ws), nbytes, _stream(stream)))
    _keep(stream, K, ws)
    return K


def work_count(state, grid, mode=HK_VOIGT_FP32C, tile_points=0, exact_pairs=False, stream=None):
    """(direct A, B, C pairs, far-field pairs, far (line, tile) pairs of the
    9-, 15- and 19-node bands, far (line, wave) pairs of the wave-level
    bands, cut-edge (line, wave) pairs and their point pairs) of one Voigt
    sum with this form and tiling."""
    c = torch.zeros(12, dtype=torch.int64, device=_dev())
    tune = _tune(tile_points, exact_pairs)
    check(lib().hk_voigt_work_count(ctypes.byref(state.st), ctypes.byref(grid), int(mode),
                                    ctypes.byref(tune), _ptr(c), _stream(stream)))
    return [int(v) for v in c.cpu().tolist()]


def pair_count(state, grid, stream=None):
    """(evaluations, region A, region B, region C) of one Voigt sum."""
    out = torch.zeros(4, dtype=torch.int64, device=_dev())
    check(lib().hk_voigt_pair_count(ctypes.byref(state.st), ctypes.byref(grid), _ptr(out),
                                    _stream(stream)))
    return [int(v) for v in out.cpu()]


def _spectra(items):
    arr = (Spectrum * len(items))()
    for i, (st, g, kmin, K) in enumerate(items):
        arr[i].st = st.st; arr[i].grid = g; arr[i].kmin = float(kmin)
        arr[i].accumulate = 0; arr[i].K = K.data_ptr()
    return arr


def autotune(state, grid, mode=HK_VOIGT_FP32C, stream=None, exact_pairs=False):
    """Fastest (tile width, lines per LDS stage) of the main pass for this
    input (hk_voigt_autotune; the library keeps the choice per problem
    signature for the process).  Returns (tile_points, stage_lines,
    {(tile, stage): ms}) -- ms 0 for a cached choice."""
    K = torch.empty(grid.nx, dtype=torch.float64, device=_dev())
    tune = _tune(0)
    nbytes = lib().hk_voigt_workspace_bytes(ctypes.byref(grid), ctypes.byref(tune))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
    best = _tune(0, exact_pairs)
    ms = (ctypes.c_float * 8)()
    check(lib().hk_voigt_autotune(ctypes.byref(state.st), ctypes.byref(grid), int(mode), _ptr(K),
                                  _ptr(ws), nbytes, ctypes.byref(best), ms, _stream(stream)))
    keys = [(t, sl) for sl in (256, 512) for t in (512, 1024, 2048, 4096)]
    return best.tile_points, best.stage_lines, dict(zip(keys, list(ms)))


def voigt_batched(items, mode=HK_VOIGT_FP32C, tile_points=0, stream=None):
    """items: list of (State, Grid, kmin, K tensor).  One persistent launch
    per pass over all spectra."""
    n = len(items)
    arr = _spectra(items)
    tune = _tune(tile_points)
    nbytes = lib().hk_voigt_batched_workspace_bytes(n, arr, ctypes.byref(tune))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
    check(lib().hk_voigt_batched(n, arr, int(mode), ctypes.byref(tune), _ptr(ws), nbytes,
                                 _stream(stream)))
    _keep(stream, ws)


class VoigtBatch:
    """A prepared batch of spectra: launch() issues kernels only, so it can be
    captured in a torch.cuda.CUDAGraph (a hipGraph) and replayed."""

    def __init__(self, items, mode=HK_VOIGT_FP32C, tile_points=0, stream=None,
                 exact_pairs=False, stage_lines=0):
        self.items = list(items)            # keeps the states and K tensors alive
        n = len(self.items)
        arr = _spectra(self.items)
        tune = _tune(tile_points, exact_pairs, stage_lines)
        nbytes = lib().hk_voigt_batched_workspace_bytes(n, arr, ctypes.byref(tune))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
        self.batch = _lib.VoigtBatch()
        check(lib().hk_voigt_batched_prepare(n, arr, int(mode), ctypes.byref(tune),
                                             _ptr(self.ws), nbytes, ctypes.byref(self.batch),
                                             _stream(stream)))

    def launch(self, stream=None):
        check(lib().hk_voigt_batched_launch(ctypes.byref(self.batch), _stream(stream)))


class KTable:
    """Prepared k-table workspace for one (grid, numax, nbins, nC, nspec);
    backward=True sizes it for hk_ktable_backward as well."""

    def __init__(self, grid, numax, nbins, nC, nspec=1, backward=False):
        self.grid, self.numax, self.nbins, self.nC, self.nspec = grid, float(numax), nbins, nC, nspec
        size = lib().hk_ktable_backward_workspace_bytes if backward else lib().hk_ktable_workspace_bytes
        nbytes = size(grid.nx, nbins, nC, nspec)
        if nbytes == 0:
            raise _lib.HKError("k-table workspace size: unsupported (nx %d, nbins %d, nC %d)"
                               % (grid.nx, nbins, nC))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
        check(lib().hk_ktable_prepare(ctypes.byref(grid), self.numax, nbins, nC, _ptr(self.ws),
                                      nbytes))
        self.Nxb = grid.nx // nbins

    def __call__(self, K, kmin=0.0, unit_scale=1.0, want_sorted=False, stream=None):
        nspec = K.numel() // self.grid.nx
        assert nspec <= self.nspec
        dev = K.device
        coeff = torch.empty(nspec * self.nbins * self.nC, dtype=torch.float64, device=dev)
        nxmin = torch.empty(nspec * self.nbins, dtype=torch.int32, device=dev)
        Ks = torch.empty_like(K) if want_sorted else None
        check(lib().hk_ktable(_ptr(K), nspec, ctypes.byref(self.grid), self.numax, self.nbins,
                              self.nC, float(kmin), float(unit_scale), _ptr(coeff), _ptr(nxmin),
                              _ptr(Ks), _ptr(self.ws), self.ws.numel(), _stream(stream)))
        _keep(stream, coeff, nxmin, Ks, self.ws)
        coeff = coeff.view(nspec, self.nbins, self.nC)
        nxmin = nxmin.view(nspec, self.nbins)
        return (coeff, nxmin, Ks) if want_sorted else (coeff, nxmin)


    def backward(self, K, Ksorted, nxmin, gcoeff, stream=None):
        """dL/dK (shaped like K) from dL/dc (hk_ktable_backward): K, Ksorted
        and nxmin are the input and outputs of the forward call."""
        nspec = K.numel() // self.grid.nx
        gc = gcoeff.to(torch.float64).contiguous()
        Kc, Ksc, nxc = K.contiguous(), Ksorted.contiguous(), nxmin.contiguous()
        gK = torch.empty_like(Kc)
        check(lib().hk_ktable_backward(_ptr(Kc), _ptr(Ksc), _ptr(nxc), nspec,
                                       ctypes.byref(self.grid), self.numax, self.nbins, self.nC,
                                       _ptr(gc), _ptr(gK), _ptr(self.ws), self.ws.numel(),
                                       _stream(stream)))
        _keep(stream, gc, Kc, Ksc, nxc, gK, self.ws)
        return gK.view_as(K)


class KTableFit(torch.autograd.Function):
    """k-table coefficients (nspec, nbins, nC) of K as a differentiable
    function; backward runs hk_ktable_backward (through the fit, ln and the
    rescale; Nxmin and the per-bin order held fixed).  kt: a KTable made with
    backward=True.  Returns (coeff, nxmin); nxmin is not differentiable."""

    @staticmethod
    def forward(ctx, K, kt, kmin=0.0, unit_scale=1.0):
        coeff, nxmin, Ks = kt(K.detach(), kmin, unit_scale, want_sorted=True)
        ctx.save_for_backward(K.detach(), Ks, nxmin)
        ctx.kt = kt
        ctx.mark_non_differentiable(nxmin)
        return coeff, nxmin

    @staticmethod
    def backward(ctx, gcoeff, gnxmin):
        K, Ks, nxmin = ctx.saved_tensors
        return ctx.kt.backward(K, Ks, nxmin, gcoeff), None, None, None


def ktable_fit(K, kt, kmin=0.0, unit_scale=1.0):
    """Differentiable k-table coefficients, see KTableFit."""
    return KTableFit.apply(K, kt, kmin, unit_scale)


def ktable(K, grid, numax, nbins, nC, kmin=0.0, unit_scale=1.0, want_sorted=False):
    kt = KTable(grid, numax, nbins, nC, nspec=max(1, K.numel() // grid.nx))
    return kt(K, kmin, unit_scale, want_sorted)


def ktable_reconstruct(coeff, nxmin, Nxb, kmin=0.0, unit_scale=1.0, stream=None):
    nspec, nbins, nC = coeff.shape
    Ky = torch.empty(nspec * nbins * Nxb, dtype=torch.float64, device=coeff.device)
    scratch = torch.empty_like(Ky)
    cc, nc = coeff.contiguous(), nxmin.contiguous()
    check(lib().hk_ktable_reconstruct(_ptr(cc), _ptr(nc), nspec, nbins, Nxb, nC, float(kmin),
                                      float(unit_scale), _ptr(Ky), _ptr(scratch),
                                      _stream(stream)))
    _keep(stream, cc, nc, Ky, scratch)
    return Ky


def ktable_plan(Nxb, nC):
    P = np.empty(nC * 2 * nC)
    check(lib().hk_ktable_plan_host(int(Nxb), int(nC), P.ctypes.data_as(_lib.P_d)))
    return P.reshape(nC, 2 * nC)


def adjoint(species, T, P, gK, grid, Q=None, cut_mode=0, cut=0.0, stream=None):
    """(dL/dT, dL/dP, dL/dS_i, dL/dvy_i) for L = sum_j gK_j K_j, in the
    species' sorted line order."""
    L = species.lines()
    tp = _tp(T, P, Q)
    c = Cut(int(cut_mode), float(cut))
    dev = gK.device
    gS = torch.empty(species.nlines, dtype=torch.float64, device=dev)
    gv = torch.empty_like(gS)
    gTP = torch.empty(2, dtype=torch.float64, device=dev)
    nbytes = lib().hk_voigt_adjoint_workspace_bytes(species.nlines)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    gKc = gK.contiguous()
    check(lib().hk_voigt_adjoint(ctypes.byref(L), ctypes.byref(tp), ctypes.byref(c),
                                 ctypes.byref(grid), _ptr(gKc), _ptr(gS), _ptr(gv),
                                 _ptr(gTP), _ptr(ws), nbytes, _stream(stream)))
    _keep(stream, gKc, gS, gv, gTP, ws)
    return gTP, gS, gv


class VoigtOpacity(torch.autograd.Function):
    """K(T, P, S, vy) on the grid (plain FP64 form) as a differentiable
    function; backward runs the hand-written adjoint kernel, which
    recomputes the profile instead of storing it.

    forward(T, P, S, vy, species, grid, Q, cut_mode, cut, kmin): T [K] and
    P [atm] are 0-dim float64 tensors, S (= S_file) and vy per line in the
    species' sorted order.  The partition function Q is held fixed."""

    @staticmethod
    def forward(ctx, T, P, S, vy, species, grid, Q=None, cut_mode=0, cut=0.0, kmin=0.0):
        sp = species.with_fields(S=S, vy=vy)
        st = line_params(sp, float(T), float(P), Q=Q, cut_mode=cut_mode, cut=cut)
        K = voigt(st, grid, kmin, HK_VOIGT_FP64)
        ctx.save_for_backward(T, P, S, vy)
        ctx.args = (species, grid, Q, cut_mode, cut)
        return K

    @staticmethod
    def backward(ctx, gK):
        T, P, S, vy = ctx.saved_tensors
        species, grid, Q, cut_mode, cut = ctx.args
        sp = species.with_fields(S=S, vy=vy)
        gTP, gS, gv = adjoint(sp, float(T), float(P), gK.to(torch.float64).contiguous(), grid,
                              Q=Q, cut_mode=cut_mode, cut=cut)
        gT = gTP[0].to(T.device).reshape(T.shape)
        gP = gTP[1].to(P.device).reshape(P.shape)
        return gT, gP, gS, gv, None, None, None, None, None, None


def opacity(T, P, S, vy, species, grid, Q=None, cut_mode=0, cut=0.0, kmin=0.0):
    """Differentiable cross section, see VoigtOpacity."""
    return VoigtOpacity.apply(T, P, S, vy, species, grid, Q, cut_mode, cut, kmin)


def binned_log_mean(K, nbins, w):
    """sum_b w_b log(mean of K over bin b) for nbins equal bins of the grid."""
    return torch.sum(w * torch.log(K.view(nbins, -1).mean(dim=1)))


def means(K, grid, T, stream=None):
    """[sum B K, sum_{K!=0} dB/K, sum B, sum dB] of one spectrum on the grid
    (device tensor of 4); Planck mean = s[0]/s[2], Rosseland = s[3]/s[1]."""
    K = K.contiguous()
    sums = torch.empty(4, dtype=torch.float64, device=K.device)
    nbytes = lib().hk_means_workspace_bytes()
    ws = torch.empty(nbytes, dtype=torch.uint8, device=K.device)
    check(lib().hk_means(_ptr(K), ctypes.byref(grid), float(T), _ptr(sums), _ptr(ws), nbytes,
                         _stream(stream)))
    _keep(stream, K, sums, ws)
    return sums


def means_x(K, x, T, stream=None):
    """Mean sums on explicit points x (trapezoid rule)."""
    K = K.contiguous(); x = x.contiguous()
    sums = torch.empty(4, dtype=torch.float64, device=K.device)
    nbytes = lib().hk_means_workspace_bytes()
    ws = torch.empty(nbytes, dtype=torch.uint8, device=K.device)
    check(lib().hk_means_x(_ptr(K), _ptr(x), K.numel(), float(T), _ptr(sums), _ptr(ws), nbytes,
                           _stream(stream)))
    _keep(stream, K, x, sums, ws)
    return sums


def sort_bins(K, seg, stream=None):
    """K with each bin [seg[b], seg[b+1]) sorted ascending."""
    K = K.contiguous()
    seg = np.ascontiguousarray(seg, dtype=np.int64)
    out = torch.empty_like(K)
    nseg = seg.size - 1
    nbytes = lib().hk_sort_bins_workspace_bytes(K.numel(), nseg)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=K.device)
    check(lib().hk_sort_bins(_ptr(K), K.numel(), seg.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                             nseg, _ptr(out), _ptr(ws), nbytes, _stream(stream)))
    _keep(stream, K, out, ws)
    return out


def transmission(Ks, nbins, Nxb, nTr, dTr, kmin=0.0, nxmin=None, stream=None):
    """Transmission function (nbins, nTr) of nbins sorted bins of Nxb values."""
    Ks = Ks.contiguous()
    Tr = torch.empty(nbins * nTr, dtype=torch.float64, device=Ks.device)
    nx = None
    if nxmin is not None:
        nx = torch.as_tensor(nxmin, dtype=torch.int32).to(Ks.device).contiguous()
    check(lib().hk_transmission(_ptr(Ks), int(nbins), int(Nxb), None if nx is None else _ptr(nx),
                                float(kmin), int(nTr), float(dTr), _ptr(Tr), _stream(stream)))
    _keep(stream, Ks, Tr, nx)
    return Tr.view(nbins, nTr)


def bin_segments(grid, numax, nbins):
    seg = (ctypes.c_int64 * (nbins + 2))()
    check(lib().hk_bin_segments(ctypes.byref(grid), float(numax), int(nbins), seg))
    return np.array(seg[:], dtype=np.int64)


def synth_lines(seed, species, n, kind=0):
    """Seeded synthetic line list (unsorted numpy arrays)."""
    out = {k: np.empty(n) for k in ("nu", "S", "EL", "gamma", "n")}
    P = _lib.P_d
    check(lib().hk_synth_lines(ctypes.c_uint64(seed), int(species), int(n), int(kind),
                               *[out[k].ctypes.data_as(P) for k in ("nu", "S", "EL", "gamma", "n")]))
    return out
