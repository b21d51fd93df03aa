"""Gradients of the k-table coefficients (hk_ktable_backward and the
KTableFit autograd function): dL/dK against central finite differences of
the oracle's sort + fit (hko_sort_segments + hko_ktable), and the chain
K(T, P, S_i, gamma_i) -> k-coefficients against the oracle's analytic
Voigt gradient and against central differences of the GPU forward."""
import numpy as np
import pytest

from helpers import random_lines, species_from
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from heliosk import engine
    return engine


def _spectrum(rng, nx, kmin, frac_low):
    """Distinct values: a shuffled geometric sequence with relative gaps
    gap = exp(15 / (nx - 1)) - 1, so a relative step below gap / 2 never
    reorders a bin (the finite differences then follow the fixed order the
    gradient assumes); with kmin > 0 a fraction of the points lies well below
    kmin (cut points, Nxmin > 0).  Returns (K, the step to use)."""
    K = np.exp(np.linspace(-40.0, -25.0, nx))[rng.permutation(nx)]
    gap = np.expm1(15.0 / (nx - 1))
    if kmin > 0:
        low = rng.random(nx) < frac_low
        nl = int(low.sum())
        K[low] = kmin * np.exp(np.linspace(-6.0, -1.0, nl))
        gap = min(gap, np.expm1(5.0 / max(nl - 1, 1)))
    return K, min(1e-4, 0.1 * gap)


@pytest.mark.parametrize("Nxb,nbins,nC,kmin,frac", [(400, 4, 6, 0.0, 0.0),
                                                    (100, 3, 8, 0.0, 0.0),
                                                    (400, 2, 6, 1e-20, 0.3),
                                                    (160, 2, 7, 1e-20, 0.45)])
def test_ktable_backward_vs_oracle_fd(E, Nxb, nbins, nC, kmin, frac):
    """L = sum w_bc c_bc; dL/dK_j on the GPU against central differences
    (relative step 1e-4, or a tenth of the smallest gap between values) of
    the oracle's sort + fit on the CPU, every point."""
    import torch
    rng = np.random.default_rng(Nxb + nC)
    nx = Nxb * nbins
    K, h = _spectrum(rng, nx, kmin, frac)
    wgt = rng.uniform(-1.0, 1.0, (nbins, nC))
    grid = E.make_grid(0.0, float(nx), 1.0)
    kt = E.KTable(grid, float(nx), nbins, nC, backward=True)
    Kd = torch.from_numpy(K).cuda()
    coeff, nxmin, Ks = kt(Kd, kmin=kmin, want_sorted=True)
    gK = kt.backward(Kd, Ks, nxmin, torch.from_numpy(wgt[None]).cuda()).cpu().numpy()
    seg = O.bin_segments(0.0, float(nx), 1.0, nx, nbins)[:nbins + 1]

    def loss(Kv):
        co, no = O.ktable(O.sort_segments(Kv, seg), Nxb, nbins, nC, kmin)
        return float(np.sum(wgt * co)), no

    L0, n0 = loss(K)
    np.testing.assert_array_equal(nxmin[0].cpu().numpy(), n0)
    fd = np.empty(nx)
    for j in range(nx):
        Kp = K.copy(); Kp[j] *= 1 + h
        Km = K.copy(); Km[j] *= 1 - h
        (Lp, npl), (Lm, nmi) = loss(Kp), loss(Km)
        assert np.array_equal(npl, n0) and np.array_equal(nmi, n0)
        fd[j] = (Lp - Lm) / (2 * h * K[j])
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(gK - fd)) / scale
    print("Nxb %d nC %d kmin %g: dL/dK vs oracle FD %.2e (max |g| %.3e)" % (Nxb, nC, kmin, err, scale))
    assert err <= 1e-5
    if kmin > 0:
        assert np.all(n0 > 0)


@pytest.mark.parametrize("Nxb,nbins", [(30000, 2), (12000, 3)])
def test_ktable_backward_both_sort_paths_vs_fd(E, Nxb, nbins):
    """Bins above and below one LDS sort (16384 points): dL/dK at 300
    random points against oracle central differences of the point's bin."""
    import torch
    rng = np.random.default_rng(Nxb)
    nx, nC = Nxb * nbins, 20
    K, h = _spectrum(rng, nx, 0.0, 0.0)
    wgt = rng.uniform(-1.0, 1.0, (nbins, nC))
    grid = E.make_grid(0.0, float(nx), 1.0)
    kt = E.KTable(grid, float(nx), nbins, nC, backward=True)
    Kd = torch.from_numpy(K).cuda()
    coeff, nxmin, Ks = kt(Kd, want_sorted=True)
    gK = kt.backward(Kd, Ks, nxmin, torch.from_numpy(wgt[None]).cuda()).cpu().numpy()
    seg = O.bin_segments(0.0, float(nx), 1.0, nx, nbins)[:nbins + 1]
    Ksort = O.sort_segments(K, seg)
    # only the bin of point j changes: evaluate that bin alone
    pick = np.sort(rng.choice(nx, 300, replace=False))
    fd = np.empty(pick.size)
    for m, j in enumerate(pick):
        b = j // Nxb
        vals = []
        for sgn in (1.0, -1.0):
            Kb = K[b * Nxb:(b + 1) * Nxb].copy()
            Kb[j - b * Nxb] *= 1 + sgn * h
            co, _ = O.ktable(np.sort(Kb), Nxb, 1, nC, 0.0)
            vals.append(float(np.sum(wgt[b] * co[0])))
        fd[m] = (vals[0] - vals[1]) / (2 * h * K[j])
    err = np.max(np.abs(gK[pick] - fd)) / np.max(np.abs(fd))
    print("Nxb %d: dL/dK vs oracle FD at 300 points %.2e (step %.1e)" % (Nxb, err, h))
    # the step is held below the 2.5e-4 gaps of 60000 values, so rounding of
    # the bin's loss (~1e-14) over 2 h K dL/dK (~1e-9) limits the differences
    # to ~1e-5: the bound is the gradient tolerance, 1e-4
    assert err <= 1e-4
    assert np.array_equal(Ks.cpu().numpy(), Ksort)


def _fixed_order_loss(E, K0, nbins, nC, wgt):
    """The k-table loss L = sum w c with each bin's order frozen at K0's (the
    function the gradient differentiates: the sort order is piecewise constant
    in K) and Nxmin = 0: c_b = G ln K_b[perm_b], G = P1 V^T + P2 [I 0] the
    fit's linear map from the least-squares plan.  Returns L(K) for spectra K
    shaped like K0 (torch, on the device)."""
    import torch
    Nxb = K0.numel() // nbins
    Kv = K0.view(nbins, Nxb)
    perm = torch.argsort(Kv, dim=1, stable=True)
    Pm = E.ktable_plan(Nxb, nC)
    x = np.arange(Nxb) * 2.0 / (Nxb - 1.0) - 1.0
    G = Pm[:, :nC] @ np.polynomial.chebyshev.chebvander(x, nC - 1).T
    G[:, :nC] += Pm[:, nC:]
    wG = torch.from_numpy(np.einsum("bc,ck->bk", wgt.view(nbins, nC).cpu().numpy(), G)).cuda()

    def L(K):
        return float(torch.sum(wG * torch.log(torch.gather(K.view(nbins, Nxb), 1, perm))))
    return L, perm, wG



def test_kcoeff_gradients_chain_vs_oracle_and_fd(E):
    """L = sum w c(K(T, P, S, gamma)) through VoigtOpacity and KTableFit:
    T, P, S_i and gamma_i gradients equal the oracle's analytic Voigt
    gradient fed with the GPU's dL/dK (1e-10), and T, P match central
    differences (h = 1e-4 relative, 1e-4) of the GPU forward with each bin's
    order held at the unperturbed spectrum's (the derivative's definition;
    that loss equals the library's at the unperturbed point to 1e-10)."""
    import torch
    d = random_lines(2000, seed=21, numin=990.0, numax=1030.0, EL=(0.0, 3000.0))
    sp = species_from(d)
    grid = E.make_grid(1000.0, 1020.0, 0.001)          # 20000 points, 4 bins of 5000
    nbins, nC = 4, 20
    kt = E.KTable(grid, 1020.0, nbins, nC, backward=True)
    wgt = torch.from_numpy(np.random.default_rng(6).uniform(-1, 1, (1, nbins, nC))).cuda()

    def leaf(v):
        return torch.tensor(float(v), dtype=torch.float64, requires_grad=True)

    T, P = leaf(1300.0), leaf(0.7)
    S, vy = sp.S.clone().requires_grad_(), sp.vy.clone().requires_grad_()
    K = E.opacity(T, P, S, vy, sp, grid, Q=d["iso_Q"], cut_mode=0, cut=5.0)
    coeff, nxmin = E.ktable_fit(K, kt)
    L = torch.sum(wgt * coeff)
    L.backward()
    # the same dL/dK into the oracle's analytic gradient
    _, _, Ks = kt(K.detach(), want_sorted=True)
    gK = kt.backward(K.detach(), Ks, nxmin, wgt).cpu().numpy()
    oT, oP, oS, ov = O.voigt_grad(d, 1300.0, 0.7, 0, 5.0, grid.numin, grid.dnu, grid.nx, gK)
    assert abs(float(T.grad) - oT) <= 1e-10 * abs(oT)
    assert abs(float(P.grad) - oP) <= 1e-10 * abs(oP)
    assert np.max(np.abs(S.grad.cpu().numpy() - oS)) <= 1e-10 * np.max(np.abs(oS))
    assert np.max(np.abs(vy.grad.cpu().numpy() - ov)) <= 1e-10 * np.max(np.abs(ov))
    assert int(nxmin.max()) == 0
    Lfix, _, _ = _fixed_order_loss(E, K.detach(), nbins, nC, wgt)
    assert abs(Lfix(K.detach()) - L.item()) <= 1e-10 * abs(L.item())
    h = 1e-4

    def spectrum(TT, PP):
        with torch.no_grad():
            return E.opacity(leaf(TT), leaf(PP), sp.S, sp.vy, sp, grid, Q=d["iso_Q"],
                             cut_mode=0, cut=5.0)

    def fd(loss):
        return ((loss(spectrum(1300.0 * (1 + h), 0.7)) - loss(spectrum(1300.0 * (1 - h), 0.7)))
                / (2 * h * 1300.0),
                (loss(spectrum(1300.0, 0.7 * (1 + h))) - loss(spectrum(1300.0, 0.7 * (1 - h))))
                / (2 * h * 0.7))

    fdT, fdP = fd(Lfix)
    print("k-coefficients: dL/dT %.8e FD %.8e; dL/dP %.8e FD %.8e"
          % (float(T.grad), fdT, float(P.grad), fdP))
    assert abs(float(T.grad) - fdT) <= 1e-4 * abs(fdT)
    assert abs(float(P.grad) - fdP) <= 1e-4 * abs(fdP)
    # re-sorting at every step: values that cross between the +-h spectra
    # (among the smallest of a bin, whose first nC values the fit also takes
    # directly) shift the difference slightly from the fixed-order derivative
    rT, rP = fd(lambda KK: float(torch.sum(wgt * kt(KK)[0])))
    print("re-sorted differences: dL/dT %.8e, dL/dP %.8e" % (rT, rP))
    assert abs(float(T.grad) - rT) <= 1e-3 * abs(rT)
    assert abs(float(P.grad) - rP) <= 1e-3 * abs(rP)


def test_kcoeff_gradients_differentiation_workload(E):
    """The differentiation workload (1e6 lines, 1e6 points, 100 bins) with
    the loss sum_bc w_bc c_bc of its 20-coefficient k-table, each bin's order
    held at the unperturbed spectrum's: T and P against central differences
    (h = 1e-4 relative), and 1000 random lines (S_i, gamma_i).  A line's
    difference is formed from its own spectrum dk: with the order fixed,
    c(K + dk) - c(K) = G log1p(dK2 / K2), dK2 = dk in sorted order, G the
    fit's linear map (the least-squares plan), so no two large sums are
    subtracted."""
    import torch
    from heliosk import configs
    w_ = configs.WORKLOADS["c3"]
    arr = configs.species_arrays(w_, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    Tv, Pbar = w_.tp[0]
    Pv = Pbar * configs.POBAR
    grid = E.make_grid(w_.numin, w_.numax, w_.dnu)
    nbins, nC, cut = w_.nbins, w_.nC, w_.cut
    Nxb = grid.nx // nbins
    kt = E.KTable(grid, w_.numax, nbins, nC, backward=True)
    wgt = torch.from_numpy(np.random.default_rng(w_.seed + 1).uniform(-1, 1, (1, nbins, nC))).cuda()

    def leaf(v):
        return torch.tensor(float(v), dtype=torch.float64, requires_grad=True)

    T, P = leaf(Tv), leaf(Pv)
    S, vy = sp.S.clone().requires_grad_(), sp.vy.clone().requires_grad_()
    K = E.opacity(T, P, S, vy, sp, grid, Q=[1.0], cut_mode=0, cut=cut)
    coeff, nxmin = E.ktable_fit(K, kt)
    L = torch.sum(wgt * coeff)
    L.backward()
    h = 1e-4
    assert int(nxmin.max()) == 0                  # kmin = 0: K2 = sorted K
    Lfix, order, wG = _fixed_order_loss(E, K.detach(), nbins, nC, wgt)
    assert abs(Lfix(K.detach()) - L.item()) <= 1e-10 * abs(L.item())

    def spectrum(TT, PP):
        with torch.no_grad():
            return E.opacity(leaf(TT), leaf(PP), sp.S, sp.vy, sp, grid, Q=[1.0], cut_mode=0,
                             cut=cut)

    fdT = (Lfix(spectrum(Tv * (1 + h), Pv)) - Lfix(spectrum(Tv * (1 - h), Pv))) / (2 * h * Tv)
    fdP = (Lfix(spectrum(Tv, Pv * (1 + h))) - Lfix(spectrum(Tv, Pv * (1 - h)))) / (2 * h * Pv)
    print("c3 k-table loss: dL/dT %.8e FD %.8e; dL/dP %.8e FD %.8e"
          % (float(T.grad), fdT, float(P.grad), fdP))
    assert abs(float(T.grad) - fdT) <= 1e-4 * abs(fdT)
    assert abs(float(P.grad) - fdP) <= 1e-4 * abs(fdP)

    # per-line differences through the fixed-order pipeline
    K2 = torch.gather(K.detach().view(nbins, Nxb), 1, order)
    nu_sorted = sp.nu.cpu().numpy()
    reach = np.nonzero(nu_sorted < w_.numax + cut)[0]
    pick = np.sort(np.random.default_rng(5).choice(reach, 1000, replace=False))
    gS, gv = S.grad.cpu().numpy(), vy.grad.cpu().numpy()
    host = {k: getattr(sp, k).cpu().numpy() for k in ("nu", "S", "EL", "vy", "n")}
    fdS, fdv = np.empty(pick.size), np.empty(pick.size)
    for m, i in enumerate(pick):
        one = E.Species(host["nu"][i:i + 1], host["S"][i:i + 1], host["EL"][i:i + 1],
                        host["vy"][i:i + 1], host["n"][i:i + 1], iso_mass=arr["iso_mass"])

        def k_of(Sv, vv):
            s1 = one.with_fields(S=torch.tensor([Sv], dtype=torch.float64, device="cuda"),
                                 vy=torch.tensor([vv], dtype=torch.float64, device="cuda"))
            st = E.line_params(s1, Tv, Pv, Q=[1.0], cut_mode=0, cut=cut)
            return E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).view(nbins, Nxb)

        S0, v0 = host["S"][i], host["vy"][i]
        k0 = k_of(S0, v0)

        def dL(k1):
            dk = torch.gather(k1 - k0, 1, order)
            return float(torch.sum(wG * torch.log1p(dk / K2)))

        fdS[m] = (dL(k_of(S0 * (1 + h), v0)) - dL(k_of(S0 * (1 - h), v0))) / (2 * h * S0)
        fdv[m] = (dL(k_of(S0, v0 * (1 + h))) - dL(k_of(S0, v0 * (1 - h)))) / (2 * h * v0)
    eS = np.abs(gS[pick] - fdS) / np.maximum(np.abs(fdS), 1e-12 * np.max(np.abs(fdS)))
    ev = np.abs(gv[pick] - fdv) / np.maximum(np.abs(fdv), 1e-12 * np.max(np.abs(fdv)))
    print("c3 k-table loss: 1000 lines, max rel err dL/dS %.3e, dL/dgamma %.3e"
          % (eS.max(), ev.max()))
    assert eS.max() <= 1e-4
    assert ev.max() <= 1e-4
