"""The differentiable cross section (torch.autograd.Function over the
hand-written adjoint kernel): gradients of a binned-opacity loss with
respect to T, P and every line's S and gamma, checked against the FP64
oracle's analytic gradient and against central finite differences
(h = 1e-4 relative), the latter at the full differentiation workload
(1e6 lines, 1e6 points, 100 bins)."""
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


def _leaf(v):
    import torch
    return torch.tensor(float(v), dtype=torch.float64, requires_grad=True)


def test_autograd_matches_oracle_gradient(E):
    import torch
    d = random_lines(500, seed=12, numin=995.0, numax=1015.0, EL=(0.0, 2000.0))
    sp = species_from(d)
    grid = E.make_grid(1000.0, 1010.0, 0.002)
    nbins = 10
    w = torch.from_numpy(np.random.default_rng(4).uniform(0, 1, nbins)).cuda()
    T, P, S, vy = _leaf(1100.0), _leaf(0.4), sp.S.clone().requires_grad_(), sp.vy.clone().requires_grad_()
    K = E.opacity(T, P, S, vy, sp, grid, Q=d["iso_Q"], cut_mode=0, cut=5.0)
    L = E.binned_log_mean(K, nbins, w)
    L.backward()
    # dL/dK_j = w_b / (Nxb mean_b), fed to the oracle's gradient
    Kn = K.detach().cpu().numpy().reshape(nbins, -1)
    gK = (w.cpu().numpy()[:, None] / (Kn.shape[1] * Kn.mean(axis=1))[:, None]
          * np.ones_like(Kn)).ravel()
    oT, oP, oS, ov = O.voigt_grad(d, 1100.0, 0.4, 0, 5.0, grid.numin, grid.dnu, grid.nx, gK)
    print("dT %.6e vs %.6e, dP %.6e vs %.6e" % (float(T.grad), oT, float(P.grad), oP))
    assert abs(float(T.grad) - oT) <= 1e-10 * abs(oT)
    assert abs(float(P.grad) - oP) <= 1e-10 * abs(oP)
    assert np.max(np.abs(S.grad.cpu().numpy() - oS)) <= 1e-10 * np.max(np.abs(oS))
    assert np.max(np.abs(vy.grad.cpu().numpy() - ov)) <= 1e-10 * np.max(np.abs(ov))


def test_autograd_differentiation_workload_vs_central_differences(E):
    """T, P and 1000 random lines (S_i, gamma_i) of the 1e6-line workload.
    A line's central difference is formed from its own spectrum:
    L(theta +- h) - L = sum_b w_b log1p(mean_b(k_i(theta +- h) - k_i) / M_b),
    which is the full loss difference without the cancellation of
    subtracting two sums over a million lines."""
    import torch
    from heliosk import configs
    w_ = configs.WORKLOADS["c3"]
    arr = configs.species_arrays(w_, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    Tv, Pbar = w_.tp[0]
    Pv = Pbar * configs.POBAR
    grid = E.make_grid(w_.numin, w_.numax, w_.dnu)
    nbins, cut = w_.nbins, w_.cut
    w = torch.from_numpy(np.random.default_rng(w_.seed).uniform(0, 1, nbins)).cuda()
    T, P = _leaf(Tv), _leaf(Pv)
    S, vy = sp.S.clone().requires_grad_(), sp.vy.clone().requires_grad_()
    K = E.opacity(T, P, S, vy, sp, grid, Q=[1.0], cut_mode=0, cut=cut)
    L = E.binned_log_mean(K, nbins, w)
    L.backward()
    h = 1e-4

    def loss(TT, PP):
        with torch.no_grad():
            KK = E.opacity(_leaf(TT), _leaf(PP), sp.S, sp.vy, sp, grid, Q=[1.0], cut_mode=0,
                           cut=cut)
            return float(E.binned_log_mean(KK, nbins, w))

    fdT = (loss(Tv * (1 + h), Pv) - loss(Tv * (1 - h), Pv)) / (2 * h * Tv)
    fdP = (loss(Tv, Pv * (1 + h)) - loss(Tv, Pv * (1 - h))) / (2 * h * Pv)
    print("c3: dL/dT %.8e FD %.8e; dL/dP %.8e FD %.8e" % (float(T.grad), fdT, float(P.grad), fdP))
    assert abs(float(T.grad) - fdT) <= 1e-4 * abs(fdT)
    assert abs(float(P.grad) - fdP) <= 1e-4 * abs(fdP)

    # 1000 random lines whose window reaches the grid
    nu_sorted = sp.nu.cpu().numpy()
    reach = np.nonzero(nu_sorted < w_.numax + cut)[0]
    pick = np.sort(np.random.default_rng(5).choice(reach, 1000, replace=False))
    M = K.detach().view(nbins, -1).mean(dim=1)
    gS, gv = S.grad.cpu().numpy(), vy.grad.cpu().numpy()
    fdS, fdv = np.empty(pick.size), np.empty(pick.size)
    host = {k: getattr(sp, k).cpu().numpy() for k in ("nu", "S", "EL", "vy", "n")}
    for m, i in enumerate(pick):
        one = E.Species(host["nu"][i:i + 1], host["S"][i:i + 1], host["EL"][i:i + 1],
                        host["vy"][i:i + 1], host["n"][i:i + 1], iso_mass=arr["iso_mass"])

        def k_of(Sv, vv):
            s1 = one.with_fields(S=torch.tensor([Sv], dtype=torch.float64, device="cuda"),
                                 vy=torch.tensor([vv], dtype=torch.float64, device="cuda"))
            st = E.line_params(s1, Tv, Pv, Q=[1.0], cut_mode=0, cut=cut)
            return E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64)

        S0, v0 = host["S"][i], host["vy"][i]
        k0 = k_of(S0, v0)

        def dL(k1):
            return float(torch.sum(w * torch.log1p((k1 - k0).view(nbins, -1).mean(dim=1) / M)))

        fdS[m] = (dL(k_of(S0 * (1 + h), v0)) - dL(k_of(S0 * (1 - h), v0))) / (2 * h * S0)
        fdv[m] = (dL(k_of(S0, v0 * (1 + h))) - dL(k_of(S0, v0 * (1 - h)))) / (2 * h * v0)
    eS = np.abs(gS[pick] - fdS) / np.maximum(np.abs(fdS), 1e-12 * np.max(np.abs(fdS)))
    ev = np.abs(gv[pick] - fdv) / np.maximum(np.abs(fdv), 1e-12 * np.max(np.abs(fdv)))
    print("c3: 1000 lines, max rel err dL/dS %.3e, dL/dgamma %.3e" % (eS.max(), ev.max()))
    assert eS.max() <= 1e-4
    assert ev.max() <= 1e-4
