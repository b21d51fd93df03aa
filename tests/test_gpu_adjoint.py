"""GPU adjoint kernel against the oracle's analytic gradient and central
finite differences of the FP64 forward sum."""
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


@pytest.mark.parametrize("T,P,cm,cut,delta", [(1200.0, 0.1, 0, 5.0, 0.0),
                                              (500.0, 3.0, 0, 2.0, -0.002),
                                              (2500.0, 0.01, 2, 30.0, 0.0)])
def test_adjoint_vs_oracle(E, T, P, cm, cut, delta):
    import torch
    d = random_lines(400, seed=3, numin=995.0, numax=1015.0, delta=delta, EL=(0.0, 2000.0))
    grid = E.make_grid(1000.0, 1010.0, 0.002)
    gK = np.random.default_rng(1).uniform(0.0, 1.0, grid.nx)
    sp = species_from(d)
    gTP, gS, gv = E.adjoint(sp, T, P, torch.from_numpy(gK).cuda(), grid, Q=d["iso_Q"],
                            cut_mode=cm, cut=cut)
    oT, oP, oS, ov = O.voigt_grad(d, T, P, cm, cut, grid.numin, grid.dnu, grid.nx, gK)
    gTP = gTP.cpu().numpy()
    assert abs(gTP[0] - oT) <= 1e-10 * abs(oT)
    assert abs(gTP[1] - oP) <= 1e-10 * abs(oP)
    s = np.max(np.abs(oS)); v = np.max(np.abs(ov))
    assert np.max(np.abs(gS.cpu().numpy() - oS)) <= 1e-10 * s
    assert np.max(np.abs(gv.cpu().numpy() - ov)) <= 1e-10 * v


@pytest.mark.parametrize("form", [1, 2])
def test_adjoint_exomol_forms_vs_oracle(E, form):
    import torch
    d = random_lines(400, seed=9, numin=995.0, numax=1015.0, A=5.0e7, EL=(0.0, 2000.0))
    grid = E.make_grid(1000.0, 1010.0, 0.002)
    gK = np.random.default_rng(3).uniform(0.0, 1.0, grid.nx)
    sp = species_from(d, form=form)
    T, P = 900.0, 0.4
    gTP, gS, gv = E.adjoint(sp, T, P, torch.from_numpy(gK).cuda(), grid, Q=d["iso_Q"],
                            cut_mode=0, cut=5.0)
    oT, oP, oS, ov = O.voigt_grad(d, T, P, 0, 5.0, grid.numin, grid.dnu, grid.nx, gK, form=form)
    gTP = gTP.cpu().numpy()
    assert abs(gTP[0] - oT) <= 1e-10 * abs(oT)
    assert abs(gTP[1] - oP) <= 1e-10 * abs(oP)
    assert np.max(np.abs(gS.cpu().numpy() - oS)) <= 1e-10 * np.max(np.abs(oS))
    assert np.max(np.abs(gv.cpu().numpy() - ov)) <= 1e-10 * np.max(np.abs(ov))


def test_adjoint_vs_central_differences(E):
    import torch
    d = random_lines(60, seed=8, numin=998.0, numax=1012.0, EL=(0.0, 500.0))
    grid = E.make_grid(1000.0, 1010.0, 0.002)
    gKn = np.random.default_rng(2).uniform(0.0, 1.0, grid.nx)
    gK = torch.from_numpy(gKn).cuda()
    T, P = 900.0, 0.5

    def loss(dd, TT, PP):
        st = E.line_params(species_from(dd), TT, PP, Q=dd["iso_Q"], cut_mode=0, cut=3.0)
        return float(torch.dot(gK, E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64)))

    gTP, gS, gv = E.adjoint(species_from(d), T, P, gK, grid, Q=d["iso_Q"], cut_mode=0, cut=3.0)
    gTP = gTP.cpu().numpy(); gS = gS.cpu().numpy(); gv = gv.cpu().numpy()
    h = 1e-4
    fdT = (loss(d, T * (1 + h), P) - loss(d, T * (1 - h), P)) / (2 * h * T)
    fdP = (loss(d, T, P * (1 + h)) - loss(d, T, P * (1 - h))) / (2 * h * P)
    assert abs(gTP[0] - fdT) <= 1e-4 * abs(fdT)
    assert abs(gTP[1] - fdP) <= 1e-4 * abs(fdP)
    # per line: L is linear in K, so the line's own spectrum gives the
    # difference of L without cancelling against the other lines
    for i in (0, 17, 59):
        one = {k: (v[i:i + 1].copy() if isinstance(v, np.ndarray) and v.shape == d["nu"].shape
                   else v) for k, v in d.items()}
        for key, g in (("S", gS), ("vy", gv)):
            dp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in one.items()}
            dm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in one.items()}
            dp[key][0] *= 1 + h; dm[key][0] *= 1 - h
            fd = (loss(dp, T, P) - loss(dm, T, P)) / (2 * h * d[key][i])
            assert abs(g[i] - fd) <= 1e-4 * abs(fd), (key, i, g[i], fd)
