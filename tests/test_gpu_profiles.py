"""GPU parity of the line profiles other than Voigt (Lorentz, Doppler, binned
Gaussian) and of plinth removal, against the FP64 oracle."""
import numpy as np
import pytest

from helpers import random_lines, rel_err, species_from, state_numpy
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from heliosk import engine
    return engine


CASES = [  # T, P, cut_mode, cut
    (1500.0, 1e-2, 0, 3.0),
    (600.0, 3.0, 1, 40.0),
    (2500.0, 0.3, 2, 6.0),
]


@pytest.mark.parametrize("profile", [2, 3, 4])
@pytest.mark.parametrize("plinth", [False, True])
@pytest.mark.parametrize("case", CASES)
def test_profile_vs_oracle(E, profile, plinth, case):
    if plinth and profile == 4:
        pytest.skip("plinth removal is not defined for profile 4")
    T, P, cm, cut = case
    d = random_lines(3000, seed=profile + 10 * cm, numin=995.0, numax=1015.0, A=1e7)
    st = E.line_params(species_from(d), T, P, Q=d["iso_Q"], cut_mode=cm, cut=cut)
    grid = E.make_grid(1000.0, 1010.0, 0.0007)
    K = E.profile(st, grid, profile, remove_plinth=plinth, kmin=1e-30).cpu().numpy()
    Ko = O.profile_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, profile,
                          remove_plinth=plinth, kmin=1e-30)
    e = rel_err(K, Ko)
    print("profile %d plinth %d %s: %.3e" % (profile, plinth, case, e))
    assert e <= 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_voigt_plinth(E, mode):
    d = random_lines(4000, seed=21, numin=995.0, numax=1015.0)
    st = E.line_params(species_from(d), 1200.0, 0.5, Q=d["iso_Q"], cut_mode=0, cut=2.0)
    grid = E.make_grid(1000.0, 1010.0, 0.001)
    K = E.profile(st, grid, 1, voigt_mode=mode).cpu().numpy()
    Kp = E.profile(st, grid, 1, remove_plinth=True, voigt_mode=mode).cpu().numpy()
    ost = state_numpy(st)
    Kop = O.profile_points(ost, grid.numin, grid.dnu, grid.nx, 1, remove_plinth=True)
    Ko = O.profile_points(ost, grid.numin, grid.dnu, grid.nx, 1)
    sub = Ko - Kop                      # the summed plinths, FP64
    # the plinth pass is FP64 in both forms: K - Kp is the plinth sum
    assert np.max(np.abs((K - Kp) - sub)) <= 1e-12 * np.max(np.abs(Ko))
    if mode == 0:
        assert rel_err(Kp, Kop) <= 1e-12
    assert np.max(np.abs(Kp - Kop)) <= (1e-12 if mode == 0 else 2e-5) * np.max(np.abs(Ko))


def test_profile_accumulate_and_errors(E):
    import torch
    d = random_lines(500, seed=5, numin=999.0, numax=1011.0)
    st = E.line_params(species_from(d), 900.0, 1.0, Q=d["iso_Q"], cut_mode=0, cut=1.0)
    grid = E.make_grid(1000.0, 1010.0, 0.001)
    K1 = E.profile(st, grid, 3, kmin=0.5)
    K2 = E.profile(st, grid, 2, K=K1.clone(), accumulate=True)
    K3 = E.profile(st, grid, 2)
    assert torch.equal(K2, K1 + K3)
    with pytest.raises(Exception, match="profile 4"):
        E.profile(st, grid, 4, remove_plinth=True)
    with pytest.raises(Exception, match="profile 5"):
        E.profile(st, grid, 5)


def _chi(T, lo, hi):
    ab = [(0.0888, -0.160, 0.00410), (0.0, 0.0526, 0.00152), (0.0232, 0.0, 0.0)]
    return [3.0, 30.0, 120.0] + [a + b * np.exp(-e * T) for a, b, e in ab] + [lo, hi]


@pytest.mark.parametrize("profile, plinth", [(1, False), (1, True), (2, False), (3, True),
                                             (4, False)])
def test_sublorentzian_profiles_vs_oracle(E, profile, plinth):
    d = random_lines(2500, seed=30 + profile, numin=2150.0, numax=2350.0, A=1e6,
                     gamma=(0.05, 0.2))
    st = E.line_params(species_from(d), 700.0, 5.0, Q=d["iso_Q"], cut_mode=0, cut=60.0)
    grid = E.make_grid(2180.0, 2320.0, 0.01)
    chi = _chi(700.0, 2200.0, 2300.0)
    K = E.profile(st, grid, profile, remove_plinth=plinth, chi=chi).cpu().numpy()
    Ko = O.profile_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, profile,
                          remove_plinth=plinth, chi=chi)
    e = rel_err(K, Ko)
    print("sub-Lorentzian profile %d plinth %d: %.3e" % (profile, plinth, e))
    assert e <= 1e-12
    Kplain = O.profile_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, profile,
                              remove_plinth=plinth)
    x = grid.numin + grid.dnu * np.arange(grid.nx)
    out = (x <= 2200.0) | (x >= 2300.0)
    assert np.max(np.abs(Ko[~out] - Kplain[~out])) > 0    # the factor acts inside the range
