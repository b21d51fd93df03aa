"""GPU parity of the line-parameter and Voigt kernels against the FP64 oracle.

Run on an MI355X: python -m pytest tests -m gpu.  The oracle is the checker
only (tests import it; the engine never does).
"""
import numpy as np
import pytest

from helpers import random_lines, rel_err, species_from, state_numpy
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11      # plain FP64 form vs the oracle
FP32C_TOL = 1e-5      # FP32 compensated form vs the oracle (the required bound)


@pytest.fixture(scope="module")
def E():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from heliosk import engine
    return engine


LP_CASES = [
    # T, P[atm], cut_mode, cut, delta, A, niso
    (296.0, 1.0, 0, 25.0, 0.0, 0.0, 1),
    (1500.0, 0.3, 0, 100.0, -0.003, 0.0, 1),
    (100.0, 1e-6, 1, 100.0, 0.0, 5e6, 3),
    (3000.0, 100.0, 2, 50.0, 0.01, 0.0, 2),
    (800.0, 2.0, 0, 0.0, 0.0, 0.0, 1),
]


@pytest.mark.parametrize("case", LP_CASES)
def test_line_params_vs_oracle(E, case):
    T, P, cm, cut, delta, A, niso = case
    d = random_lines(5000, seed=11, delta=delta, A=A, niso=niso, EL=(0.0, 20000.0))
    sp = species_from(d)
    st = E.line_params(sp, T, P, Q=d["iso_Q"], cut_mode=cm, cut=cut)
    g = state_numpy(st)
    o = O.line_params(d, T, P, cm, 1e30 if cut == 0.0 else cut)
    for k in ("nu", "S", "ialphaD", "y"):
        np.testing.assert_allclose(g[k], o[k], rtol=2e-13, atol=0, err_msg=k)
    np.testing.assert_allclose(g["hw"], o["hw"], rtol=2e-13, atol=0)


@pytest.mark.parametrize("form", [1, 2])
def test_line_params_exomol_forms_vs_oracle(E, form):
    """Sf_kernel / SfSuper_kernel semantics (ExoMol family, super-lines),
    then the plain FP64 sum of the resulting state."""
    d = random_lines(5000, seed=13, A=2.0e7, EL=(0.0, 20000.0), numin=995.0, numax=1015.0)
    sp = species_from(d, form=form)
    T, P = 1700.0, 0.3
    st = E.line_params(sp, T, P, Q=d["iso_Q"], cut_mode=0, cut=5.0)
    g = state_numpy(st)
    o = O.line_params(d, T, P, 0, 5.0, form=form)
    for k in ("nu", "S", "ialphaD", "y", "hw"):
        np.testing.assert_allclose(g[k], o[k], rtol=2e-13, atol=0, err_msg=k)
    grid = E.make_grid(1000.0, 1010.0, 0.001)
    K = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    Ko = O.voigt_points(o, grid.numin, grid.dnu, grid.nx)
    assert rel_err(K, Ko) <= 1e-12


def _grid(E, numin, numax, dnu):
    return E.make_grid(numin, numax, dnu)


def _compare(E, d, T, P, grid, cm, cut, kmin=0.0, tile_points=0, Q=None):
    sp = species_from(d)
    st = E.line_params(sp, T, P, Q=Q if Q is not None else d["iso_Q"], cut_mode=cm, cut=cut)
    ost = state_numpy(st)
    Ko = O.voigt_points(ost, grid.numin, grid.dnu, grid.nx, kmin=kmin)
    K64 = E.voigt(st, grid, kmin, E.HK_VOIGT_FP64, tile_points).cpu().numpy()
    K32 = E.voigt(st, grid, kmin, E.HK_VOIGT_FP32C, tile_points).cpu().numpy()
    return Ko, K64, K32


VOIGT_CASES = [
    # name, T, P, cut_mode, cut, nlines, (numin, numax, dnu), line range
    ("core-dominated", 1500.0, 1e-3, 0, 2.0, 200, (1000.0, 1010.0, 0.0005), (999.0, 1011.0)),
    ("pressure-wings", 1000.0, 10.0, 0, 5.0, 300, (1000.0, 1010.0, 0.001), (995.0, 1015.0)),
    ("lorentz-cut", 296.0, 1.0, 1, 30.0, 300, (1000.0, 1010.0, 0.001), (998.0, 1012.0)),
    ("doppler-cut", 3000.0, 0.01, 2, 40.0, 300, (1000.0, 1010.0, 0.001), (998.0, 1012.0)),
    ("no-cut", 700.0, 0.5, 0, 0.0, 60, (1000.0, 1004.0, 0.002), (990.0, 1014.0)),
    ("low-T-low-P", 100.0, 1e-6, 0, 1.0, 400, (1000.0, 1002.0, 0.0001), (999.5, 1002.5)),
    # many line cores overlap every point and every 256-line LDS stage
    ("dense-cores", 2000.0, 9.87, 0, 5.0, 20000, (10000.0, 10010.0, 0.001), (9995.0, 10015.0)),
    # each line core spans thousands of points (many line-core tiles)
    ("wide-cores", 3000.0, 1e-4, 0, 1.0, 100, (1000.0, 1004.0, 0.00001), (999.5, 1004.5)),
    # Doppler widths far below the grid step (x steps of ~100 per point):
    # whole-wave Gauss-Hermite paths would span x^2 up to 1e10
    ("narrow-lines", 100.0, 1e-6, 0, 100.0, 200, (0.0, 1000.0, 0.01), (0.0, 1000.0)),
]


@pytest.mark.parametrize("case", VOIGT_CASES, ids=[c[0] for c in VOIGT_CASES])
def test_voigt_vs_oracle(E, case):
    name, T, P, cm, cut, n, (a, b, dnu), (la, lb) = case
    d = random_lines(n, seed=hash(name) % 1000, numin=la, numax=lb, EL=(0.0, 3000.0))
    grid = _grid(E, a, b, dnu)
    Ko, K64, K32 = _compare(E, d, T, P, grid, cm, cut, kmin=0.0)
    e64, e32 = rel_err(K64, Ko), rel_err(K32, Ko)
    print("%s: FP64 %.3e  FP32C %.3e" % (name, e64, e32))
    assert e64 <= FP64_TOL
    assert e32 <= FP32C_TOL
    # exactly-empty points agree (shared FP64 integer windows)
    np.testing.assert_array_equal(K64 == 0.0, Ko == 0.0)
    np.testing.assert_array_equal(K32 == 0.0, Ko == 0.0)


def test_voigt_centres_near_zero(E):
    """Centres down to 1e-4 cm^-1: alpha_D -> 0, so y reaches ~1e8 and x
    ~1e9; every FP32 intermediate must stay finite."""
    rng = np.random.default_rng(31)
    d = random_lines(300, seed=31, numin=0.0, numax=12.0)
    d["nu"] = np.sort(10 ** rng.uniform(-4, np.log10(12.0), 300))
    grid = _grid(E, 0.0, 10.0, 0.001)
    Ko, K64, K32 = _compare(E, d, 1500.0, 1.0, grid, 0, 5.0)
    e64, e32 = rel_err(K64, Ko), rel_err(K32, Ko)
    print("near-zero centres: FP64 %.3e  FP32C %.3e" % (e64, e32))
    assert np.all(np.isfinite(K32)) and np.all(np.isfinite(K64))
    assert e64 <= FP64_TOL and e32 <= FP32C_TOL


def test_voigt_edges_and_kmin(E):
    # lines outside the grid, at both edges, a zero-width (P=0, A=0) Doppler
    # line, kmin > 0 and accumulate
    d = random_lines(50, seed=5, numin=999.0, numax=1001.0)
    d["nu"] = np.sort(np.r_[d["nu"][:-4], 990.0, 1000.0, 1000.99995, 1020.0])
    grid = _grid(E, 1000.0, 1001.0, 0.0001)
    Ko, K64, K32 = _compare(E, d, 500.0, 0.0, grid, 0, 0.5, kmin=1e-30)
    assert rel_err(K64, Ko) <= FP64_TOL and rel_err(K32, Ko) <= FP32C_TOL
    sp = species_from(d)
    st = E.line_params(sp, 500.0, 0.0, Q=d["iso_Q"], cut_mode=0, cut=0.5)
    K = E.voigt(st, grid, 1e-30, E.HK_VOIGT_FP64)
    K2 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, K=K.clone() * 0 + 5.0, accumulate=True)
    np.testing.assert_allclose(K2.cpu().numpy(), 5.0 + (K.cpu().numpy() - 1e-30), rtol=1e-15)


def test_voigt_empty_inputs(E):
    import torch
    d = random_lines(0, seed=1)
    grid = _grid(E, 1000.0, 1001.0, 0.001)
    sp = species_from(d)
    st = E.line_params(sp, 500.0, 1.0, Q=d["iso_Q"], cut_mode=0, cut=1.0)
    for mode in (E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C):
        K = E.voigt(st, grid, 2.5e-7, mode)
        assert torch.all(K == 2.5e-7)


def test_voigt_tiles_and_determinism(E):
    d = random_lines(500, seed=9, numin=995.0, numax=1015.0)
    grid = _grid(E, 1000.0, 1010.0, 0.001)
    sp = species_from(d)
    st = E.line_params(sp, 900.0, 2.0, Q=d["iso_Q"], cut_mode=0, cut=5.0)
    ref = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, 1024).cpu().numpy()
    np.testing.assert_array_equal(E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, 1024).cpu().numpy(), ref)
    # other tiles: the same lines in the same order, but a wave may pick a
    # different (equally valid) region formula for a point
    for tp in (512, 2048, 4096):
        assert rel_err(E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, tp).cpu().numpy(), ref) <= 1e-13
    r32 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, 2048).cpu().numpy()
    np.testing.assert_array_equal(E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, 2048).cpu().numpy(), r32)
    for tp in (512, 1024, 4096):
        assert rel_err(E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, tp).cpu().numpy(), ref) <= FP32C_TOL


def test_voigt_batched_matches_single(E):
    import torch
    items, refs = [], []
    grid = _grid(E, 1000.0, 1010.0, 0.001)
    for s, n in enumerate((40, 400, 1200)):
        d = random_lines(n, seed=20 + s, numin=990.0, numax=1020.0)
        sp = species_from(d)
        for T, P in ((300.0, 0.1), (2000.0, 5.0)):
            st = E.line_params(sp, T, P, Q=d["iso_Q"], cut_mode=0, cut=8.0)
            for mode in (E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C):
                refs.append((mode, E.voigt(st, grid, 0.0, mode).cpu().numpy()))
            items.append(st)
    for mi, mode in enumerate((E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C)):
        outs = [torch.empty(grid.nx, dtype=torch.float64, device="cuda") for _ in items]
        E.voigt_batched([(st, grid, 0.0, K) for st, K in zip(items, outs)], mode=mode)
        for k, K in enumerate(outs):
            np.testing.assert_array_equal(K.cpu().numpy(), refs[2 * k + mi][1])


def test_voigt_batch_graph_replay(E):
    """The prepared batch captured in a graph: every replay equals the eager
    launch bit for bit, with eager work between replays, and a replay after
    the line states change recomputes from the new states."""
    import torch
    grid = _grid(E, 1000.0, 1010.0, 0.001)
    sps, sts = [], []
    for s, n in enumerate((300, 1500)):
        d = random_lines(n, seed=40 + s, numin=990.0, numax=1020.0)
        sp = species_from(d)
        sps.append((sp, d))
        sts.append(E.line_params(sp, 1200.0, 2.0, Q=d["iso_Q"], cut_mode=0, cut=8.0))
    for mode in (E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C):
        outs = [torch.empty(grid.nx, dtype=torch.float64, device="cuda") for _ in sts]
        refs = [E.voigt(st, grid, 1e-30, mode).cpu().numpy() for st in sts]
        batch = E.VoigtBatch([(st, grid, 1e-30, K) for st, K in zip(sts, outs)], mode=mode)
        batch.launch()
        torch.cuda.synchronize()
        for K, r in zip(outs, refs):
            np.testing.assert_array_equal(K.cpu().numpy(), r)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            batch.launch()            # warm-up on the capture stream's device
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            batch.launch()
        for rep in range(3):
            for K in outs:
                K.fill_(-1.0)
            other = E.voigt(sts[0], grid, 0.0, mode)          # eager work between replays
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for K, r in zip(outs, refs):
                np.testing.assert_array_equal(K.cpu().numpy(), r, err_msg="replay %d" % rep)
            del other
        # new (T, P) in the same line-state memory, then replay
        for (sp, d), st in zip(sps, sts):
            E.line_params(sp, 400.0, 0.5, Q=d["iso_Q"], cut_mode=0, cut=8.0, state=st)
        refs2 = [E.voigt(st, grid, 1e-30, mode).cpu().numpy() for st in sts]
        g.replay()
        torch.cuda.synchronize()
        for K, r in zip(outs, refs2):
            np.testing.assert_array_equal(K.cpu().numpy(), r)
        for (sp, d), st in zip(sps, sts):
            E.line_params(sp, 1200.0, 2.0, Q=d["iso_Q"], cut_mode=0, cut=8.0, state=st)
        del g


def test_config0_fp64_and_fp32c_subsample(E):
    """Workload c0 (1e5 lines, 3e6 points): both forms vs the oracle on 20000
    random grid points, and FP32C vs FP64 on the whole grid."""
    from heliosk import configs
    w = configs.WORKLOADS["c0"]
    arr = configs.species_arrays(w, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    T, Pbar = w.tp[0]
    st = E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
    grid = _grid(E, w.numin, w.numax, w.dnu)
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    K32 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C).cpu().numpy()
    jj = np.sort(np.random.default_rng(0).choice(grid.nx, 20000, replace=False))
    Ko = O.voigt_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, jidx=jj)
    e64, e32, e3264 = rel_err(K64[jj], Ko), rel_err(K32[jj], Ko), rel_err(K32, K64)
    print("c0: FP64 vs oracle %.3e, FP32C vs oracle %.3e, FP32C vs FP64 (full grid) %.3e"
          % (e64, e32, e3264))
    assert e64 <= FP64_TOL and e32 <= FP32C_TOL and e3264 <= FP32C_TOL


def test_config1_precision_at_scale(E):
    """Workload c1 (1e7 lines, 3e7 points, 2000 K, 10 bar, cut 25): FP32C vs
    FP64 on the whole grid, both against the oracle on a seeded 1 % random
    subsample of the grid (300000 points), and the FP32C k-tables (300 bins x
    20) against the FP64 ones and the oracle's fit of the FP64 spectrum:
    identical Nxmin (ystart), coefficients within 1e-4."""
    import torch
    from heliosk import configs
    w = configs.WORKLOADS["c1"]
    arr = configs.species_arrays(w, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    T, Pbar = w.tp[0]
    st = E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
    grid = _grid(E, w.numin, w.numax, w.dnu)
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64)
    K32 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C)
    tau = 1e-12 * float(K64.max())
    e3264 = float(torch.max(torch.abs(K32 - K64) / torch.clamp(torch.abs(K64), min=tau)))
    jj = np.sort(np.random.default_rng(1).choice(grid.nx, grid.nx // 100, replace=False))
    Ko = O.voigt_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, jidx=jj)
    k64 = K64.cpu().numpy()[jj]; k32 = K32.cpu().numpy()[jj]
    e64, e32 = rel_err(k64, Ko), rel_err(k32, Ko)
    print("c1: FP64 vs oracle %.3e, FP32C vs oracle %.3e (%d points), FP32C vs FP64 (full grid) "
          "%.3e" % (e64, e32, jj.size, e3264))
    assert e64 <= FP64_TOL and e32 <= FP32C_TOL and e3264 <= FP32C_TOL
    kt = E.KTable(grid, w.numax, w.nbins, w.nC)
    c32, n32 = kt(K32)
    c64, n64, Ks64 = kt(K64, want_sorted=True)
    assert torch.equal(n32, n64)
    d = float(torch.max(torch.abs(c32 - c64)))
    co, no = O.ktable(Ks64.cpu().numpy(), grid.nx // w.nbins, w.nbins, w.nC, 0.0)
    np.testing.assert_array_equal(n64[0].cpu().numpy(), no)
    do = float(np.max(np.abs(c64[0].cpu().numpy() - co)))
    print("c1 k-tables: FP32C vs FP64 max |dc| %.3e, FP64 vs oracle %.3e" % (d, do))
    assert d <= 1e-4 and do <= 1e-9


@pytest.mark.parametrize("n", [1, 1000, 300001])
def test_device_line_sort_is_stable(E, n):
    """hk_lines_sort against numpy's stable argsort: the same permutation,
    bit for bit, with many equal centres (ties keep input order)."""
    import torch
    rng = np.random.default_rng(n)
    nu = np.round(rng.uniform(0.0, 3000.0, n), 1)        # ~30000 distinct values: many ties
    f = {k: rng.uniform(0.0, 1.0, n) for k in ("S", "EL", "A", "delta", "vy", "n")}
    iso = rng.integers(0, 4, n).astype(np.int32)
    sp = E.sort_lines(nu, f["S"], f["EL"], f["vy"], f["n"], A=f["A"], delta=f["delta"], iso=iso,
                      iso_mass=(1.0, 2.0, 3.0, 4.0), iso_ab=(1.0,) * 4, iso_sscale=(1.0,) * 4,
                      form=E.HK_FORM_EXOMOL)
    torch.cuda.synchronize()
    o = np.argsort(nu, kind="stable")
    np.testing.assert_array_equal(sp.nu.cpu().numpy(), nu[o])
    for k, v in f.items():
        np.testing.assert_array_equal(getattr(sp, k).cpu().numpy(), v[o], err_msg=k)
    np.testing.assert_array_equal(sp.iso.cpu().numpy(), iso[o])
    assert sp.form == E.HK_FORM_EXOMOL and sp.nlines == n


@pytest.mark.parametrize("cm, cut, dnu", [(0, 25.0, 0.001), (1, 400.0, 0.002), (2, 6000.0, 0.001),
                                          (0, 0.0, 0.004)])
def test_fp32c_far_field_vs_exact_pairs(E, cm, cut, dnu):
    """The FP32 form's far-field interpolation against the same form with
    every pair evaluated and against the FP64 sum: it must change the
    result (it is engaged) by far less than the 1e-5 bound."""
    d = random_lines(30000, seed=44, numin=900.0, numax=1100.0, gamma=(0.01, 0.1))
    st = E.line_params(species_from(d), 1500.0, 1.0, Q=d["iso_Q"], cut_mode=cm, cut=cut)
    grid = E.make_grid(950.0, 1050.0, dnu)
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    Kff = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C).cpu().numpy()
    Kex = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, exact_pairs=True).cpu().numpy()
    e_ff, e_ex, d_fe = rel_err(Kff, K64), rel_err(Kex, K64), rel_err(Kff, Kex)
    print("cut mode %d cut %g: far field vs FP64 %.3e, exact pairs vs FP64 %.3e, "
          "between %.3e" % (cm, cut, e_ff, e_ex, d_fe))
    assert e_ff <= 1e-5 and e_ex <= 1e-5
    assert 0.0 < d_fe <= 2e-6
    # deterministic
    assert np.array_equal(Kff, E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C).cpu().numpy())


@pytest.mark.parametrize("tp, cut, n", [(512, 6.0, 20), (1024, 12.0, 10), (4096, 49.0, 4)])
def test_far_field_window_ends_inside_tile(E, tp, cut, n):
    """Sparse lines whose cut windows (three tile widths each side) end inside
    tiles: the waves a far line covers take it from the far field, the wave
    with its window edge evaluates it pair by pair.  The points outside every
    window stay exactly zero, the sum matches FP64 and the every-pair form,
    the work count still adds up, and the batched form gives the same bits."""
    import torch
    d = random_lines(n, seed=48, numin=900.0, numax=1500.0, gamma=(0.01, 0.1))
    st = E.line_params(species_from(d), 1500.0, 1.0, Q=d["iso_Q"], cut_mode=0, cut=cut)
    grid = E.make_grid(1000.0, 1400.0, 0.004)
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    Kff = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, tile_points=tp).cpu().numpy()
    Kex = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, tile_points=tp, exact_pairs=True).cpu().numpy()
    assert np.array_equal(Kff == 0.0, K64 == 0.0) and np.any(K64 == 0.0)
    e_ff, d_fe = rel_err(Kff, K64), rel_err(Kff, Kex)
    tot, A, B, C = E.pair_count(st, grid)
    w = E.work_count(st, grid, E.HK_VOIGT_FP32C, tile_points=tp)
    print("tile %d: far field vs FP64 %.3e, vs every pair %.3e; direct %d far %d cut-edge %d of %d"
          % (tp, e_ff, d_fe, sum(w[:3]), w[3], w[11], tot))
    assert e_ff <= 1e-5 and 0.0 < d_fe <= 2e-6
    assert w[3] > 0 and sum(w[:4]) + w[11] == tot
    K = torch.empty(grid.nx, dtype=torch.float64, device="cuda")
    E.VoigtBatch([(st, grid, 0.0, K)], mode=E.HK_VOIGT_FP32C, tile_points=tp).launch()
    torch.cuda.synchronize()
    assert np.array_equal(K.cpu().numpy(), Kff)


def test_work_count_adds_up(E):
    """Direct pairs by region plus far-field and cut-edge pairs equal the pair
    count; with every pair evaluated (or in FP64) nothing goes through the
    far field."""
    d = random_lines(20000, seed=46, numin=900.0, numax=1100.0)
    st = E.line_params(species_from(d), 1800.0, 2.0, Q=d["iso_Q"], cut_mode=0, cut=25.0)
    grid = E.make_grid(950.0, 1050.0, 0.001)
    tot, A, B, C = E.pair_count(st, grid)
    w = E.work_count(st, grid, E.HK_VOIGT_FP32C)
    assert w[3] > 0 and all(v > 0 for v in w[4:7]) and sum(w[7:10]) > 0
    assert w[10] > 0 and w[11] > 0 and sum(w[:4]) + w[11] == tot
    for kw in (dict(mode=E.HK_VOIGT_FP32C, exact_pairs=True), dict(mode=E.HK_VOIGT_FP64)):
        assert E.work_count(st, grid, **kw) == [A, B, C] + [0] * 9


@pytest.mark.parametrize("cm, cut", [(0, 25.0), (1, 400.0)])
def test_compat_vs_oracle(E, cm, cut):
    """HK_VOIGT_COMPAT against the oracle's restatement of the reference's
    FP32 arithmetic (same float operations; region C differs by the device's
    expf/sinf/cosf last bits), and against the FP64 form by the reference's
    own region-B error (up to 1.4e-2)."""
    d = random_lines(20000, seed=47, numin=900.0, numax=1100.0)
    st = E.line_params(species_from(d), 1500.0, 1.0, Q=d["iso_Q"], cut_mode=cm, cut=cut)
    grid = E.make_grid(950.0, 1050.0, 0.002)
    Kc = E.voigt(st, grid, 0.0, E.HK_VOIGT_COMPAT).cpu().numpy()
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    Ko = O.voigt_compat_points(state_numpy(st), grid.numin, grid.dnu, grid.nx)
    e, e64 = rel_err(Kc, Ko), rel_err(Kc, K64)
    print("cut mode %d: compat vs oracle %.3e (%.1f %% identical), vs FP64 %.3e"
          % (cm, e, 100.0 * np.mean(Kc == Ko), e64))
    assert e <= 2e-6
    assert 1e-4 <= e64 <= 1.5e-2
    assert np.array_equal(Kc, E.voigt(st, grid, 0.0, E.HK_VOIGT_COMPAT).cpu().numpy())


def test_compat_without_core_is_bitwise(E):
    """At 100 atm every pair has x^2 + y^2 > 100 (regions A and B only): the
    device rounds each pair exactly as the oracle's float restatement and adds
    in the same order, so the sums agree bit for bit."""
    d = random_lines(5000, seed=48, numin=900.0, numax=1100.0)
    st = E.line_params(species_from(d), 1000.0, 100.0, Q=d["iso_Q"], cut_mode=0, cut=25.0)
    y = st.array("y").cpu().numpy()
    assert y.min() > 10.0
    grid = E.make_grid(950.0, 1050.0, 0.005)
    Kc = E.voigt(st, grid, 1e-30, E.HK_VOIGT_COMPAT).cpu().numpy()
    Ko = O.voigt_compat_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, kmin=1e-30)
    print("identical points: %d of %d, max rel %.3e" % (np.sum(Kc == Ko), Kc.size,
                                                       rel_err(Kc, Ko)))
    np.testing.assert_array_equal(Kc, Ko)


def test_compat_float_index_beyond_2_24(E):
    """On a grid of more than 2^24 points the reference converts the point
    index to float, so positions beyond 16777216 round to even indices; the
    compat form does the same, matching the oracle there."""
    d = random_lines(400, seed=49, numin=16850.0, numax=16950.0)
    st = E.line_params(species_from(d), 1500.0, 1.0, Q=d["iso_Q"], cut_mode=0, cut=5.0)
    grid = E.make_grid(0.0, 17000.0, 0.001)
    assert grid.nx > 1 << 24
    Kc = E.voigt(st, grid, 0.0, E.HK_VOIGT_COMPAT).cpu().numpy()
    j = np.arange(16840000, 16960000, 7, dtype=np.int64)
    Ko = O.voigt_compat_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, jidx=j)
    assert rel_err(Kc[j], Ko) <= 2e-6
    # an odd index rounds (half to even) onto one of its even neighbours, so
    # every line sees the same position there: about half of the (even, odd)
    # pairs of points carry identical sums
    jj = np.arange(16900000, 16900200, 2)
    assert 0.4 < np.mean(Kc[jj] == Kc[jj + 1]) < 0.6


def test_compat_cut_mode_2_accumulate_and_empty_lines(E):
    """Compat form with cut mode 2 (cut in Doppler widths), lines at nu = 0
    (empty: no window), kmin and accumulation onto an existing K."""
    d = random_lines(6000, seed=50, numin=900.0, numax=1100.0)
    d["nu"][:3] = 0.0
    st = E.line_params(species_from(d), 900.0, 0.3, Q=d["iso_Q"], cut_mode=2, cut=300.0)
    grid = E.make_grid(950.0, 1050.0, 0.002)
    Ko = O.voigt_compat_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, kmin=1e-28)
    Kc = E.voigt(st, grid, 1e-28, E.HK_VOIGT_COMPAT).cpu().numpy()
    assert rel_err(Kc, Ko) <= 2e-6
    import torch
    K0 = torch.full((grid.nx,), 5e-25, dtype=torch.float64, device="cuda")
    Ka = E.voigt(st, grid, 0.0, E.HK_VOIGT_COMPAT, K=K0, accumulate=True).cpu().numpy()
    assert rel_err(Ka - 5e-25, Ko - 1e-28) <= 2e-6


@pytest.mark.parametrize("k", [0, 15])
def test_mixed_width_lines_vs_oracle(E, k):
    """The heavy-tailed generator of the mixed-width workload (widths
    log-uniform over four decades, 5 % strong lines) at its coldest, lowest
    pressure and its hottest, highest pressure points, 2e5 lines on the full
    4e6-point grid: FP64 and FP32C against the oracle on 3000 random points,
    FP32C against FP64 everywhere."""
    from heliosk import configs
    w = configs.WORKLOADS["c4"]
    arr = configs.species_arrays(w, 0, n=200_000)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    T, Pbar = w.tp[k]
    st = E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
    grid = _grid(E, w.numin, w.numax, w.dnu)
    K64 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64).cpu().numpy()
    K32 = E.voigt(st, grid, 0.0, E.HK_VOIGT_FP32C, tile_points=512).cpu().numpy()
    jj = np.sort(np.random.default_rng(2).choice(grid.nx, 3000, replace=False))
    Ko = O.voigt_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, jidx=jj)
    e64, e32, e3264 = rel_err(K64[jj], Ko), rel_err(K32[jj], Ko), rel_err(K32, K64)
    print("mixed widths (T %.0f K, P %.3g bar): FP64 vs oracle %.3e, FP32C vs oracle %.3e, "
          "FP32C vs FP64 %.3e" % (T, Pbar, e64, e32, e3264))
    assert e64 <= FP64_TOL and e32 <= FP32C_TOL and e3264 <= FP32C_TOL


def test_batched_extreme_tp_vs_oracle(E):
    """The batched form at the batched workload's extreme points (100 K,
    1e-6 bar and 3000 K, 1e3 bar) for one of its species (1e5 lines, 16 amu)
    on the full 3e6-point grid: both forms against the oracle on 3000 random
    points, FP32C against FP64 everywhere."""
    import torch
    from heliosk import configs
    w = configs.WORKLOADS["c2"]
    arr = configs.species_arrays(w, 1)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    grid = _grid(E, w.numin, w.numax, w.dnu)
    sts = [E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
           for T, Pbar in (w.tp[0], w.tp[-1])]
    jj = np.sort(np.random.default_rng(3).choice(grid.nx, 3000, replace=False))
    Kos = [O.voigt_points(state_numpy(st), grid.numin, grid.dnu, grid.nx, jidx=jj) for st in sts]
    res = {}
    for mode in (E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C):
        outs = [torch.empty(grid.nx, dtype=torch.float64, device="cuda") for _ in sts]
        E.voigt_batched([(st, grid, 0.0, K) for st, K in zip(sts, outs)], mode=mode)
        res[mode] = [K.cpu().numpy() for K in outs]
        for K, Ko, tp in zip(res[mode], Kos, (w.tp[0], w.tp[-1])):
            e = rel_err(K[jj], Ko)
            print("batched %s at T %.0f K, P %.3g bar: vs oracle %.3e"
                  % ("FP64" if mode == E.HK_VOIGT_FP64 else "FP32C", tp[0], tp[1], e))
            assert e <= (FP64_TOL if mode == E.HK_VOIGT_FP64 else FP32C_TOL)
    for K32, K64 in zip(res[E.HK_VOIGT_FP32C], res[E.HK_VOIGT_FP64]):
        assert rel_err(K32, K64) <= FP32C_TOL


def test_side_stream_keeps_transient_buffers(E):
    """Calls queued on a stream other than torch's current one keep their
    transient workspaces until that stream is done: the side stream is held
    busy by a long sum first, the current stream then takes freed memory
    right away and zeroes it, and the side-stream results still equal the
    current-stream results bit for bit."""
    import torch
    d = random_lines(3000, seed=77, numin=990.0, numax=1030.0)
    sp = species_from(d)
    st = E.line_params(sp, 1500.0, 1.0, Q=d["iso_Q"], cut_mode=0, cut=10.0)
    grid = _grid(E, 1000.0, 1020.0, 0.0005)
    big = random_lines(200000, seed=78, numin=0.0, numax=3000.0)
    stb = E.line_params(species_from(big), 1500.0, 1.0, Q=big["iso_Q"], cut_mode=0, cut=50.0)
    gbig = _grid(E, 0.0, 3000.0, 0.001)
    modes = (E.HK_VOIGT_FP64, E.HK_VOIGT_FP32C)
    refs = {m: E.voigt(st, grid, 0.0, m).cpu() for m in modes}
    kref = E.ktable(refs[E.HK_VOIGT_FP64].cuda(), grid, 1020.0, 4, 8)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    Kb = E.voigt(stb, gbig, 0.0, E.HK_VOIGT_FP64, stream=side)      # keeps the side stream busy
    outs = {}
    for m in modes:
        outs[m] = E.voigt(st, grid, 0.0, m, stream=side)
        junk = [torch.zeros(1 << 20, dtype=torch.float64, device="cuda") for _ in range(8)]
        del junk
    kt = E.KTable(grid, 1020.0, 4, 8)
    Kin = refs[E.HK_VOIGT_FP64].cuda()
    side.wait_stream(torch.cuda.current_stream())
    c, n = kt(Kin, stream=side)
    torch.cuda.synchronize()
    for m in modes:
        assert torch.equal(outs[m].cpu(), refs[m]), m
    assert torch.equal(c.cpu(), kref[0].cpu()) and torch.equal(n.cpu(), kref[1].cpu())
    del Kb


def test_stage_lines_change_results_by_rounding_only(E):
    """512 lines per LDS stage against 256, every tile width, both forms,
    single and batched launches.  FP64 adds the lines in line order: bit-
    identical sums.  FP32C takes a stage's records path by path: the order
    changes with the stage size, so the sums agree to rounding (1e-6
    relative); the batched and single launches of one (tile, stage) give the
    same bits.  The autotuner returns
    a valid (tile, stage) pair that a second tuning of the same problem
    repeats from its cache."""
    import torch
    d = random_lines(40000, seed=48, numin=900.0, numax=1100.0)
    st = E.line_params(species_from(d), 1700.0, 3.0, Q=d["iso_Q"], cut_mode=0, cut=25.0)
    grid = _grid(E, 950.0, 1050.0, 0.001)
    for mode in (E.HK_VOIGT_FP32C, E.HK_VOIGT_FP64):
        for tp in (512, 1024, 2048, 4096):
            a = E.voigt(st, grid, 0.0, mode, tile_points=tp, stage_lines=256)
            b = E.voigt(st, grid, 0.0, mode, tile_points=tp, stage_lines=512)
            if mode == E.HK_VOIGT_FP64:
                assert torch.equal(a, b), (mode, tp)
            else:
                assert rel_err(b.cpu().numpy(), a.cpu().numpy()) <= 1e-6, (mode, tp)
        K = torch.empty(grid.nx, dtype=torch.float64, device="cuda")
        E.VoigtBatch([(st, grid, 0.0, K)], mode=mode, tile_points=1024, stage_lines=512).launch()
        torch.cuda.synchronize()
        assert torch.equal(K, E.voigt(st, grid, 0.0, mode, tile_points=1024, stage_lines=512))
    E.lib().hk_voigt_autotune_reset()
    t1, s1, ms1 = E.autotune(st, grid, E.HK_VOIGT_FP32C)
    t2, s2, ms2 = E.autotune(st, grid, E.HK_VOIGT_FP32C)
    assert t1 in (512, 1024, 2048, 4096) and s1 in (256, 512)
    assert (t2, s2) == (t1, s1) and all(v > 0 for v in ms1.values())
    assert all(v == 0 for v in ms2.values())
    print("autotuned: tile %d, stage %d; %s" % (t1, s1, ms1))
