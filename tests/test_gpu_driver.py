"""The heliosk driver end to end on the GPU: synthetic HITRAN-format inputs
(param.dat, <species>.param, partition tables, 68-byte line records) go
through the real front end, and its output files are checked against the
FP64 oracle fed with the same lines (readFile's strength scaling and
partition interpolation restated here from host.h:9-73, 598-649)."""
import os
import subprocess

import numpy as np
import pytest

from heliosk import files as F
from helpers import NA, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "helios-k_amd", "bin", "heliosk")

T_TABLE = np.arange(1.0, 3001.0, 1.0)


def q_table(k):
    return 1.0 + (0.05 + 0.01 * k) * T_TABLE ** 1.5


def q_at(T, Q):
    """Linear interpolation between the bracketing rows T0 < T <= T1."""
    j = int(np.searchsorted(T_TABLE, T, side="left"))
    if j == 0:
        return Q[0]
    t = (T - T_TABLE[j - 1]) / (T_TABLE[j] - T_TABLE[j - 1])
    return (Q[j] - Q[j - 1]) * t + Q[j - 1]


def make_species(tmp, name, cids, masses, abund, n, seed, numin, numax, s_scale=1.0):
    rng = np.random.default_rng(seed)
    lines = dict(nu=rng.uniform(numin, numax, n), S=s_scale * 10 ** rng.uniform(-25, -19, n),
                 EL=rng.uniform(0, 3000, n), A=rng.uniform(0, 2e7, n) * (rng.uniform(size=n) < 0.2),
                 delta=rng.uniform(-0.01, 0.01, n), gair=rng.uniform(0.01, 0.1, n),
                 gself=rng.uniform(0.1, 0.4, n), n=rng.uniform(0.3, 0.8, n))
    iso = rng.integers(0, len(cids), n)
    lines["cid"] = [cids[k] for k in iso]
    isos = [(c, a, 170.0 + 10 * k, 1, m) for k, (c, a, m) in enumerate(zip(cids, abund, masses))]
    F.species_files(str(tmp), name, lines, isos, T_TABLE, [q_table(k) for k in range(len(cids))],
                    limits=[int(numin) - 1, int(numax) + 1])
    return dict(lines=lines, iso=iso, masses=np.array(masses), abund=np.array(abund),
                Qtab=[q_table(k) for k in range(len(cids))])


def oracle_K(species, sscale, meanMass, T, P, q, gammaF, cut, numin, dnu, nx, kmin, units=0,
             compat=False):
    """Sum over species of the oracle cross section, with readFile's
    per-isotopologue scaling (host.h:628-644) and S2_kernel semantics
    (compat: in the reference's FP32 region arithmetic)."""
    K = np.full(nx, kmin)
    for sp, ss in zip(species, sscale):
        L = sp["lines"]
        m = sp["masses"]
        ab = sp["abund"] * (m / meanMass if units == 0 else 1.0)
        sc = np.full(len(m), ss) * (m / meanMass if units == 0 else 1.0)
        d = dict(nu=L["nu"], S=L["S"], EL=L["EL"], A=L["A"], delta=L["delta"],
                 vy=((1.0 - q) * L["gair"] + q * L["gself"]) * gammaF, n=L["n"],
                 iso=sp["iso"].astype(np.int32), iso_mass=m / NA, iso_ab=ab, iso_sscale=sc,
                 iso_Q=np.array([q_at(T, Q) for Q in sp["Qtab"]]))
        order = np.argsort(d["nu"], kind="stable")
        d = {k: (v[order] if isinstance(v, np.ndarray) and v.shape == order.shape else v)
             for k, v in d.items()}
        st = O.line_params(d, T, P, 0, cut)
        K += (O.voigt_compat_points(st, numin, dnu, nx) if compat
              else O.voigt_points(st, numin, dnu, nx, kmin=0.0))
    return K


def run_driver(cwd, *args):
    assert os.path.exists(DRIVER), "driver not built"
    print("driver: %s %s" % (cwd, " ".join(args)), flush=True)
    r = subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "Error" not in r.stdout, r.stdout
    return r.stdout


NUMIN, NUMAX, DNU, NBINS, NC = 1000.0, 1010.0, 0.001, 5, 16


def test_driver_text_outputs_vs_oracle(tmp_path):
    sp = make_species(tmp_path, "01_tst", ["11", "12"], [18.010565, 20.014811], [0.997317, 0.002],
                      3000, 1, 995.0, 1015.0)
    T, P, q, gF, cut = 1234.5, 0.7, 0.3, 1.1, 5.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="t", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doResampling=1, nC=NC,
                      doStoreFullK=1, nbins=NBINS, qalphaL=q, gammaF=gF, doStoreSK=1)
    out = run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = oracle_K([sp], [1.0], meanMass, T, P, q, gF, cut, NUMIN, DNU, nx, 0.0)
    blk = F.read_out_dat(str(tmp_path / "Out_t.dat"))
    assert len(blk) == 1 and blk[0].shape == (nx, 2)
    np.testing.assert_allclose(blk[0][:, 0], NUMIN + DNU * np.arange(nx), rtol=0, atol=1e-9)
    Kd = blk[0][:, 1]
    e = rel_err(Kd, Ko)
    print("driver K vs oracle %.3e" % e)
    assert e <= 1e-11
    # k-table rows against the oracle fit of the driver's own K
    seg = O.bin_segments(NUMIN, NUMAX, DNU, nx, NBINS)
    co, no = O.ktable(O.sort_segments(Kd, seg[:NBINS + 1]), nx // NBINS, NBINS, NC, 0.0)
    cb = F.read_cbin(str(tmp_path / "Out_t_cbin.dat"), NC)
    Nxb = nx // NBINS
    np.testing.assert_array_equal(cb["ystart"], np.minimum(no / (Nxb - 1.0), 1.0))
    ref = np.where((no == Nxb)[:, None], 0.0, co)
    assert np.max(np.abs(cb["coeff"] - ref)) <= 1e-10 * np.max(np.abs(ref))
    # K per bin: the reconstruction from the fitted coefficients
    kb = F.read_out_dat(str(tmp_path / "Out_t_bin.dat"))
    assert len(kb) == NBINS
    Ky = O.reconstruct(cb["coeff"], no, Nxb, 0.0)
    got = np.concatenate([b[:, 1] for b in kb])
    assert rel_err(got, Ky) <= 1e-12
    np.testing.assert_allclose(kb[0][:, 0], np.array(["%g" % (j / (Nxb - 1.0)) for j in range(Nxb)],
                                                      dtype=float))
    assert "Time for K(x)" in out


def test_driver_binary_mixture_pressures_and_kread(tmp_path):
    # Units = 1 writes K * M/N_A in float32: line strengths per gram (S * N_A / M)
    # keep that product in float32's normal range
    s1 = make_species(tmp_path, "01_a", ["11", "12"], [18.010565, 20.014811], [0.997317, 0.002],
                      2000, 2, 995.0, 1015.0, s_scale=NA / 18.0)
    s2 = make_species(tmp_path, "02_b", ["21"], [43.98983], [0.984204], 1500, 3, 990.0, 1020.0,
                      s_scale=NA / 44.0)
    (tmp_path / "P.txt").write_text("0.7\n3.0\n")
    (tmp_path / "species.txt").write_text("01_a 0.6\n02_b 0.4\n")
    T, q, gF, cut = 987.0, 1.0, 1.0, 4.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="m", T=T, P=1.0, PFile="P.txt",
                      SpeciesFile="species.txt", numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut,
                      doResampling=1, nC=NC, doStoreFullK=2, nbins=NBINS, qalphaL=q, gammaF=gF,
                      Units=1, kmin=1e-25)
    run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(0.6 * np.sum(s1["abund"] * s1["masses"]) + 0.4 * np.sum(s2["abund"] * s2["masses"]))
    unit = meanMass / NA
    Kf = F.read_out_bin(str(tmp_path / "Out_m.bin"), nx)
    assert Kf.shape == (2, nx)
    cb = F.read_cbin(str(tmp_path / "Out_m_cbin.dat"), NC)
    assert cb["coeff"].shape == (2 * NBINS, NC)
    np.testing.assert_array_equal(cb["TP"], np.repeat([[T, 0.7], [T, 3.0]], NBINS, axis=0))
    # Units = 1: kmin is divided by the unit scale and printed so (heliosk.cu:697-704)
    np.testing.assert_allclose(cb["kmin"], 1e-25 / unit, rtol=1e-15)
    for iP, P in enumerate((0.7, 3.0)):
        Ko = oracle_K([s1, s2], [0.6, 0.4], meanMass, T, P, q, gF, cut, NUMIN, DNU, nx,
                      1e-25 / unit, units=1)
        e = rel_err(Kf[iP].astype(np.float64), Ko * unit)
        print("P %g: float32 output vs oracle %.3e" % (P, e))
        assert e <= 1.2e-7
    # k-table-only run from the binary cross sections (doStoreFullK = -2)
    os.rename(tmp_path / "Out_m_cbin.dat", tmp_path / "first_cbin.dat")
    F.write_param_dat(str(tmp_path / "param.dat"), name="m", T=T, P=1.0, PFile="P.txt",
                      SpeciesFile="species.txt", numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut,
                      doResampling=1, nC=NC, doStoreFullK=-2, nbins=NBINS, qalphaL=q, gammaF=gF,
                      Units=1, kmin=1e-25, pathToK="./")
    run_driver(str(tmp_path))
    cb2 = F.read_cbin(str(tmp_path / "Out_m_cbin.dat"), NC)
    # the same k-tables up to the float32 rounding of the stored K
    np.testing.assert_array_equal(cb2["ystart"], cb["ystart"])
    assert np.max(np.abs(cb2["coeff"] - cb["coeff"])) <= 1e-5 * np.max(np.abs(cb["coeff"]))


def test_driver_fp32c_precision_key(tmp_path):
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 2500, 5, 995.0, 1015.0)
    T, P, cut = 1800.0, 2.0, 5.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="f", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=NBINS,
                      precision=1)
    run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = oracle_K([sp], [1.0], meanMass, T, P, 1.0, 1.0, cut, NUMIN, DNU, nx, 0.0)
    Kd = F.read_out_dat(str(tmp_path / "Out_f.dat"))[0][:, 1]
    e = rel_err(Kd, Ko)
    print("driver FP32C K vs oracle %.3e" % e)
    assert 0.0 < e <= 1e-5


def test_driver_compat_precision_key(tmp_path):
    """precision = 2: the driver sums in the reference's own FP32 region
    arithmetic (HK_VOIGT_COMPAT); its K follows the oracle's restatement of
    it, and differs from the FP64 oracle by the reference's region-B level."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 2500, 5, 995.0, 1015.0)
    T, P, cut = 1800.0, 2.0, 5.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="f", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=NBINS,
                      precision=2)
    out = run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    args = ([sp], [1.0], meanMass, T, P, 1.0, 1.0, cut, NUMIN, DNU, nx, 0.0)
    Kc = oracle_K(*args, compat=True)
    K64 = oracle_K(*args)
    Kd = F.read_out_dat(str(tmp_path / "Out_f.dat"))[0][:, 1]
    e, e64 = rel_err(Kd, Kc), rel_err(Kd, K64)
    print("driver compat K vs oracle compat %.3e, vs FP64 oracle %.3e" % (e, e64))
    assert e <= 2e-6
    assert 1e-4 <= e64 <= 1.5e-2


def _exomol_species(tmp, name, database, nfiles_limits, n, seed, T_table, ab=1.0, mass=18.010565,
                    gamma=0.07, nexp=0.5, npfcol=3):
    """ExoMol-family inputs: <name>.param with the #ExoMol block, a partition
    table (npfcol columns) and binary line files of 2, 4 or 5 doubles per line,
    named as ISO.h:125-175 names them."""
    rng = np.random.default_rng(seed)
    nd = F.EXOMOL_DOUBLES[database]
    lim = nfiles_limits
    nfiles = len(lim) - 1
    cols = dict(nu=[], S=[], EL=[], A=[], gN=[])
    counts = []
    for k in range(nfiles):
        nu = np.sort(rng.uniform(lim[k], lim[k + 1], n))
        c = dict(nu=nu, S=10 ** rng.uniform(-24, -19, n), EL=rng.uniform(0, 5000, n),
                 A=rng.uniform(0, 5e7, n) * (rng.uniform(size=n) < 0.3),
                 gN=rng.uniform(0, 2e7, n) * (rng.uniform(size=n) < 0.3))
        stem = F.species_param_name(name, database, nfiles, lim[k], lim[k + 1])
        F.write_exomol_bin(str(tmp / (stem + "bin")), database, c["nu"], c["S"], c["EL"], c["A"],
                           c["gN"])
        for key in cols:
            cols[key].append(c[key] if nd > 2 or key in ("nu", "S") else np.zeros(n))
        counts.append(n)
    Q = 1.0 + 0.3 * T_table ** 1.4
    with open(tmp / ("%s.pf" % name), "w") as f:
        for t, q in zip(T_table, Q):
            f.write(("%8.1f %16.6f %d\n" % (t, q, 1)) if npfcol == 3 else ("%8.1f %16.6f\n" % (t, q)))
    F.write_species_param(str(tmp / (name + ".param")), name,
                          [("0", ab, 0.0, 0, mass, "%s.pf" % name)], counts, lim,
                          database=database, npfcol=npfcol, default_gamma=gamma, default_n=nexp)
    lines = {k: np.concatenate(v) for k, v in cols.items()}
    if nd < 5:
        lines["gN"] = np.zeros_like(lines["nu"])
    if nd < 4:
        lines["A"] = np.zeros_like(lines["nu"])
        lines["EL"] = np.zeros_like(lines["nu"])
    return dict(lines=lines, Q=Q, ab=ab, mass=mass, gamma=gamma, nexp=nexp)


def oracle_exomol_K(sp, form, T, P, gammaF, cut, numin, dnu, nx):
    """Sf_kernel / SfSuper_kernel (voigt.h:365-410) with the L_kernel* line
    set-up (voigt.h:484-578): Units = 0 scaling of one isotopologue, whose
    abundance-weighted molar mass is the mean mass (Ab m / meanMass = 1 and
    Sscale m / meanMass = 1 / Ab)."""
    L = sp["lines"]
    nl = L["nu"].size
    q = q_at_table(T, sp["Q"])
    d = dict(nu=L["nu"], S=L["S"], EL=L["EL"], A=L["A"] + L["gN"], delta=np.zeros(nl),
             vy=np.full(nl, sp["gamma"] * gammaF), n=np.full(nl, sp["nexp"]), iso=None,
             iso_mass=np.array([sp["mass"] / NA]), iso_ab=np.array([1.0]),
             iso_sscale=np.array([1.0 / sp["ab"]]), iso_Q=np.array([q]))
    order = np.argsort(d["nu"], kind="stable")
    d = {k: (v[order] if isinstance(v, np.ndarray) and v.shape == order.shape else v)
         for k, v in d.items()}
    st = O.line_params(d, T, P, 0, cut, form=form)
    return O.voigt_points(st, numin, dnu, nx)


def q_at_table(T, Q):
    return q_at(T, Q)


@pytest.mark.parametrize("database, limits, ab, pfile", [
    (2, [990, 1005, 1020], 1.0, True),       # ExoMol, two files, two pressures
    (20, [990, 1020], 0.9, False),           # ExoMol super-lines
    (30, [990, 1020], 1.0, False),           # Kurucz atoms (A + Gamma_nat)
])
def test_driver_exomol_family_vs_oracle(tmp_path, database, limits, ab, pfile):
    name = "1H2-16O__TST"
    sp = _exomol_species(tmp_path, name, database, limits, 1500, 20 + database, T_TABLE, ab=ab)
    T, gF, cut = 1400.0, 1.2, 5.0
    Ps = [0.5, 4.0] if pfile else [2.0]
    kw = {}
    if pfile:
        (tmp_path / "P.txt").write_text("".join("%g\n" % v for v in Ps))
        kw["PFile"] = "P.txt"
    F.write_param_dat(str(tmp_path / "param.dat"), name="x", T=T, P=Ps[0], Species_Name=name,
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=2, nbins=NBINS,
                      gammaF=gF, **kw)
    out = run_driver(str(tmp_path))
    nfiles = len(limits) - 1
    assert "File range 0 to %d" % (nfiles - 1) in out
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    Kf = F.read_out_bin(str(tmp_path / "Out_x.bin"), nx)
    assert Kf.shape == (len(Ps), nx)
    form = 2 if database == 20 else 1
    for iP, P in enumerate(Ps):
        Ko = oracle_exomol_K(sp, form, T, P, gF, cut, NUMIN, DNU, nx)
        e = rel_err(Kf[iP].astype(np.float64), Ko)
        print("database %d, P %g: float32 output vs oracle %.3e" % (database, P, e))
        assert e <= 1.2e-7


def test_driver_chunked_stream_matches_oracle(tmp_path):
    """Three HITRAN files of unsorted lines streamed in chunks of 777 lines
    (chunks cross file boundaries; each is sorted on the device): the same
    cross sections as one chunk, and both against the oracle."""
    sp = make_species(tmp_path, "01_chk", ["11", "12", "13"], [18.010565, 20.014811, 19.01478],
                      [0.997317, 0.002, 0.000371884], 4000, 31, 995.0, 1015.0)
    F.species_files(str(tmp_path), "01_chk", sp["lines"],
                    [(c, a, 170.0 + 10 * k, 1, m) for k, (c, a, m) in
                     enumerate(zip(["11", "12", "13"], sp["abund"], sp["masses"]))],
                    T_TABLE, [q_table(k) for k in range(3)], limits=[990, 1000, 1007, 1020])
    assert (tmp_path / "01_chk_00990-01000.bin").exists()
    T, P, cut = 1500.0, 1.3, 5.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="c", T=T, P=P, Species_Name="01_chk",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=NBINS)
    out = run_driver(str(tmp_path), "-maxlines", "777")
    assert "Line chunk 4: 777 lines" in out and "Line chunk 5: 115 lines" in out, out
    K1 = F.read_out_dat(str(tmp_path / "Out_c.dat"))[0][:, 1]
    run_driver(str(tmp_path))
    K2 = F.read_out_dat(str(tmp_path / "Out_c.dat"))[0][:, 1]
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = oracle_K([sp], [1.0], meanMass, T, P, 1.0, 1.0, cut, NUMIN, DNU, nx, 0.0)
    e1, e2, e12 = rel_err(K1, Ko), rel_err(K2, Ko), rel_err(K1, K2)
    print("chunked %.3e, one chunk %.3e, between %.3e" % (e1, e2, e12))
    assert e1 <= 1e-11 and e2 <= 1e-11 and e12 <= 1e-13


def _read_rows(path):
    return [l.split() for l in open(path) if l.strip()]


def test_driver_means_transmission_and_bins(tmp_path):
    """doMean = 1, doTransmission = 2 and doStoreSK = 1 without resampling:
    means of K(x) and the transmission of each sorted bin, against the oracle
    applied to the driver's own K (text outputs carry all digits) and to the
    oracle's K."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 3000, 41, 995.0, 1015.0)
    T, P, cut, nTr, dTr = 1100.0, 0.8, 5.0, 300, 0.07
    F.write_param_dat(str(tmp_path / "param.dat"), name="t", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=NBINS,
                      doMean=1, doTransmission=2, nTr=nTr, dTr=dTr, doStoreSK=1)
    out = run_driver(str(tmp_path))
    assert "Time for Transmission" in out and "Time for mean K(x)" in out
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    Nxb = nx // NBINS
    Kd = F.read_out_dat(str(tmp_path / "Out_t.dat"))[0][:, 1]
    from test_oracle import mean_lines
    means = np.loadtxt(tmp_path / "Out_t_mean.dat")
    np.testing.assert_allclose(means, mean_lines(O.means(Kd, NUMIN, DNU, T), T, DNU), rtol=1e-13)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = oracle_K([sp], [1.0], meanMass, T, P, 1.0, 1.0, cut, NUMIN, DNU, nx, 0.0)
    np.testing.assert_allclose(means, mean_lines(O.means(Ko, NUMIN, DNU, T), T, DNU), rtol=1e-11)
    # K(y) per bin = the sorted bins of K(x)
    kb = F.read_out_dat(str(tmp_path / "Out_t_bin.dat"))
    Ky = np.concatenate([b[:, 1] for b in kb])
    np.testing.assert_array_equal(Ky, np.concatenate([np.sort(Kd[b * Nxb:(b + 1) * Nxb])
                                                      for b in range(NBINS)]))
    To = O.transmission(Ky, NBINS, Nxb, nTr, dTr)
    for b in range(NBINS):
        tr = np.loadtxt(tmp_path / ("Out_t_tr%04d.dat" % b))
        assert tr.shape == (nTr, 2)
        np.testing.assert_allclose(tr[:, 0], np.exp((np.arange(nTr) - nTr // 2) * dTr), rtol=1e-15)
        np.testing.assert_allclose(tr[:, 1], To[b], rtol=1e-13)


def test_driver_resampled_transmission_uses_last_nxmin(tmp_path):
    """Resampling, two pressures and doTransmission = 1: the transmission of
    each pressure is taken from its reconstructed K(y) and blended with the
    Nxmin of the last pressure fitted (heliosk.cu:2130-2193, 2400-2410).
    Cut in Lorentz widths makes the empty part of each bin depend on P."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 400, 43, 1000.0, 1004.0)
    T, cut, nTr, dTr, nC, kmin = 700.0, 30.0, 64, 0.3, 12, 1e-28
    (tmp_path / "P.txt").write_text("0.05\n1.5\n")
    F.write_param_dat(str(tmp_path / "param.dat"), name="r", T=T, P=1.0, PFile="P.txt",
                      Species_Name="01_tst", numin=NUMIN, numax=NUMAX, dnu=DNU, cutMode=1,
                      cut=cut, doStoreFullK=0, nbins=NBINS, doResampling=1, nC=nC, kmin=kmin,
                      doTransmission=1, nTr=nTr, dTr=dTr)
    run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    Nxb = nx // NBINS
    cb = F.read_cbin(str(tmp_path / "Out_r_cbin.dat"), nC)
    # ystart = min(Nxmin / (Nxb - 1), 1); a bin of kmin only (Nxmin = Nxb) prints zeros
    empty = (cb["ystart"] == 1.0) & np.all(cb["coeff"] == 0.0, axis=1)
    nxmin = np.where(empty, Nxb, np.rint(cb["ystart"] * (Nxb - 1))).astype(np.int32)
    nxmin = nxmin.reshape(2, NBINS)
    assert not np.array_equal(nxmin[0], nxmin[1]), nxmin      # the quirk is visible
    coeff = cb["coeff"].reshape(2, NBINS, nC)
    rows = np.array([[float(v) for v in r] for r in _read_rows(tmp_path / "Out_r_tr.dat")])
    assert rows.shape == (2 * NBINS * nTr, 5)
    for iP in range(2):
        Ky = O.reconstruct(coeff[iP], nxmin[iP], Nxb, kmin)
        To = O.transmission(Ky, NBINS, Nxb, nTr, dTr, kmin=kmin, nxmin=nxmin[1])
        got = rows[iP * NBINS * nTr:(iP + 1) * NBINS * nTr, 1].reshape(NBINS, nTr)
        np.testing.assert_allclose(got, To, rtol=1e-12)
        np.testing.assert_array_equal(rows[iP * NBINS * nTr:(iP + 1) * NBINS * nTr, 4],
                                      np.tile(np.arange(nTr), NBINS))


def test_driver_output_edges(tmp_path):
    """doStoreSK = 2 with an output-edges file, two pressures: the per-bin
    rows are the edge-interval means of the plain K(y) rows."""
    from test_driver_cpu import edge_means
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 1500, 47, 995.0, 1015.0)
    (tmp_path / "P.txt").write_text("0.3\n2.0\n")
    edges = [0.0, 0.1, 0.35, 0.6, 0.85, 0.999999, 1.0]
    (tmp_path / "edges.txt").write_text("".join("%r\n" % e for e in edges))
    common = dict(name="e", T=900.0, P=1.0, PFile="P.txt", Species_Name="01_tst", numin=NUMIN,
                  numax=NUMAX, dnu=DNU, cut=5.0, doStoreFullK=0, nbins=NBINS, doStoreSK=2)
    F.write_param_dat(str(tmp_path / "param.dat"), **common)
    run_driver(str(tmp_path))
    plain = [np.array([[float(v) for v in r] for r in _read_rows(tmp_path / ("Out_e_bin%04d.dat" % b))])
             for b in range(NBINS)]
    F.write_param_dat(str(tmp_path / "param.dat"), OutputEdgesFile="edges.txt", **common)
    out = run_driver(str(tmp_path))
    assert "Use output edges in file edges.txt" in out
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    Nxb = nx // NBINS
    for b in range(NBINS):
        got = [[float(v) for v in r] for r in _read_rows(tmp_path / ("Out_e_bin%04d.dat" % b))]
        expect = []
        for iP, P in enumerate((0.3, 2.0)):
            Kb = plain[b][iP * Nxb:(iP + 1) * Nxb, 1]
            for k, (mid, v) in enumerate(edge_means(None, Kb, edges)):
                expect.append([float("%g" % mid), v, 900.0, P, k])
        assert len(got) == len(expect) == 2 * (len(edges) - 1)
        np.testing.assert_allclose(np.array(got), np.array(expect), rtol=1e-15)


@pytest.mark.parametrize("profile, plinth", [(2, 0), (3, 1), (4, 0), (1, 1)])
def test_driver_profiles_and_plinth(tmp_path, profile, plinth):
    sp = make_species(tmp_path, "01_tst", ["11", "12"], [18.010565, 20.014811], [0.997317, 0.002],
                      2000, 50 + profile, 995.0, 1015.0)
    T, P, cut = 1300.0, 0.6, 2.0
    F.write_param_dat(str(tmp_path / "param.dat"), name="pr", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=NBINS,
                      profile=profile, removePlinth=plinth)
    run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    Kd = F.read_out_dat(str(tmp_path / "Out_pr.dat"))[0][:, 1]
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    L = sp["lines"]
    m = sp["masses"]
    d = dict(nu=L["nu"], S=L["S"], EL=L["EL"], A=L["A"], delta=L["delta"], vy=L["gself"],
             n=L["n"], iso=sp["iso"].astype(np.int32), iso_mass=m / NA,
             iso_ab=sp["abund"] * m / meanMass, iso_sscale=m / meanMass,
             iso_Q=np.array([q_at(T, Q) for Q in sp["Qtab"]]))
    st = O.line_params(d, T, P, 0, cut)
    Ko = O.profile_points(st, NUMIN, DNU, nx, profile, remove_plinth=bool(plinth))
    e = rel_err(Kd, Ko)
    print("driver profile %d plinth %d vs oracle %.3e" % (profile, plinth, e))
    assert e <= 1e-11


def test_driver_points_per_bin(tmp_path):
    """"Nnu per bin": each of the equal bins gets its own uniform points
    (setX_kernel, voigt.h:751-767); K, the trapezoid means and the per-bin
    K(y) against the oracle summed on each bin's grid."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 2000, 61, 995.0, 1015.0)
    T, P, cut, nbins, Nxb = 1400.0, 0.9, 5.0, 5, 1733
    F.write_param_dat(str(tmp_path / "param.dat"), name="x", T=T, P=P, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=cut, doStoreFullK=1, nbins=nbins,
                      Nnu_per_bin=Nxb, doMean=1, doStoreSK=1)
    run_driver(str(tmp_path))
    blk = F.read_out_dat(str(tmp_path / "Out_x.dat"))[0]
    assert blk.shape == (nbins * Nxb, 2)
    bnd = [NUMIN + i * (NUMAX - NUMIN) / nbins for i in range(nbins)] + [NUMAX]
    x = np.concatenate([bnd[b] + np.arange(Nxb) * ((bnd[b + 1] - bnd[b]) / Nxb)
                        for b in range(nbins)])
    np.testing.assert_allclose(blk[:, 0], x, rtol=2e-16, atol=0)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = np.concatenate([oracle_K([sp], [1.0], meanMass, T, P, 1.0, 1.0, cut, bnd[b],
                                  (bnd[b + 1] - bnd[b]) / Nxb, Nxb, 0.0) for b in range(nbins)])
    e = rel_err(blk[:, 1], Ko)
    print("per-bin points: K vs oracle %.3e" % e)
    assert e <= 1e-11
    from test_oracle import mean_lines
    means = np.loadtxt(tmp_path / "Out_x_mean.dat")
    # the reference multiplies the trapezoid integrals by param dnu once more
    np.testing.assert_allclose(means, mean_lines(O.means_x(blk[:, 1], blk[:, 0], T), T, DNU),
                               rtol=1e-12)
    kb = F.read_out_dat(str(tmp_path / "Out_x_bin.dat"))
    assert len(kb) == nbins
    for b in range(nbins):
        np.testing.assert_array_equal(kb[b][:, 1], np.sort(blk[b * Nxb:(b + 1) * Nxb, 1]))


def test_driver_bins_file(tmp_path):
    """Irregular bins from a bins file on the uniform grid: the per-bin
    files hold each bin's own number of sorted points (the useIndividualBins
    branch, heliosk.cu:2292-2345)."""
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 1500, 63, 995.0, 1015.0)
    bnd = [1000.0, 1001.5, 1004.0, 1007.25, 1010.0]
    (tmp_path / "bins.txt").write_text("".join("%r\n" % v for v in bnd))
    (tmp_path / "P.txt").write_text("0.4\n2.5\n")
    F.write_param_dat(str(tmp_path / "param.dat"), name="b", T=800.0, P=1.0, PFile="P.txt",
                      Species_Name="01_tst", numin=0.0, numax=1.0, dnu=DNU, cut=5.0,
                      doStoreFullK=1, nbins=7, binsFile="bins.txt", doStoreSK=2)
    out = run_driver(str(tmp_path))
    assert "Use bins in file bins.txt" in out
    info = open(tmp_path / "Info_b.dat").read()
    assert "nbins = 4" in info and "use Individual bins: bins.txt" in info
    blocks = F.read_out_dat(str(tmp_path / "Out_b.dat"))
    assert len(blocks) == 2 and blocks[0].shape == (10000, 4)
    for b in range(4):
        rows = np.array([[float(v) for v in r] for r in _read_rows(tmp_path / ("Out_b_bin%04d.dat" % b))])
        x = blocks[0][:, 0]
        sel = (x >= bnd[b]) & (x < bnd[b + 1])
        N = int(sel.sum())
        assert rows.shape == (2 * N, 5)
        for iP in range(2):
            Kb = np.sort(blocks[iP][sel, 1])
            r = rows[iP * N:(iP + 1) * N]
            np.testing.assert_array_equal(r[:, 1], Kb)
            np.testing.assert_allclose(r[:, 0], [float("%g" % (j / (N - 1.0))) for j in range(N)])
            np.testing.assert_array_equal(r[:, 4], np.arange(N))


@pytest.mark.parametrize("T, units", [(1500.0, 0), (700.0, 1), (3200.0, 0)])
def test_driver_cia(tmp_path, T, units):
    """ciaSystem = H2-H: K from the HITRAN CIA table (readCiaFile,
    host.h:670-789) instead of line lists, two pressures, against the
    oracle's restatement; T above every set gives kmin."""
    rng = np.random.default_rng(int(T))
    sets = []
    for Ts in (500.0, 1000.0, 2000.0, 3000.0):
        nu = np.round(np.concatenate([[995.0], np.sort(rng.uniform(995.0, 1012.0, 300)), [1012.0]]), 4)
        sets.append((Ts, nu, 10 ** rng.uniform(-47, -44, nu.size)))
    F.write_cia(str(tmp_path / "H2-H_2011.cia"), sets, formula="H2-H")
    (tmp_path / "P.txt").write_text("0.5\n3.0\n")
    F.write_param_dat(str(tmp_path / "param.dat"), name="c", T=T, P=1.0, PFile="P.txt",
                      ciaSystem="H2-H", Species_Name="none", numin=NUMIN, numax=NUMAX, dnu=DNU,
                      cut=5.0, doStoreFullK=1, nbins=NBINS, kmin=1e-40, Units=units)
    run_driver(str(tmp_path))
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    x = NUMIN + DNU * np.arange(nx)
    blocks = F.read_out_dat(str(tmp_path / "Out_c.dat"))
    unit = 1.00794 / NA if units else 1.0
    kmin = 1e-40 / unit
    for iP, P in enumerate((0.5, 3.0)):
        Ko = O.cia_K(str(tmp_path / "H2-H_2011.cia"), 4, blocks[iP][:, 0], T, P, 1.00794, kmin)
        np.testing.assert_allclose(blocks[iP][:, 0], x, rtol=1e-15)
        np.testing.assert_allclose(blocks[iP][:, 1], Ko * unit, rtol=1e-14, atol=0)
        if T > 3000.0:
            assert np.all(blocks[iP][:, 1] == kmin * unit)


def test_driver_sublorentzian(tmp_path):
    """subLorentzianfile: Perrin-Hartmann wings inside the file's range, on
    the per-bin points the reference uses with this option; against the
    oracle on each bin's grid with B_i at T rounded to float as there."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 1500, 71, 2180.0, 2320.0)
    (tmp_path / "chi.dat").write_text(
        "sigma1 3.0\nsigma2 30.0\nsigma3 120.0\n\nB\talpha \tbeta \tepsilon\n"
        "B1\t0.0888\t-0.160\t0.00410\nB2\t0.0\t0.0526\t0.00152\nB3\t0.0232\t0.0\t0.0\n\n"
        "range in cm^-1\n2210 2290\n")
    T, P, cut, numin, numax, dnu, nbins = 650.0, 3.0, 40.0, 2200.0, 2300.0, 0.01, 4
    F.write_param_dat(str(tmp_path / "param.dat"), name="s", T=T, P=P, Species_Name="01_tst",
                      numin=numin, numax=numax, dnu=dnu, cut=cut, doStoreFullK=1, nbins=nbins,
                      subLorentzianfile="chi.dat")
    out = run_driver(str(tmp_path))
    assert "Use sub-Lorentzian file chi.dat" in out
    blk = F.read_out_dat(str(tmp_path / "Out_s.dat"))[0]
    f32 = lambda v: float(np.float32(v))
    ab = [(f32(0.0888), f32(-0.160), f32(0.00410)), (0.0, f32(0.0526), f32(0.00152)),
          (f32(0.0232), 0.0, 0.0)]
    chi = [3.0, 30.0, 120.0] + [f32(a + b * np.exp(-e * T)) for a, b, e in ab] + [2210.0, 2290.0]
    nx = int((numax - numin) / dnu + 0.5)
    Nxb = nx // nbins
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    L, m = sp["lines"], sp["masses"]
    d = dict(nu=L["nu"], S=L["S"], EL=L["EL"], A=L["A"], delta=L["delta"], vy=L["gself"],
             n=L["n"], iso=sp["iso"].astype(np.int32), iso_mass=m / NA,
             iso_ab=sp["abund"] * m / meanMass, iso_sscale=m / meanMass,
             iso_Q=np.array([q_at(T, Q) for Q in sp["Qtab"]]))
    st = O.line_params(d, T, P, 0, cut)
    bnd = [numin + i * (numax - numin) / nbins for i in range(nbins)] + [numax]
    Ko = np.concatenate([O.profile_points(st, bnd[b], (bnd[b + 1] - bnd[b]) / Nxb, Nxb, 1, chi=chi)
                         for b in range(nbins)])
    e = rel_err(blk[:, 1], Ko)
    print("driver sub-Lorentzian vs oracle %.3e" % e)
    assert e <= 1e-11


def _run_any(cwd):
    """The driver's stdout, errors allowed (the reference exits with status 0
    after printing Error: ...)."""
    r = subprocess.run([DRIVER], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    return r.stdout


def test_driver_per_bin_cbin_files_and_append(tmp_path):
    """doResampling = 2 writes one _cbin%04d.dat per bin (heliosk.cu:2166-2187):
    for each pressure the bin's row -- kmin, ystart, nC coefficients, T, P --
    ending with a newline except in the last bin's file; the rows equal
    doResampling = 1's.  ReplaceFiles = 0 appends a second run's rows,
    ReplaceFiles = 1 starts the files again."""
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 2000, 31, 995.0, 1015.0)
    (tmp_path / "P.txt").write_text("0.5\n2.0\n")
    common = dict(name="r", T=1400.0, P=1.0, PFile="P.txt", Species_Name="01_tst", numin=NUMIN,
                  numax=NUMAX, dnu=DNU, cut=5.0, nC=NC, doStoreFullK=0, nbins=NBINS)
    F.write_param_dat(str(tmp_path / "param.dat"), doResampling=1, **common)
    run_driver(str(tmp_path))
    rows = F.read_cbin(str(tmp_path / "Out_r_cbin.dat"), NC)        # pressure-major rows
    F.write_param_dat(str(tmp_path / "param.dat"), doResampling=2, ReplaceFiles=1, **common)
    run_driver(str(tmp_path))

    def bin_tokens():
        out = []
        for b in range(NBINS):
            with open(tmp_path / ("Out_r_cbin%04d.dat" % b)) as f:
                text = f.read()
            out.append((text, np.array([float(v) for v in text.split()])))
        return out

    first = bin_tokens()
    width = 2 + NC + 2
    for b, (text, tok) in enumerate(first):
        assert tok.size == 2 * width
        for iP in range(2):
            got = tok[iP * width:(iP + 1) * width]
            ref = np.concatenate([[rows["kmin"][iP * NBINS + b], rows["ystart"][iP * NBINS + b]],
                                  rows["coeff"][iP * NBINS + b], rows["TP"][iP * NBINS + b]])
            np.testing.assert_array_equal(got, ref)
        assert text.count("\n") == (2 if b < NBINS - 1 else 0)
    # a second run appends (ReplaceFiles = 0) ...
    F.write_param_dat(str(tmp_path / "param.dat"), doResampling=2, ReplaceFiles=0, **common)
    run_driver(str(tmp_path))
    for (t1, k1), (t2, k2) in zip(first, bin_tokens()):
        assert t2 == t1 + t1
    # ... and ReplaceFiles = 1 starts over
    F.write_param_dat(str(tmp_path / "param.dat"), doResampling=2, ReplaceFiles=1, **common)
    run_driver(str(tmp_path))
    for (t1, _), (t2, _) in zip(first, bin_tokens()):
        assert t2 == t1


def test_driver_replace_files_zero_appends_cross_sections(tmp_path):
    """ReplaceFiles = 0 opens Out_<name>.dat / .bin for appending
    (heliosk.cu:373-379): two runs leave two copies, ReplaceFiles = 1 one."""
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 800, 33, 995.0, 1015.0)
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    for fullk, reader in ((1, lambda: F.read_out_dat(str(tmp_path / "Out_a.dat"))),
                          (2, lambda: F.read_out_bin(str(tmp_path / "Out_a.bin"), nx))):
        common = dict(name="a", T=900.0, P=1.0, Species_Name="01_tst", numin=NUMIN, numax=NUMAX,
                      dnu=DNU, cut=5.0, doStoreFullK=fullk, nbins=NBINS)
        F.write_param_dat(str(tmp_path / "param.dat"), ReplaceFiles=1, **common)
        run_driver(str(tmp_path))
        one = reader()
        F.write_param_dat(str(tmp_path / "param.dat"), ReplaceFiles=0, **common)
        run_driver(str(tmp_path))
        two = reader()
        assert len(two) == 2 * len(one)
        for blk in two:
            np.testing.assert_array_equal(blk, one[0])
        F.write_param_dat(str(tmp_path / "param.dat"), ReplaceFiles=1, **common)
        run_driver(str(tmp_path))
        assert len(reader()) == len(one)


def test_driver_fullk_text_reload(tmp_path):
    """doStoreFullK = -1 reads the cross sections back from
    <pathToK>Out_<name>.dat (heliosk.cu:1862-1935) and runs only the k-table:
    %.20g keeps every double, so the k-tables equal the first run's bit for
    bit, also with two pressures (four columns)."""
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 1500, 35, 995.0, 1015.0)
    (tmp_path / "P.txt").write_text("0.3\n4.0\n")
    os.mkdir(tmp_path / "kdir")
    common = dict(name="k", T=1600.0, P=1.0, PFile="P.txt", Species_Name="01_tst", numin=NUMIN,
                  numax=NUMAX, dnu=DNU, cut=5.0, nC=NC, doResampling=1, nbins=NBINS, kmin=1e-24)
    F.write_param_dat(str(tmp_path / "param.dat"), doStoreFullK=1, **common)
    run_driver(str(tmp_path))
    first = open(tmp_path / "Out_k_cbin.dat").read()
    os.rename(tmp_path / "Out_k.dat", tmp_path / "kdir" / "Out_k.dat")
    os.remove(tmp_path / "Out_k_cbin.dat")
    F.write_param_dat(str(tmp_path / "param.dat"), doStoreFullK=-1, pathToK="kdir/", **common)
    out = run_driver(str(tmp_path))
    assert open(tmp_path / "Out_k_cbin.dat").read() == first
    # a file with too few rows is refused with the reference's message
    blk = open(tmp_path / "kdir" / "Out_k.dat").read().splitlines()
    (tmp_path / "kdir" / "Out_k.dat").write_text("\n".join(blk[:100]) + "\n")
    out = _run_any(str(tmp_path))
    assert "Error: Input file too short kdir/Out_k.dat" in out, out


@pytest.mark.parametrize("T,msg", [(0.5, "Error: partition function not valid for given temperature. T1: 1, T: 0.5"),
                                   (3500.0, "Error: partition function not valid for given temperature. T0: 3000, T: 3500"),
                                   (1.0, None), (2999.5, None)])
def test_driver_partition_range(tmp_path, T, msg):
    """readPartition's rules (host.h:9-73): T below the table's first row or
    beyond its last is refused with the reference's messages; T on the first
    row and between the last two rows are interpolated (q_at)."""
    sp = make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 300, 37, 995.0, 1015.0)
    F.write_param_dat(str(tmp_path / "param.dat"), name="q", T=T, P=1.0, Species_Name="01_tst",
                      numin=NUMIN, numax=NUMAX, dnu=DNU, cut=5.0, doStoreFullK=1, nbins=NBINS)
    out = _run_any(str(tmp_path))
    if msg is not None:
        assert msg in out, out
        return
    assert "Error" not in out, out
    nx = int((NUMAX - NUMIN) / DNU + 0.5)
    meanMass = float(np.sum(sp["abund"] * sp["masses"]))
    Ko = oracle_K([sp], [1.0], meanMass, T, 1.0, 1.0, 1.0, 5.0, NUMIN, DNU, nx, 0.0)
    assert rel_err(F.read_out_dat(str(tmp_path / "Out_q.dat"))[0][:, 1], Ko) <= 1e-11


def test_driver_partition_unreadable(tmp_path):
    """An empty partition table cannot be read; a table with a text row stops
    there (the reference's fscanf loop), so T beyond it is not covered."""
    make_species(tmp_path, "01_tst", ["11"], [18.010565], [0.997317], 300, 38, 995.0, 1015.0)
    F.write_param_dat(str(tmp_path / "param.dat"), name="q", T=1500.0, P=1.0,
                      Species_Name="01_tst", numin=NUMIN, numax=NUMAX, dnu=DNU, cut=5.0,
                      doStoreFullK=1, nbins=NBINS)
    qfile = "q_01_tst_0.txt"                     # F.species_files' partition table
    assert os.path.exists(tmp_path / qfile)
    (tmp_path / qfile).write_text("")
    assert "Error: partition function can not be read" in _run_any(str(tmp_path))
    (tmp_path / qfile).write_text("1 2\n100 3\nabc def\n2000 4\n")
    out = _run_any(str(tmp_path))
    assert "Error: partition function not valid for given temperature. T0: 100, T: 1500" in out, out
