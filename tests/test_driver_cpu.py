"""CPU checks of the heliosk driver's front end: the file writers round-trip,
and param.dat / console parsing and validation reject what the reference
rejects (with its messages and its exit status 0) before any device is
touched."""
import os
import subprocess

import numpy as np
import pytest

from heliosk import files as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "helios-k_amd", "bin", "heliosk")


def run_driver(cwd, *args):
    if not os.path.exists(DRIVER):
        pytest.fail("driver not built: %s (run __graft_entry__.build())" % DRIVER)
    r = subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_hitran_record_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    n = 50
    d = dict(nu=np.sort(rng.uniform(0, 100, n)), S=10 ** rng.uniform(-25, -19, n),
             EL=rng.uniform(0, 3000, n), A=rng.uniform(0, 1, n), delta=rng.uniform(-0.01, 0.01, n),
             gair=rng.uniform(0.01, 0.1, n), gself=rng.uniform(0.1, 0.5, n), n=rng.uniform(0.3, 0.8, n))
    cid = ["11" if k % 3 else "12" for k in range(n)]
    p = tmp_path / "x.bin"
    F.write_hitran_bin(str(p), cid, **d)
    assert os.path.getsize(p) == 68 * n
    rec = F.read_hitran_bin(str(p))
    for k, v in d.items():
        np.testing.assert_array_equal(rec[k], v)
    assert rec["cid"][1] == b" 11" and rec["cid"][0] == b" 12"
    raw = open(p, "rb").read()
    assert raw[:4] == b" 12\0"


@pytest.mark.parametrize("database, nd", [(2, 4), (20, 2), (30, 5), (31, 5), (32, 5)])
def test_exomol_record_layout(tmp_path, database, nd):
    rng = np.random.default_rng(database)
    cols = [rng.uniform(0, 1, 7) for _ in range(5)]
    p = tmp_path / "x.bin"
    F.write_exomol_bin(str(p), database, *cols)
    raw = np.fromfile(p, dtype="<f8")
    assert raw.size == 7 * nd
    np.testing.assert_array_equal(raw.reshape(7, nd), np.stack(cols[:nd], axis=1))


def test_line_file_names():
    # ISO.h:125-175
    assert F.species_param_name("01_hit16", 0, 2, 0, 250) == "01_hit16_00000-00250."
    assert F.species_param_name("1H2-16O__BT2", 2, 16, 250, 500) == "1H2-16O__BT2__00250-00500."
    assert F.species_param_name("CO2_cdsd", 3, 3, 500, 1000) == "CO2_cdsd__00500_01000."
    assert F.species_param_name("x", 20, 3, 0, 10) == "x."
    assert F.species_param_name("x", 2, 1, 0, 10) == "x."


def _setup(tmp_path, **kw):
    F.write_param_dat(str(tmp_path / "param.dat"), **kw)


@pytest.mark.parametrize("kw, msg", [
    (dict(profile=5), "Error: profile 5 is not supported"),
    (dict(profile=4, removePlinth=1), "Error, remove plinth is not supported for profile 4"),
    (dict(numin=0.0, numax=10.0, dnu=0.01, nbins=7), "Error: range cannot be divided evenly in bins."),
])
def test_driver_rejects_like_the_reference(tmp_path, kw, msg):
    _setup(tmp_path, **kw)
    rc, out = run_driver(str(tmp_path))
    assert rc == 0                         # the reference exits 0 after "Error: ..."
    assert msg in out, out


def test_driver_rejects_unordered_output_edges(tmp_path):
    (tmp_path / "edges.txt").write_text("0.0\n0.5\n0.4\n1.0\n")
    _setup(tmp_path, OutputEdgesFile="edges.txt", doStoreSK=1)
    rc, out = run_driver(str(tmp_path))
    assert "Use output edges in file edges.txt" in out
    assert "Error; output edges not monotonic growing" in out, out
    _setup(tmp_path, OutputEdgesFile="missing.txt", doStoreSK=1)
    rc, out = run_driver(str(tmp_path))
    assert "Error: output edges file not found: missing.txt" in out, out


def edge_means(y, K, edges):
    """Out_<name>_bin rows with an output-edges file (heliosk.cu:2266-2281):
    the running mean restarts at each edge e with y_j <= e <= y_{j+1} (point j
    included); the first edge opens the first interval."""
    rows, ie, n, acc = [], 0, 0, 0.0
    N = len(K)
    for j in range(N):
        y0, y1 = j / (N - 1.0), (j + 1) / (N - 1.0)
        n += 1
        acc += K[j]
        if ie < len(edges) and y0 <= edges[ie] <= y1:
            if ie > 0:
                rows.append((0.5 * (edges[ie] + edges[ie - 1]), acc / n))
            ie += 1
            n, acc = 0, 0.0
    return rows


def test_edge_means_hand_case():
    K = np.arange(11, dtype=float) ** 2
    rows = edge_means(np.linspace(0, 1, 11), K, [0.0, 0.25, 0.5, 1.0])
    assert rows == [(0.125, (1 + 4) / 2), (0.375, (9 + 16) / 2),
                    (0.75, (25 + 36 + 49 + 64 + 81) / 5)]


def test_driver_bins_and_per_bin_points_errors(tmp_path):
    (tmp_path / "bins.txt").write_text("1000\n1001.5\n1004\n1010\n")
    for kw, msg in [
        (dict(binsFile="bins.txt", doResampling=1),
         "Error: The resampling function is not supported for the bin-file option"),
        (dict(binsFile="bins.txt", doTransmission=1),
         "Error: The transmission function is not supported for the bin-file option"),
        (dict(Nnu_per_bin=100, doResampling=1),
         "Error: The resampling function is not supported for unequal spacing option"),
        (dict(Nnu_per_bin=100, doTransmission=2),
         "Error: The transmission function is not supported for unequal spacing option"),
        (dict(binsFile="missing.txt"), "Error: bins file not found: missing.txt"),
        (dict(binsFile="bins.txt", dnu=0.001),      # 10000 points, 3 bins
         "Error: range cannot be divided evenly in bins. 10000 3"),
    ]:
        _setup(tmp_path, **kw)
        rc, out = run_driver(str(tmp_path))
        assert msg in out, (kw, out)
    (tmp_path / "bins.txt").write_text("1000\n1004\n1002\n1010\n")
    _setup(tmp_path, binsFile="bins.txt")
    rc, out = run_driver(str(tmp_path))
    assert "Use bins in file bins.txt" in out
    assert "Error; bin boundaries not monotonic growing" in out, out


def test_driver_sublorentz_file_errors(tmp_path):
    _setup(tmp_path, subLorentzianfile="nochi.dat")
    rc, out = run_driver(str(tmp_path))
    assert "Use sub-Lorentzian file nochi.dat" in out
    assert "Error: sub-Lorentzian file not found: nochi.dat" in out, out
    (tmp_path / "chi.dat").write_text("sigma1 3.0\nsigma2 30.0\n")
    _setup(tmp_path, subLorentzianfile="chi.dat")
    rc, out = run_driver(str(tmp_path))
    assert "Error in sub-Lorentzian file read chi.dat" in out, out


def test_driver_undefined_key(tmp_path):
    (tmp_path / "param.dat").write_text("name = a\nT = 300\nnot a key = 3\n")
    rc, out = run_driver(str(tmp_path))
    assert rc == 0 and "Undefined line in param.dat file: line 2" in out, out


def test_driver_console_overrides(tmp_path):
    _setup(tmp_path, numin=0.0, numax=10.0, dnu=0.01, nbins=5)
    rc, out = run_driver(str(tmp_path), "-nbins", "7")
    assert "range cannot be divided evenly in bins. 1000 7" in out, out
    rc, out = run_driver(str(tmp_path), "-bogus", "1")
    assert "Error: Console arguments not valid!" in out, out
    rc, out = run_driver(str(tmp_path), "-maxlines", "0")
    assert "Error: maxlines must lie in 1 .. 2147483647" in out, out
    rc, out = run_driver(str(tmp_path), "-prec", "3")
    assert "Error: precision must be 0, 1 or 2" in out, out
    rc, out = run_driver(str(tmp_path), "-dnu", "0.003")
    assert "range cannot be divided evenly in bins. 3333 5" in out, out


def test_driver_empty_path_values(tmp_path):
    # "pathToData =" and "pathToK =" with no value: the reader takes the next
    # key's value (host.h:128-138, 197-205)
    _setup(tmp_path, numin=2.0, numax=10.0, dnu=0.01, nbins=3)
    assert "pathToData = \nnumin = 2.0" in (tmp_path / "param.dat").read_text()
    rc, out = run_driver(str(tmp_path))
    assert "range cannot be divided evenly in bins. 800 3" in out, out
