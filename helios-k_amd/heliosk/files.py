"""Readers and writers of the driver's file formats.

Writers produce inputs in the reference's layouts, so synthetic line lists
run through the real driver path:
  * HITRAN/HITEMP binary line files: 68-byte records, char cid[4] then
    nu, S, EL, A, delta, gamma_air, gamma_self, n as float64 (host.h:605-615);
  * <species>.param files (ISO.h:16-219);
  * partition-function tables, two columns T Q (host.h:9-73);
  * param.dat (host.h:75-300).
Readers parse the driver's outputs: Out_<name>.bin (float32 K, P-major,
heliosk.cu:1889-1911), Out_<name>.dat (heliosk.cu:1839-1861), the k-table
rows of Out_<name>_cbin.dat (heliosk.cu:2166-2181) and the per-bin K of
Out_<name>_bin.dat (heliosk.cu:2245-2283).
"""
import os

import numpy as np

HITRAN_RECORD = np.dtype([("cid", "S4"), ("nu", "<f8"), ("S", "<f8"), ("EL", "<f8"), ("A", "<f8"),
                          ("delta", "<f8"), ("gair", "<f8"), ("gself", "<f8"), ("n", "<f8")])
assert HITRAN_RECORD.itemsize == 68

# param.dat keys in the reference's order (param.dat:1-35)
PARAM_DEFAULTS = [
    ("name", "i"), ("T", 1500.0), ("P", 1.0), ("PFile", "-"), ("Species Name", "species"),
    ("SpeciesFile", "-"), ("ciaSystem", "-"), ("pathToData", ""), ("numin", 0.0), ("numax", 1000.0),
    ("dnu", 0.01), ("Nnu per bin", 0), ("cutMode", 0), ("cut", 25.0), ("doResampling", 0),
    ("nC", 20), ("doTransmission", 0), ("nTr", 1000), ("dTr", 0.05), ("doStoreFullK", 1),
    ("pathToK", ""), ("doStoreSK", 0), ("nbins", 1), ("binsFile", "-"), ("OutputEdgesFile", "-"),
    ("kmin", 0.0), ("qalphaL", 1.0), ("gammaF", 1.0), ("doMean", 0), ("Units", 0),
    ("ReplaceFiles", 1), ("profile", 1), ("subLorentzianfile", "-"), ("removePlinth", 0),
    ("doTuning", 1),
]


def cid_bytes(cid):
    """Isotopologue id as stored in a record: "%3s" plus the terminating NUL."""
    return ("%3s" % cid).encode()[:3] + b"\0"


def write_hitran_bin(path, cid, nu, S, EL, A, delta, gair, gself, n):
    nl = len(nu)
    rec = np.zeros(nl, dtype=HITRAN_RECORD)
    rec["cid"] = [cid_bytes(c) for c in (cid if np.ndim(cid) else [cid] * nl)]
    for k, v in (("nu", nu), ("S", S), ("EL", EL), ("A", A), ("delta", delta), ("gair", gair),
                 ("gself", gself), ("n", n)):
        rec[k] = v
    rec.tofile(path)


def read_hitran_bin(path):
    return np.fromfile(path, dtype=HITRAN_RECORD)


def write_partition(path, T, Q):
    with open(path, "w") as f:
        for t, q in zip(T, Q):
            f.write("%12.4f %20.10f\n" % (t, q))


def write_species_param(path, name, isotopologues, nlines, limits, database=0, molecule=1,
                        npfcol=2, default_gamma=0.0, default_n=0.0):
    """isotopologues: [(cid, abundance, Q296, g, molar mass, partition file)];
    nlines: lines per file; limits: nfiles+1 file limits (cm^-1);
    default_gamma / default_n: the ExoMol-family line widths (ISO.h:188-204)."""
    with open(path, "w") as f:
        f.write("Database = %d\n" % database)
        f.write("Molecule number = %d\n" % molecule)
        f.write("Name = %s\n" % name)
        f.write("Number of Isotopologues = %d\n" % len(isotopologues))
        f.write("#ID Abundance      Q(296K)   g     Molar Mass(g)  partition file :\n")
        for cid, ab, q, g, m, qf in isotopologues:
            f.write(" %s %.10g %.10g %d %.12g %s\n" % (cid, ab, q, g, m, qf))
        f.write("Number of columns in partition File = %d\n" % npfcol)
        f.write("Number of line/transition files = %d\n" % len(nlines))
        f.write("Number of lines per file :\n")
        for k in nlines:
            f.write("%d\n" % k)
        f.write("Line file limits :\n")
        for v in limits:
            f.write("%d\n" % v)
        f.write("#ExoMol :\n")
        f.write("Number of states = 0\n")
        f.write("Number of columns in transition files = 0\n")
        f.write("Default value of Lorentzian half-width for all lines = %.12g\n" % default_gamma)
        f.write("Default value of temperature exponent for all lines = %.12g\n" % default_n)
        f.write("Version = 0\n")


# doubles per line of the ExoMol-family binary files (heliosk.cu:853-873)
EXOMOL_DOUBLES = {2: 4, 20: 2, 30: 5, 31: 5, 32: 5}


def write_exomol_bin(path, database, nu, S, EL=None, A=None, gamma_nat=None):
    """Binary line file of an ExoMol-family database: per line nu, S
    [, E'', A [, Gamma_nat]] as little-endian doubles (the layout
    prepareExomol.cpp writes and host.h:656-667 reads)."""
    nd = EXOMOL_DOUBLES[database]
    cols = [nu, S, EL, A, gamma_nat][:nd]
    rec = np.empty((len(nu), nd), dtype="<f8")
    for k, c in enumerate(cols):
        rec[:, k] = c
    rec.tofile(path)


def species_param_name(base, database, nfiles, lo, hi):
    """Line-file stem of file [lo, hi) of a species (ISO.h:125-175)."""
    if nfiles > 1 and database in (0, 1, 2, 3):
        sep = "__" if database in (2, 3) else "_"
        dash = "_" if database == 3 else "-"
        return "%s%s%05d%s%05d." % (base, sep, lo, dash, hi)
    return base + "."


def write_cia(path, sets, formula="H2-He"):
    """HITRAN collision-induced absorption file: per set a 100-character
    header (formula 20, numin 10, numax 10, point count 7, T 7, max 10 and
    36 further characters) and "nu cia" rows (the layout readCiaFile,
    host.h:670-789, reads).  sets: [(T, nu array, cia array)]."""
    with open(path, "w") as f:
        for T, nu, cia in sets:
            head = "%-20s%10.4f%10.4f%7d%7.1f%10.3E" % (formula, nu[0], nu[-1], len(nu), T,
                                                       max(cia))
            f.write(head.ljust(100)[:100] + "\n")
            for a, b in zip(nu, cia):
                f.write(" %10.4f %.12E\n" % (a, b))


def write_param_dat(path, **overrides):
    """param.dat with the reference's keys; overrides use the key text with
    spaces replaced by underscores (e.g. Species_Name="01_hit16")."""
    vals = dict(PARAM_DEFAULTS)
    for k, v in overrides.items():
        key = k.replace("_", " ") if k.replace("_", " ") in vals else k
        if key not in vals and key != "precision":
            raise KeyError(k)
        vals[key] = v
    with open(path, "w") as f:
        for k, _ in PARAM_DEFAULTS:
            f.write("%s = %s\n" % (k, vals[k]))
        if "precision" in vals:
            f.write("precision = %s\n" % vals["precision"])


def read_out_bin(path, nx):
    K = np.fromfile(path, dtype="<f4")
    return K.reshape(-1, nx)


def read_out_dat(path):
    """Blocks of "x K [T P]" rows separated by blank lines -> list of arrays."""
    blocks, cur = [], []
    with open(path) as f:
        for line in f:
            s = line.split()
            if s:
                cur.append([float(v) for v in s])
            elif cur:
                blocks.append(np.array(cur))
                cur = []
    if cur:
        blocks.append(np.array(cur))
    return blocks


def read_cbin(path, nC):
    """k-table rows: kmin, ystart, nC coefficients [, T, P]."""
    rows = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if s:
                rows.append([float(v) for v in s])
    out = np.array(rows)
    return {"kmin": out[:, 0], "ystart": out[:, 1], "coeff": out[:, 2:2 + nC],
            "TP": out[:, 2 + nC:] if out.shape[1] > 2 + nC else None}


def species_files(dirpath, name, lines, isotopologues, T_table, Q_tables, limits=None):
    """Write <name>.param, the partition tables and <name>.bin for one
    species.  lines: dict of arrays nu, S, EL, A, delta, gair, gself, n and
    per-line cid; isotopologues: [(cid, abundance, Q296, g, molar mass)]."""
    isos = []
    for k, (cid, ab, q, g, m) in enumerate(isotopologues):
        qf = "q_%s_%d.txt" % (name, k)
        write_partition(os.path.join(dirpath, qf), T_table, Q_tables[k])
        isos.append((cid, ab, q, g, m, qf))
    nl = len(lines["nu"])
    lim = limits if limits is not None else [0, int(np.ceil(np.max(lines["nu"]))) + 1 if nl else 1]
    nfiles = len(lim) - 1
    # one file per [lim[k], lim[k+1]) range, lines in their given order
    fidx = np.clip(np.searchsorted(np.asarray(lim[1:-1], dtype=float), lines["nu"], side="right"),
                   0, nfiles - 1) if nfiles > 1 else np.zeros(nl, dtype=int)
    counts = []
    cid = np.asarray(lines["cid"]) if np.ndim(lines["cid"]) else np.array([lines["cid"]] * nl)
    for k in range(nfiles):
        sel = np.nonzero(fidx == k)[0]
        counts.append(sel.size)
        stem = species_param_name(name, 0, nfiles, lim[k], lim[k + 1])
        write_hitran_bin(os.path.join(dirpath, stem + "bin"), list(cid[sel]),
                         *[np.asarray(lines[f])[sel] for f in ("nu", "S", "EL", "A", "delta", "gair",
                                                                "gself", "n")])
    write_species_param(os.path.join(dirpath, name + ".param"), name, isos, counts, lim)
