"""Copies the reference's own example outputs into compact test fixtures.

Source: exoclime/HELIOS-K 2.08, directory plots/ (GPL-3.0; output data of
the reference's documented example runs, written by HELIOS-K 2.0-2.05).
Run in the build container where the reference tree is mounted:

    python tests/golden/gen_reference_fixtures.py [/root/reference]

It writes data only (numbers read from the text outputs, no code):
  p004_K.npy               K column of plots/p004/Out_r.dat (= Out_i.dat), 30000 pts, 1 bin
  p004_sorted.sha256       sha256 of the float64 K column of plots/p004/Out_i_bin.dat
  p004_Out_r_cbin.dat      plots/p004/Out_r_cbin.dat (k-table, kmin = 0, nC = 20)
  p004_Out_r_bin_K.npy     K column of plots/p004/Out_r_bin.dat (reconstruction)
  p005_Out_i_bin.dat       plots/p005/Out_i_bin.dat (sorted K, 500 pts)
  p005_Out_r_cbin.dat      plots/p005/Out_r_cbin.dat (k-table, kmin = 1e-7)
  p005_Out_r_bin.dat       plots/p005/Out_r_bin.dat (reconstruction)
  p009_p{1,3}_K.npz        K columns of plots/p009/Out_p{1,3}.dat (1e5 pts, 5 bins)
  p009_p{1,3}_sorted.sha256  sha256 of the float64 K columns of Out_p*_bin000{0..4}.dat
  p009_p{1,3}_mean.dat     plots/p009/Out_p{1,3}_mean.dat (Planck / Rosseland means)
  p006_bin{0,4}.npz        K column of plots/p006/Out_i_bin000{0,4}.dat (sorted K, 60000 pts)
  p006_tr{0,4}.npy         plots/p006/Out_i_tr000{0,4}.dat (m, transmission), 1000 rows
"""
import hashlib
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def kcol(path):
    return np.loadtxt(path, usecols=(1,), dtype=np.float64)


def main(ref):
    p = os.path.join(ref, "plots")
    np.save(os.path.join(HERE, "p004_K.npy"), kcol(os.path.join(p, "p004", "Out_r.dat")))
    with open(os.path.join(HERE, "p004_sorted.sha256"), "w") as f:
        f.write(hashlib.sha256(kcol(os.path.join(p, "p004", "Out_i_bin.dat")).tobytes()).hexdigest() + "\n")
    np.save(os.path.join(HERE, "p004_Out_r_bin_K.npy"), kcol(os.path.join(p, "p004", "Out_r_bin.dat")))
    shutil.copyfile(os.path.join(p, "p004", "Out_r_cbin.dat"), os.path.join(HERE, "p004_Out_r_cbin.dat"))
    for f in ("Out_i_bin.dat", "Out_r_cbin.dat", "Out_r_bin.dat"):
        shutil.copyfile(os.path.join(p, "p005", f), os.path.join(HERE, "p005_" + f))
    for run in ("p1", "p3"):
        K = kcol(os.path.join(p, "p009", "Out_%s.dat" % run))
        np.savez_compressed(os.path.join(HERE, "p009_%s_K.npz" % run), K=K)
        h = hashlib.sha256()
        for b in range(5):
            h.update(kcol(os.path.join(p, "p009", "Out_%s_bin%04d.dat" % (run, b))).tobytes())
        with open(os.path.join(HERE, "p009_%s_sorted.sha256" % run), "w") as f:
            f.write(h.hexdigest() + "\n")
    for run in ("p1", "p3"):
        shutil.copyfile(os.path.join(p, "p009", "Out_%s_mean.dat" % run),
                        os.path.join(HERE, "p009_%s_mean.dat" % run))
    for b in (0, 4):
        np.savez_compressed(os.path.join(HERE, "p006_bin%d.npz" % b),
                            K=kcol(os.path.join(p, "p006", "Out_i_bin%04d.dat" % b)))
        np.save(os.path.join(HERE, "p006_tr%d.npy" % b),
                np.loadtxt(os.path.join(p, "p006", "Out_i_tr%04d.dat" % b), dtype=np.float64))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
