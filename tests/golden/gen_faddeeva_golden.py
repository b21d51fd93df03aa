"""Writes faddeeva_mpmath.csv: w(z) = exp(-z^2) erfc(-iz) at 50 digits.

Run in the build container (mpmath present):  python tests/golden/gen_faddeeva_golden.py
The grid covers the core (|z|^2 < 100), the Gauss-Hermite band and the far
wings used by the kernels, with y down to 1e-12 (near-Doppler lines) and
up to 1e5 (high-pressure lines).
"""
import os

import mpmath as mp

mp.mp.dps = 50
XS = ["0", "1e-6", "1e-3", "0.1", "0.5", "1", "1.5", "2", "3", "3.5", "4.5", "5", "6",
      "7", "7.9", "8.1", "9.9", "10.1", "12", "20", "35", "50", "64.1", "100", "300",
      "1000", "3000", "1e4", "1e5", "1e6"]
YS = ["0", "1e-12", "1e-8", "1e-6", "1e-4", "1e-3", "0.01", "0.05", "0.1", "0.3", "0.5",
      "1", "2", "3", "5", "7.9", "8.1", "10", "12", "30", "100", "1e3", "1e5"]


def w(z):
    # exp(-z^2) erfc(-iz); for large |z| use the integral identity via mpmath's
    # own extra precision (workdps) so the product does not cancel.
    with mp.workdps(50 + int(2 * mp.log10(1 + abs(z) ** 2)) + 20):
        return mp.exp(-z * z) * mp.erfc(-1j * z)


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "faddeeva_mpmath.csv")
    rows = []
    for xs in XS:
        for ys in YS:
            x = mp.mpf(xs); y = mp.mpf(ys)
            r2 = x * x + y * y
            if r2 > mp.mpf("1e13"):
                continue
            v = w(mp.mpc(x, y))
            rows.append("%s,%s,%s,%s" % (mp.nstr(x, 20), mp.nstr(y, 20),
                                          mp.nstr(v.real, 25), mp.nstr(v.imag, 25)))
    with open(out, "w") as f:
        f.write("# x,y,re_w,im_w  (mpmath %s, 50 digits)\n" % mp.__version__)
        f.write("\n".join(rows) + "\n")
    print("wrote", len(rows), "rows to", out)


if __name__ == "__main__":
    main()
