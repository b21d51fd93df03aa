"""Shared helpers of the GPU tests: small random line lists and comparisons."""
import numpy as np

NA = 6.0221412927e23


def random_lines(n, seed=0, numin=1000.0, numax=1100.0, delta=0.0, A=0.0, niso=1,
                 gamma=(0.01, 0.1), logS=(-25.0, -19.0), EL=(0.0, 5000.0)):
    rng = np.random.default_rng(seed)
    nu = np.sort(rng.uniform(numin, numax, n))
    iso = rng.integers(0, niso, n).astype(np.int32) if niso > 1 else None
    masses = np.array([18.010565, 20.0148, 19.0148, 17.0])[:niso] / NA
    return dict(nu=nu, S=10 ** rng.uniform(*logS, n), EL=rng.uniform(*EL, n),
                A=np.full(n, float(A)), delta=np.full(n, float(delta)),
                vy=rng.uniform(*gamma, n), n=rng.uniform(0.3, 0.8, n), iso=iso,
                iso_mass=masses, iso_ab=np.array([0.997, 0.002, 3.7e-4, 1e-4])[:niso],
                iso_sscale=np.ones(niso), iso_Q=np.array([174.58, 176.05, 1052.14, 900.0])[:niso])


def species_from(d, form=0):
    from heliosk.engine import Species
    return Species(d["nu"], d["S"], d["EL"], d["vy"], d["n"], A=d["A"], delta=d["delta"],
                   iso=d["iso"], iso_mass=tuple(d["iso_mass"]), iso_ab=tuple(d["iso_ab"]),
                   iso_sscale=tuple(d["iso_sscale"]), form=form)


def state_numpy(st):
    return {k: st.array(k).cpu().numpy() for k in ("nu", "S", "ialphaD", "y", "hw")}


def rel_err(K, Kref, tau_rel=1e-12):
    """max_j |K - K*| / max(|K*|, tau), tau = tau_rel * max|K*| (relative metric
    with a floor that keeps exactly-empty points out)."""
    K = np.asarray(K); Kref = np.asarray(Kref)
    tau = tau_rel * np.max(np.abs(Kref)) if Kref.size else 0.0
    return float(np.max(np.abs(K - Kref) / np.maximum(np.abs(Kref), tau))) if K.size else 0.0
