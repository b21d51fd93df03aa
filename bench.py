#!/usr/bin/env python3
"""Benchmark of the line-by-line path on one MI355X.

Workload (default): 1e7 synthetic lines, grid 0-30000 cm^-1 at 0.001 (3e7
points), T = 2000 K, P = 10 bar, cut 25 cm^-1, 300 bins x 20 Chebyshev
coefficients -- the precision-at-scale configuration.  One step = line
parameters + FP32-compensated Voigt sum + k-table over the whole grid, with
the line list already resident in HBM.  value = algorithmic Voigt
evaluations (line-point pairs inside the cut windows, counted exactly by the
library) per second of step time -- the pairs the reference evaluates one
by one.  The FP32C form sums the lines far from a tile through an 11-node
Chebyshev far field; "voigt_work" gives the pairs evaluated directly and
those delivered through the far field, and the roofline counts only the
operations actually done.

Also reported: the plain FP64 form's evaluations/s on the same input, the
k-table points/s, FP32C-vs-FP64 max relative error on the full grid, the
Voigt kernel's roofline fraction (fixed per-region operation model, see
FLOPS_PER_EVAL) and a CPU baseline (the FP64 oracle on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "helios-k_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

# Operations per evaluation of the FP64 plain method, counted on its formulas
# with divisions, exponentials and trigonometric calls weighted 4, for the
# reference's regions (A: r2 >= 1e6 one-term Lorentz wing; B: 100 <= r2 < 1e6
# rational; C: r2 < 100 Algorithm 916 with ~9 series terms).
FLOPS_PER_EVAL = {"A": 10.0, "B": 22.0, "C": 400.0}
# Far field of the FP32C form: one FP64 node evaluation (4-term complex
# asymptotic series with a refined reciprocal, ~40 FP64 operations, counted
# twice at the FP32 rate) per node, 9 nodes per (line, tile) in band 1, 15
# in band 2 and 19 in band 3; Clenshaw sums per point.
FAR_NODES = 9
FAR_NODES2 = 15
FAR_NODES3 = 19
FLOPS_PER_NODE = 80.0
FLOPS_PER_POINT_FAR = 25.0
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32
PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64
PEAK_HBM_GBS = 8000.0
# The two launches of one FP32C Voigt sum at the default 1024-point tiles.
VOIGT_KERNELS = ("void hk::voigt_tiles_kernel<1, 4>(hk::VArgs)",
                 "void hk::voigt_core_kernel<1>(hk::VArgs)")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hbm_traffic(kernels, workload):
    """HBM bytes of one launch of each of `kernels`, summed, from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE runs of this
    benchmark, each scaled by the factor the same counter reads for a 1 GiB
    reduction / fill), or None when no summary for this workload exists."""
    path = os.path.join(ROOT, "profiles", "hbm_counters_%s.md" % workload)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    a = text.find("```json")
    if a < 0:
        return None
    b = text.find("```", a + 7)
    d = json.loads(text[a + 7:b])
    total = 0.0
    for name in kernels:
        k = d["kernels"].get(name)
        if k is None:
            return None
        total += (k["fetch_kib"] * d["fetch_factor"] + k["write_kib"] * d["write_factor"]) * 1024.0
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device copy of this many MiB first (HBM counter calibration)")
    ap.add_argument("--cpu-lines", type=int, default=100000,
                    help="lines of the CPU-baseline sample (first lines by generator index; "
                         "the default is about 10 s on 16 threads at the default workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (the path would shard "
              "by wavenumber range or by (T,P) spectrum, with no exchange between GPUs)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from heliosk import configs
    from heliosk import engine as E

    w = configs.WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("bench: workload %s (%s)" % (args.workload, w.note))
    if len(w.lines) > 1 or len(w.tp) > 1:
        return bench_batched(args, w)
    if args.workload == "c3":
        return bench_adjoint(args, w)
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        total = float(src.sum())          # a reduction that reads the buffer once
        torch.cuda.synchronize()
        log("bench: copy probe of %d MiB done (%g)" % (args.copy_probe_mb, total))
        del src, dst
    t0 = time.time()
    arr = configs.species_arrays(w, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    T, Pbar = w.tp[0]
    P = Pbar * configs.POBAR
    grid = E.make_grid(w.numin, w.numax, w.dnu)
    kt = E.KTable(grid, w.numax, w.nbins, w.nC)
    st = E.State(sp.nlines)
    K = torch.empty(grid.nx, dtype=torch.float64, device=dev)
    nbytes = E.lib().hk_voigt_workspace_bytes(E.ctypes.byref(grid), E.ctypes.byref(E._tune(0)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    log("bench: setup %.1f s (%d lines, %d points)" % (time.time() - t0, sp.nlines, grid.nx))

    E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut, state=st)
    counts = E.pair_count(st, grid)
    evals = counts[0]
    region = {"A": counts[1], "B": counts[2], "C": counts[3]}
    work = E.work_count(st, grid, E.HK_VOIGT_FP32C)
    direct = {"A": work[0], "B": work[1], "C": work[2]}
    flops = (sum(direct[r] * FLOPS_PER_EVAL[r] for r in direct)
             + (work[4] * FAR_NODES + work[5] * FAR_NODES2 + work[6] * FAR_NODES3)
             * FLOPS_PER_NODE
             + grid.nx * FLOPS_PER_POINT_FAR)
    log("bench: %.4e evaluations per step, regions A/B/C = %s" % (evals, region))
    log("bench: FP32C work: direct pairs A/B/C = %s, far-field pairs %.4e, far (line, tile) "
        "%.4e (9 nodes) + %.4e (15) + %.4e (19)" % (direct, work[3], work[4], work[5], work[6]))

    def step(mode, ev=None):
        E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut, state=st)
        if ev is not None:
            ev[0].record()
        E.voigt(st, grid, 0.0, mode, K=K, workspace=ws)
        if ev is not None:
            ev[1].record()
        return kt(K, kmin=0.0)

    mode = E.HK_VOIGT_FP32C
    for i in range(args.warmup):
        step(mode)
        torch.cuda.synchronize()
        log("bench: warmup step %d done" % i)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        coeff, nxmin = step(mode, evs[i])
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t_start) / args.steps
    voigt_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    log("bench: FP32C step %.3f ms, Voigt kernel %.3f ms" % (t_step * 1e3, voigt_ms))
    K32 = K.clone()

    # k-table alone (sort + fit) on the resident K
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        kt(K32, kmin=0.0)
    e1.record()
    torch.cuda.synchronize()
    ktable_ms = e0.elapsed_time(e1) / 3

    extra = other_paths(E, torch, arr, sp, st, grid, K32, kt, w, T, evals)

    # plain FP64 form on the same input
    fp64_ms = None
    if args.fp64_steps > 0:
        E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, K=K, workspace=ws)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.fp64_steps):
            E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, K=K, workspace=ws)
        e1.record()
        torch.cuda.synchronize()
        fp64_ms = e0.elapsed_time(e1) / args.fp64_steps
        log("bench: FP64 Voigt kernel %.3f ms" % fp64_ms)
        Kr = K
        tau = 1e-12 * float(Kr.max())
        rel = float(torch.max(torch.abs(K32 - Kr) / torch.clamp(torch.abs(Kr), min=tau)))
    else:
        rel = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, arr, T, P, grid, args.cpu_lines)

    achieved = flops / (voigt_ms * 1e-3) / 1e12
    out = {
        "metric": "Voigt evals/s + k-table pts/s, batched FP32-comp vs plain FP64 at <=1e-5 rel err",
        "value": evals / t_step,
        "unit": "Voigt evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (compensated) profile, f64 line params and k-fit",
        "data": "synthetic (seeded generator, %d lines)" % sp.nlines,
        "config": {"workload": "%s: %s" % (w.name, w.note), "lines": sp.nlines,
                   "grid_points": grid.nx, "T_K": T, "P_bar": Pbar, "cut_cm-1": w.cut,
                   "bins": w.nbins, "nC": w.nC, "parallelism": "single GPU"},
        "voigt_evals_per_step": evals,
        "voigt_regions_ABC": [region["A"], region["B"], region["C"]],
        "voigt_work": {"direct_pairs_ABC": [direct["A"], direct["B"], direct["C"]],
                       "far_field_pairs": work[3],
                       "far_line_tiles": work[4] + work[5] + work[6],
                       "far_node_evaluations": (work[4] * FAR_NODES + work[5] * FAR_NODES2
                                                + work[6] * FAR_NODES3)},
        "voigt_kernel_ms": voigt_ms,
        "voigt_evals_per_s_kernel": evals / (voigt_ms * 1e-3),
        "fp64_voigt_kernel_ms": fp64_ms,
        "fp64_voigt_evals_per_s": evals / (fp64_ms * 1e-3) if fp64_ms else None,
        "speedup_fp32c_vs_fp64": fp64_ms / voigt_ms if fp64_ms else None,
        "fp32c_vs_fp64_max_rel_err": rel,
        "precision_ok": (rel is not None and rel <= 1e-5),
        "ktable_ms": ktable_ms,
        "ktable_points_per_s": grid.nx / (ktable_ms * 1e-3),
        "ktable_hbm_GBs_algorithmic": 8.0 * grid.nx / (ktable_ms * 1e-3) / 1e9,
        "other_paths": extra,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": hbm_traffic(VOIGT_KERNELS, args.workload),
                     "kernel": "one Voigt sum = main pass + line-core pass (%s)"
                               % " + ".join(VOIGT_KERNELS),
                     "model": "ops/direct pair A=%g B=%g C=%g (FP64 formulas, div/exp/trig x4) "
                              "+ %g per far-field node + %g per point"
                              % (FLOPS_PER_EVAL["A"], FLOPS_PER_EVAL["B"], FLOPS_PER_EVAL["C"],
                                 FLOPS_PER_NODE, FLOPS_PER_POINT_FAR)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def other_paths(E, torch, arr, sp, st, grid, K, kt, w, T, evals):
    """The paths around the Voigt sum on the same workload, timed with
    events (3 runs each): device line sort of the raw list, Planck/Rosseland
    sums, transmission of the sorted bins, the Lorentz profile gather, and
    the Voigt sum in the reference's own FP32 arithmetic (compat form)."""
    out = {}

    def timed(fn, n=3):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    dev = K.device
    raw = {k: torch.from_numpy(arr[k]).to(dev) for k in ("nu", "S", "EL", "vy", "n")}
    n = sp.nlines
    ms = timed(lambda: E.sort_lines(raw["nu"], raw["S"], raw["EL"], raw["vy"], raw["n"],
                                    iso_mass=arr["iso_mass"]))
    out["line_sort"] = {"ms": ms, "lines_per_s": n / (ms * 1e-3),
                        "note": "raw list (unsorted, 7 FP64 fields) -> stable radix sort by nu "
                                "+ gather; includes output allocation"}
    ms = timed(lambda: E.means(K, grid, T), n=5)
    out["means"] = {"ms": ms, "points_per_s": grid.nx / (ms * 1e-3),
                    "hbm_GBs_algorithmic": 8.0 * grid.nx / (ms * 1e-3) / 1e9}
    _, _, Ks = kt(K, kmin=0.0, want_sorted=True)
    nTr, Nxb = 1000, grid.nx // w.nbins
    ms = timed(lambda: E.transmission(Ks, w.nbins, Nxb, nTr, 0.05))
    out["transmission"] = {"ms": ms, "nTr": nTr,
                           "exp_per_s": w.nbins * Nxb * nTr / (ms * 1e-3)}
    Kl = torch.empty_like(K)
    ms = timed(lambda: E.profile(st, grid, 2, K=Kl), n=1)
    out["lorentz_profile"] = {"ms": ms, "pairs_per_s": evals / (ms * 1e-3),
                              "note": "FP64 gather, same cut windows as the Voigt sum"}
    Kc = torch.empty_like(K)
    ms = timed(lambda: E.voigt(st, grid, 0.0, E.HK_VOIGT_COMPAT, K=Kc), n=1)
    out["compat_voigt"] = {"ms": ms, "pairs_per_s": evals / (ms * 1e-3),
                           "note": "the reference's FP32 region formulas pair by pair, FP64 "
                                   "sums (HK_VOIGT_COMPAT)"}
    log("bench: other paths %s" % json.dumps({k: round(v["ms"], 3) for k, v in out.items()}))
    return out


def bench_batched(args, w):
    """Many species x many (T, P) in one graph-captured launch per pass
    (autotuned tile size), FP32C; then the k-tables of every spectrum."""
    import numpy as np
    import torch

    from heliosk import configs
    from heliosk import engine as E

    dev = torch.device("cuda", 0)
    t0 = time.time()
    grid = E.make_grid(w.numin, w.numax, w.dnu)
    species = []
    for s_ in range(len(w.lines)):
        arr = configs.species_arrays(w, s_)
        species.append(E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"],
                                 iso_mass=arr["iso_mass"]))
    items, states = [], []
    nspec = len(species) * len(w.tp)
    Kall = torch.empty(nspec, grid.nx, dtype=torch.float64, device=dev)
    evals = 0
    k = 0
    for sp in species:
        for T, Pbar in w.tp:
            st = E.line_params(sp, T, Pbar * configs.POBAR, Q=[1.0], cut_mode=0, cut=w.cut)
            evals += E.pair_count(st, grid)[0]
            states.append(st)
            items.append((st, grid, 0.0, Kall[k]))
            k += 1
    torch.cuda.synchronize()
    log("bench: setup %.1f s (%d spectra, %.4e evaluations)" % (time.time() - t0, nspec, evals))
    # autotune on the largest species at the middle (T, P)
    mid = (len(species) - 1) * len(w.tp) + len(w.tp) // 2
    tile, tms = E.autotune(states[mid], grid, E.HK_VOIGT_FP32C)
    log("bench: autotuned tile %d points (%s ms)" % (tile, {k_: round(v, 3) for k_, v in tms.items()}))
    res = {}
    for mode, name in ((E.HK_VOIGT_FP32C, "fp32c"), (E.HK_VOIGT_FP64, "fp64")):
        if mode == E.HK_VOIGT_FP64 and args.fp64_steps <= 0:
            continue
        batch = E.VoigtBatch(items, mode=mode, tile_points=tile)
        batch.launch()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            batch.launch()
        steps = args.steps if mode == E.HK_VOIGT_FP32C else args.fp64_steps
        for _ in range(args.warmup if mode == E.HK_VOIGT_FP32C else 0):
            g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / steps
        log("bench: %s batched Voigt %.3f ms per replay" % (name, res[name]))
        if name == "fp32c":
            K32 = Kall.clone()
        del g, batch
    rel = None
    if "fp64" in res:
        tau = 1e-12 * torch.amax(Kall, dim=1, keepdim=True)
        rel = float(torch.max(torch.abs(K32 - Kall) / torch.maximum(torch.abs(Kall), tau)))
    kt = E.KTable(grid, w.numax, w.nbins, w.nC, nspec=nspec)
    kt(K32, kmin=0.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    kt(K32, kmin=0.0)
    e1.record()
    torch.cuda.synchronize()
    ktable_ms = e0.elapsed_time(e1)
    t_step = (res["fp32c"] + ktable_ms) * 1e-3
    out = {
        "metric": "Voigt evals/s + k-table pts/s, batched FP32-comp vs plain FP64 at <=1e-5 rel err",
        "value": evals / t_step, "unit": "Voigt evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (compensated) profile, f64 line params and k-fit",
        "data": "synthetic (seeded generator, %d species, %d lines)" % (len(species), sum(w.lines)),
        "config": {"workload": "%s: %s" % (w.name, w.note), "spectra": nspec,
                   "grid_points": grid.nx, "cut_cm-1": w.cut, "bins": w.nbins, "nC": w.nC,
                   "tile_points": tile, "parallelism": "single GPU"},
        "voigt_evals_per_step": evals,
        "voigt_batched_ms": res["fp32c"],
        "voigt_evals_per_s_kernel": evals / (res["fp32c"] * 1e-3),
        "fp64_voigt_batched_ms": res.get("fp64"),
        "speedup_fp32c_vs_fp64": res["fp64"] / res["fp32c"] if "fp64" in res else None,
        "fp32c_vs_fp64_max_rel_err": rel,
        "precision_ok": (rel is not None and rel <= 1e-5),
        "ktable_ms": ktable_ms,
        "ktable_points_per_s": nspec * grid.nx / (ktable_ms * 1e-3),
        "autotune_ms": tms,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args, w):
    """Differentiation workload: FP64 forward + hand-written adjoint of a
    binned log-mean loss, through the torch.autograd.Function."""
    import numpy as np
    import torch

    from heliosk import configs
    from heliosk import engine as E

    dev = torch.device("cuda", 0)
    arr = configs.species_arrays(w, 0)
    sp = E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"], iso_mass=arr["iso_mass"])
    T, Pbar = w.tp[0]
    P = Pbar * configs.POBAR
    grid = E.make_grid(w.numin, w.numax, w.dnu)
    wts = torch.from_numpy(np.random.default_rng(w.seed).uniform(0, 1, w.nbins)).to(dev)
    st = E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut)
    evals = E.pair_count(st, grid)[0]

    def step():
        Tt = torch.tensor(T, dtype=torch.float64, requires_grad=True)
        Pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
        S = sp.S.clone().requires_grad_()
        vy = sp.vy.clone().requires_grad_()
        K = E.opacity(Tt, Pt, S, vy, sp, grid, Q=[1.0], cut_mode=0, cut=w.cut)
        L = E.binned_log_mean(K, w.nbins, wts)
        L.backward()
        return L

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    out = {
        "metric": "Voigt evals/s (forward + adjoint) of a binned-opacity loss",
        "value": 2 * evals / t_step, "unit": "Voigt evaluations/s (forward and adjoint counted)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, %d lines)" % sp.nlines,
        "config": {"workload": "%s: %s" % (w.name, w.note), "bins": w.nbins,
                   "parallelism": "single GPU"},
        "voigt_evals_per_step": evals, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(w, arr, T, P, grid, nlines):
    """The FP64 oracle (line parameters + Voigt sum over the full grid) on the
    first `nlines` lines by generator index, timed on this host's cores."""
    import numpy as np
    from heliosk import configs
    from oracle import oracle as O
    ncores = min(16, os.cpu_count() or 1)
    sub = {k: (v[:nlines] if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
    ol = configs.oracle_lines(sub)
    t0 = time.perf_counter()
    ost = O.line_params(ol, T, P, 0, w.cut)
    K = O.voigt_points(ost, grid.numin, grid.dnu, grid.nx, nthreads=ncores)
    dt = time.perf_counter() - t0
    # the same sample in the reference's FP32-evaluation / FP64-accumulation
    # arithmetic (the oracle's compat restatement)
    t1 = time.perf_counter()
    O.voigt_compat_points(ost, grid.numin, grid.dnu, grid.nx, nthreads=ncores)
    dtc = time.perf_counter() - t1
    nu = ost["nu"]
    # evaluations of the sample: window points of each line
    j = (nu[:, None] + np.array([-w.cut, w.cut])[None, :] - grid.numin) / grid.dnu
    lo = np.clip(np.floor(j[:, 0]) + 1, 0, grid.nx); hi = np.clip(np.floor(j[:, 1]) + 1, 0, grid.nx)
    ev = float(np.sum(np.maximum(hi - lo, 0)))
    log("bench: CPU oracle %d lines, %.3e evals in %.1f s (compat arithmetic %.1f s) on %d "
        "threads" % (nlines, ev, dt, dtc, ncores))
    return {"value": ev / dt, "unit": "Voigt evaluations/s", "cores": ncores, "kind": "port",
            "sample": "FP64 oracle (Algorithm 916 + continued fraction, compensated sum) over the "
                      "full %d-point grid for the first %d lines of the workload" % (grid.nx, nlines),
            "seconds": dt, "checksum": float(np.sum(K)),
            "compat_value": ev / dtc, "compat_seconds": dtc,
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
