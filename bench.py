#!/usr/bin/env python3
"""Benchmark of the line-by-line path on one MI355X.

Headline workload (default, --workload c2): the batched uneven workload --
6 synthetic species of 1e4, 1e5, 5e5, 1e6, 3e6 and 5e6 lines (masses 2, 16,
18, 28, 44, 48 amu) x 64 (T, P) points (T log-spaced 100-3000 K, P
log-spaced 1e-6-1e3 bar, paired), grid 0-30000 cm^-1 at 0.01 (3e6 points),
cut 100 cm^-1, k-tables of 300 bins x 20 Chebyshev coefficients.  One step
= line parameters of all 384 (species, T, P) states + the batched
FP32-compensated Voigt sum (one persistent launch per pass over the tiles
of every spectrum) + the k-tables of all 384 spectra, each stage a captured
hipGraph, replayed with the line lists resident in HBM.

value = Voigt evaluations per second of step time.  One evaluation is a
(line, grid point) pair inside the line's cut window, counted exactly by the
library (hk_voigt_pair_count): the pairs the reference evaluates one by
one.  The FP32C form delivers most pairs through its Chebyshev far field
("voigt_work" gives the split), so value is a pair-equivalent rate;
"speedup_fp32c_exact_vs_fp64" gives the FP32C/FP64 ratio with every pair
evaluated directly (hk_tune.exact_pairs = 1), beside the far-field ratio.

Roofline of the Voigt sum (the persistent main and line-core kernels): the
VALU floating-point operations per replay that the hardware counters
SQ_INSTS_VALU_FLOPS_FP32/FP64 (+ _TRANS) record (committed summary
profiles/valu_counters_<workload>.md) divided by the replay time measured
here with HIP events, against the FP32 vector peak.  The k-table stage's HBM
rate is reported from its algorithmic bytes (8 B per point read) and from
the FETCH_SIZE / WRITE_SIZE summary profiles/hbm_counters_<workload>.md.

CPU baseline (rank 0, N = 1): the FP64 oracle (this repository's C
restatement of the path) compiled with -O3 -march=native -fopenmp on the
host that runs the benchmark, on the host threads the run is given
(OMP_NUM_THREADS): the Voigt sum of the first 1 % of each species' lines
over the full grid at the middle (T, P), and the k-tables (per-bin sort +
fit) of the first 1 % of bins.

--workload c1 / c0 runs the single-spectrum form (1e7 lines, 3e7 points at
0.001 cm^-1, 2000 K, 10 bar, cut 25 for c1); the default run also appends a
compact c1 block ("precision_at_scale").  --workload c4 runs the batched
form on the mixed-width stress workload, c3 the forward + adjoint step.
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

METRIC = "Voigt evals/s + k-table pts/s, batched FP32-comp vs plain FP64 at <=1e-5 rel err"
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (counts packed FP32)
PEAK_FP32_ISSUE_TFLOPS = 78.6  # without packed FP32 (this build): 256 CU x 4 SIMD x 16 lanes x 2 x 2.4 GHz
PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64
PEAK_HBM_GBS = 8000.0
CLOCK_GHZ = 2.4
N_CU = 256
# Kernels are named by the prefix of their demangled name as rocprofv3 prints it.
# One FP32C Voigt sum of the single-spectrum form at the default 1024-point tiles.
VOIGT_KERNELS_SINGLE = ("void hk::voigt_tiles_kernel<1, 4, 1>(", "void hk::voigt_core_kernel<1>(")
# The batched form's two persistent kernels: main pass at tile 512..4096 points
# (256 x points per thread) with 256 or 512 lines per LDS stage (256 x stages), and
# the line-core pass.
PERSISTENT_MAIN = "void hk::voigt_persistent_kernel<1, %d, %d>("
PERSISTENT_CORE = "void hk::voigt_persistent_kernel<1, 0, 1>("
# The k-table stage: bins of at most 16384 points are sorted and fitted by one
# kernel; longer bins by LDS-sorted runs, splitters, buckets and a combine.
KTABLE_KERNELS = ("(anonymous namespace)::build_segs_kernel(",
                  "hk::bin_sort_fit_kernel(", "hk::run_sort_kernel(", "hk::merge_lds_kernel(",
                  "hk::splitter_kernel(", "hk::bucket_kernel(", "hk::bucket_combine_kernel(")
WAVE = 64
FLOP_COUNTERS = ("SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FLOPS_FP64",
                 "SQ_INSTS_VALU_FLOPS_FP32_TRANS", "SQ_INSTS_VALU_FLOPS_FP64_TRANS")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_json(kind, workload):
    """The JSON block of a committed counter summary (profiles/<kind>_<workload>.md),
    or None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", "%s_%s.md" % (kind, workload))
    if not os.path.exists(path):
        return None
    text = open(path).read()
    a = text.find("```json")
    if a < 0:
        return None
    b = text.find("```", a + 7)
    return json.loads(text[a + 7:b])


def _matching(d, prefixes, required):
    """Entries of a counter summary whose kernel name starts with one of
    `prefixes`; None when a required prefix matches nothing."""
    out = []
    for i, p in enumerate(prefixes):
        hit = [v for k, v in d["kernels"].items() if k.startswith(p)]
        if not hit and i < required:
            return None
        out.extend(hit)
    return out


def hbm_traffic(kernels, workload, required=None):
    """HBM bytes of one launch of each kernel named by a prefix in `kernels`,
    summed, from the committed FETCH_SIZE / WRITE_SIZE summary, each counter
    scaled by the factor it reads for a 1 GiB reduction / fill; None without a
    summary for the workload.  The first `required` prefixes (all by default)
    must be present."""
    d = profile_json("hbm_counters", workload)
    if d is None:
        return None
    ks = _matching(d, kernels, len(kernels) if required is None else required)
    if ks is None:
        return None
    return sum((k["fetch_kib"] * d["fetch_factor"] + k["write_kib"] * d["write_factor"]) * 1024.0
               for k in ks)


def counter_flops(kernels, workload):
    """(VALU FLOPs, VALU instructions, SQ_ACTIVE_INST_VALU, SQ_BUSY_CU_CYCLES) per
    launch of `kernels`, summed, from the committed counter summary, or None."""
    d = profile_json("valu_counters", workload)
    if d is None:
        return None
    ks = _matching(d, kernels, len(kernels))
    if ks is None:
        return None
    out = [0.0, 0.0, 0.0, 0.0]
    for k in ks:
        # the FLOPS counters count per wave64 instruction (1 per add or mul, 2
        # per FMA): lane FLOPs are 64 x the count (the 1 GiB FP64 reduction of
        # the copy probe, 1.34e8 adds, reads 2.12e6)
        out[0] += WAVE * sum(k[c] for c in FLOP_COUNTERS)
        out[1] += k["SQ_INSTS_VALU"]
        out[2] += k["SQ_ACTIVE_INST_VALU"]
        out[3] += k["SQ_BUSY_CU_CYCLES"]
    return out


def roofline_valu(kernels, workload, ms, label):
    """Counter-based VALU roofline of one launch of `kernels` taking `ms`."""
    c = counter_flops(kernels, workload)
    if c is None:
        return None
    flops, insts, active, busy = c
    achieved = flops / (ms * 1e-3) / 1e12
    # one wave64 VALU instruction per SIMD per 4 cycles without packed math
    issue_rate = N_CU * 4 * CLOCK_GHZ * 1e9 / 4.0
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "frac_of_unpacked_peak": achieved / PEAK_FP32_ISSUE_TFLOPS,
            "valu_issue_utilisation": insts / (ms * 1e-3 * issue_rate),
            "active_inst_valu_per_busy_cu_cycle": active / busy if busy else None,
            "traffic": hbm_traffic(kernels, workload),
            "kernel": label, "flops_per_launch": flops,
            "source": "64 lanes x (SQ_INSTS_VALU_FLOPS_FP32 + FP64 + FP32_TRANS + FP64_TRANS) "
                      "per launch (profiles/valu_counters_%s.md; the counters count per wave "
                      "instruction, lanes masked off included) / launch time measured here"
                      % workload}


def omp_threads():
    """Host threads of the CPU baseline: the run's CPU share as OMP_NUM_THREADS
    gives it, else every core."""
    v = os.environ.get("OMP_NUM_THREADS")
    try:
        return max(1, int(v)) if v else (os.cpu_count() or 1)
    except ValueError:
        return os.cpu_count() or 1


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other paths / the appended single-spectrum block")
    ap.add_argument("--profile", action="store_true",
                    help="one FP32C step launched eagerly (no graphs), nothing else: the "
                         "command the rocprofv3 counter summaries run")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device copy of this many MiB first (HBM counter calibration)")
    ap.add_argument("--cpu-lines", type=int, default=100000,
                    help="single-spectrum form: lines of the CPU-baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (the path would shard "
              "by wavenumber range or by (T,P) spectrum, with no exchange between GPUs)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from heliosk import configs

    w = configs.WORKLOADS[args.workload]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("bench: workload %s (%s)" % (args.workload, w.note))
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        total = float(src.sum())          # a reduction that reads the buffer once
        torch.cuda.synchronize()
        log("bench: copy probe of %d MiB done (%g)" % (args.copy_probe_mb, total))
        del src, dst
    if args.workload == "c3":
        out = bench_adjoint(args, w)
    elif len(w.lines) > 1 or len(w.tp) > 1:
        out = bench_batched(args, w, args.workload)
        if out is not None and args.workload == "c2" and not args.no_extra and not args.profile:
            sub = argparse.Namespace(**vars(args))
            sub.steps, sub.warmup, sub.no_cpu, sub.no_extra = 3, 1, True, True
            out["precision_at_scale"] = bench_single(sub, configs.WORKLOADS["c1"], "c1")
    else:
        out = bench_single(args, w, args.workload)
    if out is not None:
        print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# Single spectrum (c0, c1)
# ---------------------------------------------------------------------------
def bench_single(args, w, wname):
    """One step = line parameters + FP32C Voigt sum + k-table of one spectrum."""
    import numpy as np
    import torch

    from heliosk import configs
    from heliosk import engine as E

    dev = torch.device("cuda", 0)
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
    log("bench: %s setup %.1f s (%d lines, %d points)" % (wname, time.time() - t0, sp.nlines,
                                                          grid.nx))

    E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut, state=st)
    counts = E.pair_count(st, grid)
    evals = counts[0]
    work = E.work_count(st, grid, E.HK_VOIGT_FP32C)

    def step(mode, ev=None, exact=False):
        E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut, state=st)
        if ev is not None:
            ev[0].record()
        E.voigt(st, grid, 0.0, mode, K=K, workspace=ws, exact_pairs=exact)
        if ev is not None:
            ev[1].record()
        return kt(K, kmin=0.0)

    mode = E.HK_VOIGT_FP32C
    if args.profile:
        step(mode)
        torch.cuda.synchronize()
        log("bench: profile step done")
        return None
    for i in range(args.warmup):
        step(mode)
        torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        coeff32, nxmin32 = step(mode, evs[i])
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t_start) / args.steps
    voigt_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    log("bench: %s FP32C step %.3f ms, Voigt %.3f ms" % (wname, t_step * 1e3, voigt_ms))
    K32 = K.clone()

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    ktable_ms = timed(lambda: kt(K32, kmin=0.0), 3)
    exact_ms = timed(lambda: E.voigt(st, grid, 0.0, mode, K=K, workspace=ws, exact_pairs=True), 1)
    Kx = K.clone()
    extra = None if args.no_extra else other_paths(E, torch, arr, sp, st, grid, K32, kt, w, T,
                                                   evals)
    fp64_ms = rel = rel_exact = kt_cmp = None
    if args.fp64_steps > 0:
        fp64_ms = timed(lambda: E.voigt(st, grid, 0.0, E.HK_VOIGT_FP64, K=K, workspace=ws),
                        args.fp64_steps)
        log("bench: %s FP64 Voigt %.3f ms" % (wname, fp64_ms))
        tau = 1e-12 * float(K.max())
        rel = float(torch.max(torch.abs(K32 - K) / torch.clamp(torch.abs(K), min=tau)))
        rel_exact = float(torch.max(torch.abs(Kx - K) / torch.clamp(torch.abs(K), min=tau)))
        coeff64, nxmin64 = kt(K, kmin=0.0)
        kt_cmp = ktable_agreement(E, torch, coeff32, nxmin32, coeff64, nxmin64, kt.Nxb)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_single(w, arr, T, P, grid, args.cpu_lines)

    out = {
        "metric": METRIC,
        "value": evals / t_step,
        "unit": "Voigt evaluations/s (line-point pairs in the cut windows per second of step; "
                "pair-equivalent, far-field pairs included)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (compensated) profile, f64 line params and k-fit",
        "data": "synthetic (seeded generator, %d lines)" % sp.nlines,
        "config": {"workload": "%s: %s" % (w.name, w.note), "lines": sp.nlines,
                   "grid_points": grid.nx, "T_K": T, "P_bar": Pbar, "cut_cm-1": w.cut,
                   "bins": w.nbins, "nC": w.nC, "parallelism": "single GPU"},
        "voigt_evals_per_step": evals,
        "voigt_regions_ABC": counts[1:4],
        "voigt_work": {"direct_pairs_ABC": work[0:3], "far_field_pairs": work[3],
                       "far_line_tiles_9_15_19_nodes": work[4:7],
                       "far_line_waves_9_15_19_nodes": work[7:10],
                       "cut_edge_line_waves_and_pairs": work[10:12]},
        "voigt_kernel_ms": voigt_ms,
        "voigt_evals_per_s_kernel": evals / (voigt_ms * 1e-3),
        "fp32c_exact_pairs_voigt_ms": exact_ms,
        "fp64_voigt_kernel_ms": fp64_ms,
        "fp64_voigt_evals_per_s": evals / (fp64_ms * 1e-3) if fp64_ms else None,
        "speedup_fp32c_vs_fp64": fp64_ms / voigt_ms if fp64_ms else None,
        "speedup_fp32c_exact_vs_fp64": fp64_ms / exact_ms if fp64_ms else None,
        "fp32c_vs_fp64_max_rel_err": rel,
        "fp32c_exact_vs_fp64_max_rel_err": rel_exact,
        "precision_ok": (rel is not None and rel <= 1e-5),
        "ktable_fp32c_vs_fp64": kt_cmp,
        "ktable_ms": ktable_ms,
        "ktable_points_per_s": grid.nx / (ktable_ms * 1e-3),
        "ktable_hbm_GBs_algorithmic": 8.0 * grid.nx / (ktable_ms * 1e-3) / 1e9,
        "other_paths": extra,
        "roofline": roofline_valu(VOIGT_KERNELS_SINGLE, wname, voigt_ms,
                                  "one FP32C Voigt sum = main pass + line-core pass"),
        "cpu_baseline": cpu,
    }
    return out


def ktable_agreement(E, torch, c32, n32, c64, n64, Nxb):
    """FP32C k-tables against FP64 k-tables of the same spectra: identical
    Nxmin (so ystart), max |delta c|, and the max relative difference of the
    reconstructed K(y) = exp(Chebyshev sum) (hk_ktable_reconstruct)."""
    same = bool(torch.equal(n32, n64))
    # bins with Nxmin = Nxb have no fit (printed as 0.0 by the reference)
    fitted = n64 < Nxb
    dc = float(torch.max(torch.abs(c32 - c64)[fitted])) if bool(fitted.any()) else 0.0
    nspec = c32.shape[0]
    worst = 0.0
    step = max(1, (1 << 27) // (c32.shape[1] * Nxb))      # spectra per chunk (~1 GiB of K(y))
    for a in range(0, nspec, step):
        b = min(nspec, a + step)
        y32 = E.ktable_reconstruct(c32[a:b], n32[a:b], Nxb)
        y64 = E.ktable_reconstruct(c64[a:b], n64[a:b], Nxb)
        y32 = y32.view(b - a, -1, Nxb)[fitted[a:b]]
        y64 = y64.view(b - a, -1, Nxb)[fitted[a:b]]
        if y64.numel():
            worst = max(worst, float(torch.max(torch.abs(y32 - y64) / torch.abs(y64))))
        del y32, y64
    return {"nxmin_identical": same, "max_abs_coeff_diff": dc, "max_rel_Ky_diff": worst,
            "fitted_bins": int(fitted.sum()), "bins": int(fitted.numel()),
            "ok": same and dc <= 1e-4 and worst <= 1e-4}


def other_paths(E, torch, arr, sp, st, grid, K, kt, w, T, evals):
    """The paths around the Voigt sum on the same workload, timed with
    events: device line sort of the raw list, Planck/Rosseland sums,
    transmission of the sorted bins, the Lorentz profile gather, and the
    Voigt sum in the reference's own FP32 arithmetic (compat form)."""
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
                        "note": "raw list (unsorted, 7 FP64 fields) -> stable sort by nu "
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


# ---------------------------------------------------------------------------
# Batched (c2, c4): many species x many (T, P)
# ---------------------------------------------------------------------------
def bench_batched(args, w, wname):
    """One step = line parameters of every (species, T, P) state + the batched
    FP32C Voigt sum + the k-tables of every spectrum, as three captured graphs
    replayed back to back (events between them split the step)."""
    import numpy as np
    import torch

    from heliosk import configs
    from heliosk import engine as E

    dev = torch.device("cuda", 0)
    t0 = time.time()
    grid = E.make_grid(w.numin, w.numax, w.dnu)
    species, arrays = [], []
    for s_ in range(len(w.lines)):
        arr = configs.species_arrays(w, s_)
        arrays.append(arr)
        species.append(E.Species(arr["nu"], arr["S"], arr["EL"], arr["vy"], arr["n"],
                                 iso_mass=arr["iso_mass"]))
    nspec = len(species) * len(w.tp)
    Kall = torch.empty(nspec, grid.nx, dtype=torch.float64, device=dev)
    lp_items, states, items = [], [], []
    for sp in species:
        for T, Pbar in w.tp:
            st = E.State(sp.nlines)
            lp_items.append((sp, T, Pbar * configs.POBAR, st))
            states.append(st)
            items.append((st, grid, 0.0, Kall[len(items)]))

    def run_lp():
        for sp, T, P, st in lp_items:
            E.line_params(sp, T, P, Q=[1.0], cut_mode=0, cut=w.cut, state=st)

    run_lp()
    torch.cuda.synchronize()
    evals, regions = 0, [0, 0, 0]
    work = [0] * 12
    for st in states:
        c = E.pair_count(st, grid)
        evals += c[0]
        regions = [a + b for a, b in zip(regions, c[1:4])]
    log("bench: setup %.1f s (%d spectra, %.4e evaluations)" % (time.time() - t0, nspec, evals))
    # tile width and lines per LDS stage: autotuned on the largest species at
    # the middle (T, P)
    mid = (len(species) - 1) * len(w.tp) + len(w.tp) // 2
    tile, stage, tms = E.autotune(states[mid], grid, E.HK_VOIGT_FP32C)
    log("bench: autotuned tile %d points, %d lines per stage (%s ms)"
        % (tile, stage, {"%dx%d" % k_: round(v, 3) for k_, v in tms.items()}))
    for st in states:
        work = [a + b for a, b in zip(work, E.work_count(st, grid, E.HK_VOIGT_FP32C,
                                                         tile_points=tile))]
    kt = E.KTable(grid, w.numax, w.nbins, w.nC, nspec=nspec)
    b32 = E.VoigtBatch(items, mode=E.HK_VOIGT_FP32C, tile_points=tile, stage_lines=stage)
    torch.cuda.synchronize()

    if args.profile:
        run_lp()
        b32.launch()
        kt(Kall, kmin=0.0)
        torch.cuda.synchronize()
        log("bench: profile step done")
        return None

    def capture(fn):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            res = fn()
        return g, res

    g_lp, _ = capture(run_lp)
    g_v32, _ = capture(b32.launch)
    g_kt, (coeff_g, nxmin_g) = capture(lambda: kt(Kall, kmin=0.0))
    for i in range(args.warmup):
        g_lp.replay(); g_v32.replay(); g_kt.replay()
        torch.cuda.synchronize()
        log("bench: warmup step %d done" % i)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        e = evs[i]
        e[0].record(); g_lp.replay(); e[1].record(); g_v32.replay(); e[2].record()
        g_kt.replay(); e[3].record()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t_start) / args.steps
    lp_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    voigt_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    ktable_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in evs]))
    log("bench: step %.1f ms = line params %.2f + Voigt %.1f + k-tables %.2f ms"
        % (t_step * 1e3, lp_ms, voigt_ms, ktable_ms))
    coeff32, nxmin32 = coeff_g.clone(), nxmin_g.clone()
    K32 = Kall.clone()

    def replay_ms(g, n):
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            g.replay()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    # every pair evaluated directly (no far field), same tile width
    bx = E.VoigtBatch(items, mode=E.HK_VOIGT_FP32C, tile_points=tile, exact_pairs=True,
                      stage_lines=stage)
    g_x, _ = capture(bx.launch)
    exact_ms = replay_ms(g_x, 1)
    log("bench: FP32C every pair %.1f ms per replay" % exact_ms)
    Kx = Kall.clone()
    del g_x, bx
    res = {"fp64": None}
    rel = rel_exact = kt_cmp = None
    if args.fp64_steps > 0:
        b64 = E.VoigtBatch(items, mode=E.HK_VOIGT_FP64, tile_points=tile, stage_lines=stage)
        g_64, _ = capture(b64.launch)
        res["fp64"] = replay_ms(g_64, args.fp64_steps)
        log("bench: FP64 batched Voigt %.1f ms per replay" % res["fp64"])
        del g_64, b64
        tau = 1e-12 * torch.amax(Kall, dim=1, keepdim=True)
        den = torch.maximum(torch.abs(Kall), tau)
        rel = float(torch.max(torch.abs(K32 - Kall) / den))
        rel_exact = float(torch.max(torch.abs(Kx - Kall) / den))
        del Kx, den
        coeff64, nxmin64 = kt(Kall, kmin=0.0)
        kt_cmp = ktable_agreement(E, torch, coeff32, nxmin32, coeff64, nxmin64, kt.Nxb)
        log("bench: FP32C vs FP64 %.3e (every pair %.3e); k-tables %s" % (rel, rel_exact, kt_cmp))

    cpu = None
    if not args.no_cpu:
        i0 = len(w.tp) // 2
        nbins_kt = -(-nspec * w.nbins // 100)                       # >= 1 % of the bins
        nspec_kt = max(1, -(-nbins_kt // w.nbins))
        cpu = cpu_baseline_batched(w, arrays, grid, i0, Kall[:nspec_kt].cpu().numpy(),
                                   coeff64[:nspec_kt].cpu().numpy() if kt_cmp else None)

    main_k = PERSISTENT_MAIN % (tile // 256, stage // 256)
    roof = roofline_valu((main_k, PERSISTENT_CORE), wname, voigt_ms,
                         "batched FP32C Voigt sum = persistent main pass + persistent line-core "
                         "pass (one replay over all %d spectra)" % nspec)
    kt_traffic = hbm_traffic(KTABLE_KERNELS, wname, required=0)
    out = {
        "metric": METRIC,
        "value": evals / t_step,
        "unit": "Voigt evaluations/s (line-point pairs in the cut windows per second of step; "
                "pair-equivalent, far-field pairs included)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (compensated) profile, f64 line params and k-fit",
        "data": "synthetic (seeded generator, %d species, %d lines)" % (len(species),
                                                                         sum(w.lines)),
        "config": {"workload": "%s: %s" % (w.name, w.note), "species_lines": list(w.lines),
                   "tp_points": len(w.tp), "spectra": nspec, "grid_points": grid.nx,
                   "cut_cm-1": w.cut, "bins": w.nbins, "nC": w.nC, "tile_points": tile,
                   "stage_lines": stage, "parallelism": "single GPU"},
        "step_ms_split": {"line_params": lp_ms, "voigt": voigt_ms, "ktables": ktable_ms},
        "voigt_evals_per_step": evals,
        "voigt_regions_ABC": regions,
        "voigt_work": {"direct_pairs_ABC": work[0:3], "far_field_pairs": work[3],
                       "far_line_tiles_9_15_19_nodes": work[4:7],
                       "far_line_waves_9_15_19_nodes": work[7:10],
                       "cut_edge_line_waves_and_pairs": work[10:12]},
        "voigt_batched_ms": voigt_ms,
        "voigt_evals_per_s_kernel": evals / (voigt_ms * 1e-3),
        "fp32c_exact_pairs_batched_ms": exact_ms,
        "fp64_voigt_batched_ms": res["fp64"],
        "fp64_voigt_evals_per_s": evals / (res["fp64"] * 1e-3) if res["fp64"] else None,
        "speedup_fp32c_vs_fp64": res["fp64"] / voigt_ms if res["fp64"] else None,
        "speedup_fp32c_exact_vs_fp64": res["fp64"] / exact_ms if res["fp64"] else None,
        "fp32c_vs_fp64_max_rel_err": rel,
        "fp32c_exact_vs_fp64_max_rel_err": rel_exact,
        "precision_ok": (rel is not None and rel <= 1e-5),
        "ktable_fp32c_vs_fp64": kt_cmp,
        "ktable_ms": ktable_ms,
        "ktable_points_per_s": nspec * grid.nx / (ktable_ms * 1e-3),
        "ktable_hbm_GBs_algorithmic": 8.0 * nspec * grid.nx / (ktable_ms * 1e-3) / 1e9,
        "ktable_hbm_GBs_counted": kt_traffic / (ktable_ms * 1e-3) / 1e9 if kt_traffic else None,
        "autotune_ms": {"%dx%d" % k_: v for k_, v in tms.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    return out


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
    return {
        "metric": "Voigt evals/s (forward + adjoint) of a binned-opacity loss",
        "value": 2 * evals / t_step, "unit": "Voigt evaluations/s (forward and adjoint counted)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, %d lines)" % sp.nlines,
        "config": {"workload": "%s: %s" % (w.name, w.note), "bins": w.nbins,
                   "parallelism": "single GPU"},
        "voigt_evals_per_step": evals, "roofline": None, "cpu_baseline": None,
    }


# ---------------------------------------------------------------------------
# CPU baseline: the oracle built for this host (-O3 -march=native -fopenmp)
# ---------------------------------------------------------------------------
def _window_evals(nu, cut, grid):
    import numpy as np
    j = (nu[:, None] + np.array([-cut, cut])[None, :] - grid.numin) / grid.dnu
    lo = np.clip(np.floor(j[:, 0]) + 1, 0, grid.nx)
    hi = np.clip(np.floor(j[:, 1]) + 1, 0, grid.nx)
    return float(np.sum(np.maximum(hi - lo, 0)))


def cpu_baseline_single(w, arr, T, P, grid, nlines):
    """The oracle's line parameters + FP64 Voigt sum over the full grid for the
    first `nlines` lines by generator index, plus the same sample in the
    reference's FP32-evaluation / FP64-accumulation arithmetic."""
    import numpy as np
    from heliosk import configs
    from oracle import oracle as O
    L, flags = O.lib_native()
    nth = omp_threads()
    sub = {k: (v[:nlines] if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
    ol = configs.oracle_lines(sub)
    t0 = time.perf_counter()
    ost = O.line_params(ol, T, P, 0, w.cut, L=L)
    K = O.voigt_points(ost, grid.numin, grid.dnu, grid.nx, nthreads=nth, L=L)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    O.voigt_compat_points(ost, grid.numin, grid.dnu, grid.nx, nthreads=nth, L=L)
    dtc = time.perf_counter() - t1
    ev = _window_evals(ost["nu"], w.cut, grid)
    log("bench: CPU oracle %d lines, %.3e evals in %.1f s (compat %.1f s) on %d threads"
        % (nlines, ev, dt, dtc, nth))
    return {"value": ev / dt, "unit": "Voigt evaluations/s", "cores": nth, "kind": "port",
            "sample": "FP64 oracle (Algorithm 916 + continued fraction, compensated sum) over the "
                      "full %d-point grid for the first %d lines of the workload" % (grid.nx, nlines),
            "seconds": dt, "checksum": float(np.sum(K)),
            "compat_value": ev / dtc, "compat_seconds": dtc, "build": flags,
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


def cpu_baseline_batched(w, arrays, grid, itp, Ksample, coeff_gpu):
    """Voigt leg: the oracle's line parameters + FP64 Voigt sum over the full
    grid for the first 1 % of each species' lines at (T, P) point itp.
    k-table leg: the oracle's per-bin sort and fit (findCut, rescale, ln,
    literal QR least squares) of the spectra in Ksample (>= 1 % of the
    bins), checked against the GPU's FP64 coefficients of the same spectra."""
    import numpy as np
    from heliosk import configs
    from oracle import oracle as O
    L, flags = O.lib_native()
    nth = omp_threads()
    T, Pbar = w.tp[itp]
    P = Pbar * configs.POBAR
    ev, dt, nl = 0.0, 0.0, 0
    for s, arr in enumerate(arrays):
        n = max(1, arr["nu"].size // 100)
        sub = {k: (v[:n] if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
        ol = configs.oracle_lines(sub)
        t0 = time.perf_counter()
        ost = O.line_params(ol, T, P, 0, w.cut, L=L)
        O.voigt_points(ost, grid.numin, grid.dnu, grid.nx, nthreads=nth, L=L)
        dt += time.perf_counter() - t0
        ev += _window_evals(ost["nu"], w.cut, grid)
        nl += n
    log("bench: CPU Voigt leg %d lines, %.3e evals in %.1f s on %d threads" % (nl, ev, dt, nth))
    nspec_kt = Ksample.shape[0]
    Nxb = grid.nx // w.nbins
    seg = O.bin_segments(grid.numin, w.numax, grid.dnu, grid.nx, w.nbins)[:w.nbins + 1]
    tk = 0.0
    dmax = 0.0
    for s in range(nspec_kt):
        t0 = time.perf_counter()
        Ks = O.sort_segments(Ksample[s], seg, L=L)
        co, _ = O.ktable(Ks, Nxb, w.nbins, w.nC, 0.0, L=L)
        tk += time.perf_counter() - t0
        if coeff_gpu is not None:
            dmax = max(dmax, float(np.max(np.abs(co - coeff_gpu[s]))))
    npts = nspec_kt * w.nbins * Nxb
    log("bench: CPU k-table leg %d bins (%d points) in %.2f s" % (nspec_kt * w.nbins, npts, tk))
    return {"value": ev / dt, "unit": "Voigt evaluations/s", "cores": nth, "kind": "port",
            "sample": "Voigt: FP64 oracle over the full %d-point grid for the first 1 %% of each "
                      "species' lines (%d lines) at T=%.1f K, P=%.3g bar; k-table: the oracle's "
                      "per-bin sort + fit of %d spectra (%d of %d bins)"
                      % (grid.nx, nl, T, Pbar, nspec_kt, nspec_kt * w.nbins,
                         len(arrays) * len(w.tp) * w.nbins),
            "seconds": dt,
            "ktable_value": npts / tk, "ktable_unit": "k-table points/s", "ktable_seconds": tk,
            "ktable_max_abs_coeff_diff_vs_gpu_fp64": dmax if coeff_gpu is not None else None,
            "build": flags, "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


if __name__ == "__main__":
    main()
