/*
 * heliosk.h -- C ABI of the MI355X line-by-line opacity engine.
 *
 * Plain C, plain pointers and sizes.  Every device buffer is owned by the
 * caller (as every device buffer of the reference is, host.h:792-876,
 * heliosk.cu:712-725); entry points that need scratch memory take a
 * workspace whose size a *_workspace_bytes() query returns.  Every call
 * takes an explicit stream (a hipStream_t; NULL = the default stream) and
 * returns HK_OK (0) or an error code; hk_last_error() gives the text of the
 * calling thread's last error.
 *
 * Reference seams replaced (exoclime/HELIOS-K 2.08):
 *   hk_lines_sort        thrust::sort_by_key + Sort_kernel/Copy_kernel
 *                        (heliosk.cu:1147-1171, voigt.h:843-866)
 *   hk_line_params       S2_kernel (voigt.h:331-363) + the per-line part of
 *                        readFile (host.h:641-644) + the cut of S3_kernel
 *                        (voigt.h:436-482)
 *   hk_voigt_accumulate  Line6A_Call/Line6B_Call/Line6C_Call (voigt.h:1820,
 *                        1968, 2004) with nuLimits/iiLimits (voigt.h:2045-2341)
 *                        and AddKStreams_kernel (voigt.h:728)
 *   hk_voigt_batched     the species/file/pressure loops of main
 *                        (heliosk.cu:784-1790) around those calls
 *   hk_ktable            the per-bin sort and resampling sequence of main
 *                        (heliosk.cu:2046-2193: thrust sorts, findCut,
 *                        rescale, copyK2, lnK, Vandermonde/QR, leastSquare)
 *   hk_ktable_reconstruct expfx_kernel + rescale(f=-1) (heliosk.cu:2189-2193)
 *   hk_profile_accumulate Line6fA/B/C_kernel<profile 1-4> + Plinth_kernel
 *                        (voigt.h:1091-1262)
 *   hk_means             Mean_kernel + IntegrateMean_kernel (resample.h:635-771)
 *   hk_transmission      Integrate_kernel (resample.h:541-614)
 *   hk_voigt_adjoint     no reference counterpart (gradient of the binned
 *                        opacity; see the function comment)
 */
#ifndef HELIOSK_H
#define HELIOSK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *hk_stream_t;      /* hipStream_t */

enum {
    HK_OK = 0,
    HK_ERR_ARG = 1,             /* invalid argument (message says which) */
    HK_ERR_HIP = 2,             /* HIP runtime error */
    HK_ERR_WORKSPACE = 3,       /* workspace too small */
    HK_ERR_UNSUPPORTED = 4
};

#define HK_MAX_ISO 16

/* Voigt evaluation forms. */
enum {
    HK_VOIGT_FP64 = 0,          /* plain FP64 profile, FP64 accumulation */
    HK_VOIGT_FP32C = 1,         /* FP32 profile, compensated FP32 tile sums */
    HK_VOIGT_COMPAT = 2         /* the reference's own FP32 region formulas
                                   (Line6fA/B/C_kernel, voigt.h:1177-1706:
                                   A S1/r2, B its rational approximation,
                                   C Algorithm 916 summed to def_TOLF), FP64
                                   accumulation; hk_voigt_accumulate only */
};

/* Line profiles (param.dat "profile", voigt.h:1235-1262). */
enum {
    HK_PROFILE_VOIGT = 1,
    HK_PROFILE_LORENTZ = 2,
    HK_PROFILE_DOPPLER = 3,
    HK_PROFILE_BINNED_GAUSS = 4     /* Gaussian averaged over each grid cell */
};

/* Line-list forms: how the file strength and widths become S(T), alpha_L(T,P).
 *  HK_FORM_HITRAN    HITRAN/HITEMP (S2_kernel, voigt.h:331-363): Boltzmann and
 *                    stimulated-emission factors relative to T0 = 296 K,
 *                    pressure shift delta P, P in atm.
 *  HK_FORM_EXOMOL    ExoMol / Kurucz / NIST / VALD (Sf_kernel, voigt.h:365-397):
 *                    exp(-c2 E/T) (1 - exp(-c2 nu/T)), no shift, P / 0.986923.
 *  HK_FORM_SUPERLINE ExoMol super-lines (SfSuper_kernel, voigt.h:399-420):
 *                    S as stored (no Q), alpha_L = vy (P/0.986923) (T0/T)^n. */
enum { HK_FORM_HITRAN = 0, HK_FORM_EXOMOL = 1, HK_FORM_SUPERLINE = 2 };

/* Raw line list of one species, device SoA, sorted by nu ascending.
 * Per line: nu [cm^-1], S = S_file (the strength as stored in the line
 * file), EL [cm^-1], A [s^-1], delta [cm^-1/atm], vy = ((1-q) g_air +
 * q g_self) * gammaF [cm^-1/atm], n (temperature exponent), iso (index
 * into the isotopologue tables; NULL means 0 for every line).
 * Strength at (T,P) follows readFile (host.h:641-644) and the form's
 * kernel: S_h = S_file / Q[iso] * iso_ab[iso] * iso_sscale[iso] (no Q for
 * HK_FORM_SUPERLINE).  A is the Einstein A (ExoMol-family lists: A plus the
 * natural-broadening gamma), added to alpha_L as A / (4 pi c). */
typedef struct {
    int64_t nlines;
    const double *nu, *S, *EL, *A, *delta, *vy, *n;
    const int32_t *iso;
    int niso;
    int form;                       /* HK_FORM_* */
    double iso_mass[HK_MAX_ISO];    /* mass of one molecule [g] (M / N_A) */
    double iso_ab[HK_MAX_ISO];
    double iso_sscale[HK_MAX_ISO];
    double max_abs_delta;           /* max |delta| over the lines */
} hk_lines;

/* One thermodynamic point. */
typedef struct {
    double T;                       /* K */
    double P;                       /* atm */
    double Q[HK_MAX_ISO];           /* partition function at T, per isotopologue */
} hk_tp;

/* Line cut: cut_mode 0 = cut in cm^-1, 1 = in Lorentz half widths,
 * 2 = in Doppler widths (S3_kernel, voigt.h:442-450).  cut == 0 means
 * no cut (heliosk.cu:941). */
typedef struct {
    int cut_mode;
    double cut;
} hk_cut;

/* Uniform grid: x_j = fma(j, dnu, numin), j = 0 .. nx-1 (setX_kernel). */
typedef struct {
    double numin, dnu;
    int64_t nx;
} hk_grid;

/* Line state at one (T,P), written by hk_line_params.  Device SoA. */
typedef struct {
    int64_t nlines;
    double *nu;         /* shifted centre nu + delta P */
    double *S;          /* S(T) * (1/alpha_D), as S2_kernel leaves it */
    double *ialphaD;    /* 1 / alpha_D */
    double *y;          /* alpha_L / alpha_D */
    double *hw;         /* cut half width [cm^-1]; < 0 marks an empty line */
    double *ey;         /* erfcx(y), used by the line-core series */
    double *blk_max;    /* max S over each block of 256 lines, ceil(n/256) */
    double *radius;     /* radius[0] = max_i (hw_i + |nu_i - key_i|), radius[1] =
                           max_i (min(hw_i, 10 alpha_D,i) + |nu_i - key_i|) */
    const double *key;  /* sorted search key (the raw nu), set by hk_line_params */
} hk_line_state;

/* Launch tuning; zero fields select the defaults. */
typedef struct {
    int tile_points;    /* grid points per workgroup tile: 512, 1024
                           (default), 2048 or 4096 */
    int exact_pairs;    /* HK_VOIGT_FP32C: 1 evaluates every (line, point)
                           pair; 0 (default) sums lines far from a tile,
                           or from a wave's quarter of it, through
                           Chebyshev far fields */
    int stage_lines;    /* lines staged through LDS per step of the main
                           pass: 256 (default) or 512 (the reference's fixed
                           line blocks def_nlA/B/C, define.h:56-58); FP64
                           results do not depend on it, FP32C results only
                           by rounding (a stage's lines are added path by
                           path, so the order changes with the stage) */
} hk_tune;

const char *hk_last_error(void);
const char *hk_version(void);

/* Bytes of device memory an hk_line_state of n lines needs (one
 * allocation holding nu, S, ialphaD, y, hw, ey, blk_max and radius). */
size_t hk_line_state_bytes(int64_t nlines);
/* Lay out an hk_line_state inside caller memory of hk_line_state_bytes. */
int hk_line_state_init(hk_line_state *st, int64_t nlines, void *mem, size_t bytes);

/* Order a raw line list by centre: the device sort of main
 * (thrust::sort_by_key on nu, then the Sort_kernel / Copy_kernel gathers of
 * every field, heliosk.cu:1147-1171, voigt.h:843-866).  Stable: lines with
 * equal nu keep their input order.  `in` may be in any order; the arrays
 * `out` points to (nu, S, EL, A, delta, vy, n and, when in->iso is not NULL,
 * iso) receive the sorted fields and must be distinct from in's, each with
 * room for in->nlines entries.  out's scalar fields (nlines, niso, form, the
 * isotopologue tables, max_abs_delta) are set from in.  At most 2^31 - 1
 * lines per call; the workspace holds two permutations and the radix-sort
 * scratch. */
size_t hk_lines_sort_workspace_bytes(int64_t nlines);
int hk_lines_sort(const hk_lines *in, hk_lines *out, void *workspace, size_t bytes,
                  hk_stream_t stream);

/* Line parameters at one (T,P) -- S2_kernel semantics, FP64. */
int hk_line_params(const hk_lines *lines, const hk_tp *tp, const hk_cut *cut,
                   hk_line_state *st, hk_stream_t stream);

/* Voigt sum onto the grid: K_j = kmin + sum_i S_i Re w(x_ij, y_i)/sqrt(pi)
 * (accumulate == 0) or K_j += sum (accumulate != 0), over the lines whose
 * cut window holds j: |x_j - nu_i| < hw_i in FP64.  Deterministic: each
 * point is owned by one thread that adds the lines in index order.
 * HK_VOIGT_COMPAT instead evaluates each pair as the reference's FP32
 * kernels do (x = va + j vb in float, kept when x^2 < vcut2, region formulas
 * of Line6fA/B/C_kernel, S3_kernel's FP32 line constants) and adds the
 * values in FP64; its region-B formula is within 1.4e-2 of the exact
 * profile, as the reference's is. */
size_t hk_voigt_workspace_bytes(const hk_grid *grid, const hk_tune *tune);
int hk_voigt_accumulate(const hk_line_state *st, const hk_grid *grid,
                        double kmin, int accumulate, int mode,
                        const hk_tune *tune, double *K,
                        void *workspace, size_t workspace_bytes,
                        hk_stream_t stream);

/* hk_voigt_accumulate with no workspace argument and no device allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The tile tables live in a static device array of
 * HK_VOIGT_NOALLOC_SCRATCH_BYTES held by every GPU the library runs on until
 * the process ends; grid->nx may be at most HK_VOIGT_NOALLOC_MAX_NX (larger
 * grids return HK_ERR_ARG).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other host threads) race and
 * must be serialised by the caller.  Same kernels and results as
 * hk_voigt_accumulate, bit for bit. */
#define HK_VOIGT_NOALLOC_MAX_NX ((int64_t)1 << 25)
#define HK_VOIGT_NOALLOC_SCRATCH_BYTES (((size_t)1 << 23) + 512)
int hk_voigt_accumulate_noalloc(const hk_line_state *st, const hk_grid *grid,
                                double kmin, int accumulate, int mode,
                                const hk_tune *tune, double *K, hk_stream_t stream);

/* Algorithmic work of one Voigt sum: counts[0] = line-point pairs inside
 * the cut windows (the evaluations), counts[1..3] = those with
 * x^2+y^2 >= 1e6, in [100, 1e6), < 100 (the reference's regions A, B, C,
 * voigt.h:1223-1230, 1490, 1665).  counts: 4 device uint64. */
int hk_voigt_pair_count(const hk_line_state *st, const hk_grid *grid,
                        unsigned long long *counts, hk_stream_t stream);

/* Work actually done by one hk_voigt_accumulate call with this mode and
 * tuning (tile width set by tile_points): counts (12 device uint64) =
 * [direct pairs in regions A, B, C (as hk_voigt_pair_count), pairs summed
 * through the FP32 far fields, far (line, tile) pairs of band 1 -- each 9
 * node evaluations --, of band 2 -- 15 each --, of band 3 -- 19 each --,
 * far (line, wave) pairs of the wave-level far field in bands 1, 2, 3 (the
 * same node counts), cut-edge (line, wave) pairs and their point pairs].  A
 * far line whose window ends inside a tile counts as a far (line, tile)
 * pair; its pairs in the wave that holds the window's end are cut-edge
 * pairs when that wave takes them through its per-wave expansion (counted
 * so whenever the line qualifies), direct otherwise.  The direct, far-field
 * and cut-edge pairs add up to counts[0] of hk_voigt_pair_count. */
int hk_voigt_work_count(const hk_line_state *st, const hk_grid *grid, int mode,
                        const hk_tune *tune, unsigned long long *counts, hk_stream_t stream);

/* Autotuner (replaces the reference's launch self-tuning, heliosk.cu:1363-
 * 1401): times hk_voigt_accumulate on this input for each tile size in
 * {512, 1024, 2048, 4096} points and each stage of {256, 512} lines (one
 * untimed and one timed launch each, device events) and writes the fastest
 * pair into *best; ms[8] (may be NULL) receives the timings, tile-major
 * within each stage size (ms[4 * (stage == 512) + tile index]).  K is overwritten; the workspace must be sized
 * for the smallest tile (hk_voigt_workspace_bytes covers every size).
 * Waits for the stream.  Results of different tile sizes agree to
 * rounding, not bit for bit, so the choice is cached for the process per
 * problem signature (form, exact_pairs, grid, line count, and the lines'
 * cut reach and median y, each to 3 significant digits): a repeated tuning
 * of the same problem returns the first choice without timing (ms = 0), and
 * reruns give the same bits.  best->exact_pairs is read: the sizes are
 * timed with that setting.  hk_voigt_autotune_reset empties the cache. */
int hk_voigt_autotune(const hk_line_state *st, const hk_grid *grid, int mode,
                      double *K, void *workspace, size_t workspace_bytes,
                      hk_tune *best, float *ms, hk_stream_t stream);
void hk_voigt_autotune_reset(void);

/* Many spectra in one persistent launch (work queue over spectrum tiles). */
typedef struct {
    hk_line_state st;
    hk_grid grid;
    double kmin;
    int accumulate;
    double *K;          /* device, grid.nx values */
} hk_spectrum;

size_t hk_voigt_batched_workspace_bytes(int nspec, const hk_spectrum *spec,
                                        const hk_tune *tune);
int hk_voigt_batched(int nspec, const hk_spectrum *spec, int mode,
                     const hk_tune *tune, void *workspace,
                     size_t workspace_bytes, hk_stream_t stream);

/* The same in two steps.  _prepare uploads the batch tables into the
 * workspace (returns after the upload) and fills *batch; _launch issues
 * kernels only (no allocation, copy or host synchronisation), so it can be
 * captured in a hipGraph and replayed: each replay recomputes every
 * spectrum from the current contents of the line states. */
typedef struct {
    void *workspace;
    int nspec, mode, tile_points, stage_lines;
    int64_t items_main, items_core;
    size_t off_args, off_cargs, off_items, off_citems, off_counter;
} hk_voigt_batch;

int hk_voigt_batched_prepare(int nspec, const hk_spectrum *spec, int mode,
                             const hk_tune *tune, void *workspace,
                             size_t workspace_bytes, hk_voigt_batch *batch,
                             hk_stream_t stream);
int hk_voigt_batched_launch(const hk_voigt_batch *batch, hk_stream_t stream);

/* k-distribution tables.  K holds nspec spectra of grid->nx points; bins are
 * equal in wavenumber over [numin, numax) and hold Nxb = nx/nbins points.
 * hk_ktable_prepare fills the workspace with the per-(Nxb, nC) least-squares
 * map once; hk_ktable then sorts each bin ascending (Ksorted, which may be
 * NULL), finds Nxmin, rescales, takes ln(K * unit_scale) and fits nC
 * Chebyshev coefficients per bin (coeff[(s*nbins + b)*nC + c]). */
size_t hk_ktable_workspace_bytes(int64_t nx, int nbins, int nC, int nspec);
int hk_ktable_prepare(const hk_grid *grid, double numax, int nbins, int nC,
                      void *workspace, size_t workspace_bytes);
int hk_ktable(const double *K, int nspec, const hk_grid *grid, double numax,
              int nbins, int nC, double kmin, double unit_scale,
              double *coeff, int32_t *nxmin, double *Ksorted,
              void *workspace, size_t workspace_bytes, hk_stream_t stream);
/* Gradient of the k-table coefficients with respect to the spectra (no
 * reference counterpart: the backward of the sequence hk_ktable replaces,
 * heliosk.cu:2046-2193 and resample.h:243-481).  gcoeff = dL/dc in
 * hk_ktable's coeff layout; writes gK = dL/dK (nspec x nx).  K is the input
 * of the hk_ktable call, Ksorted and nxmin its outputs (the sorted bins are
 * required).  The derivative runs through the least-squares fit, ln and the
 * rescale (linear between sorted values); Nxmin and the sort order are
 * piecewise constant in K and held fixed; values equal within a bin share
 * the mean gradient of their ranks.  The workspace, of
 * hk_ktable_backward_workspace_bytes, must have been hk_ktable_prepare'd
 * (it may also serve the forward call).  Bins must be the fixed Nxb chunks
 * of the grid (HK_ERR_UNSUPPORTED otherwise). */
size_t hk_ktable_backward_workspace_bytes(int64_t nx, int nbins, int nC, int nspec);
int hk_ktable_backward(const double *K, const double *Ksorted, const int32_t *nxmin,
                       int nspec, const hk_grid *grid, double numax, int nbins, int nC,
                       const double *gcoeff, double *gK, void *workspace,
                       size_t workspace_bytes, hk_stream_t stream);
/* K(y) from the coefficients: exp(Chebyshev sum) on y = k/Nxb, mapped back
 * over [Nxmin, Nxb) (expfx_kernel + rescale f=-1).  scratch: nspec*nbins*Nxb
 * doubles of device memory. */
int hk_ktable_reconstruct(const double *coeff, const int32_t *nxmin, int nspec,
                          int nbins, int Nxb, int nC, double kmin,
                          double unit_scale, double *Ky, double *scratch,
                          hk_stream_t stream);
/* The least-squares map of hk_ktable_prepare on the host: P is nC x 2nC,
 * c = P[:, :nC] m + P[:, nC:] b0 with m_i = sum_k T_i(x_k) b_k and
 * b0 = b[0:nC]; it reproduces leastSquare_kernel (NC-1 reflections). */
int hk_ktable_plan_host(int Nxb, int nC, double *P);

/* Sum of one line list with any profile, optionally with each line's
 * plinth removed (Line6fA/B/C_kernel<profile> and Plinth_kernel,
 * voigt.h:1091-1262; S1 as S3_kernel forms it, voigt.h:458-470).  In Doppler
 * units x = (x_j - nu)/alpha_D, per pair inside the cut:
 *   HK_PROFILE_VOIGT         S Re w(x, y)/sqrt(pi)  (hk_voigt_accumulate in
 *                            voigt_mode: HK_VOIGT_FP64, HK_VOIGT_FP32C or
 *                            HK_VOIGT_COMPAT)
 *   HK_PROFILE_LORENTZ       (S y/pi) / (x^2 + y^2)
 *   HK_PROFILE_DOPPLER       (S/sqrt(pi)) exp(-x^2)
 *   HK_PROFILE_BINNED_GAUSS  (S/vb) (erf(x + vb/2) - erf(x - vb/2))/2, vb = dnu/alpha_D
 * remove_plinth subtracts from every pair the line's profile at its cut
 * (profiles 1-3; the reference rejects it for profile 4, and so does this
 * call).  Profiles 2-4 and the plinth are FP64.  Workspace: the Voigt
 * workspace or the tile table, whichever is larger, plus one double per
 * line. */
/* Sub-Lorentzian line wings (Perrin & Hartmann 1989; the subLorentzianfile
 * option, host.h:489-519, sLChi voigt.h:47-72): for range[0] < x < range[1]
 * every pair's profile value is multiplied by chi(|x - nu|) x / nu, with
 * chi = 1 below sigma[0], then exp(-B[0](d - sigma[0])), continued with
 * slopes B[1] above sigma[1] and B[2] above sigma[2]; B at the run's T. */
typedef struct {
    double sigma[3];
    double B[3];
    double range[2];
} hk_chi;

size_t hk_profile_workspace_bytes(const hk_grid *grid, int64_t nlines);
/* chi: NULL, or the sub-Lorentzian factor; with it every profile (Voigt
 * through the FP64 Faddeeva function, voigt_mode then unused) runs in the
 * FP64 gather. */
int hk_profile_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin,
                          int accumulate, int profile, int remove_plinth, int voigt_mode,
                          const hk_chi *chi, double *K, void *workspace,
                          size_t workspace_bytes, hk_stream_t stream);

/* Planck and Rosseland mean sums of one spectrum on a uniform grid
 * (Mean_kernel + IntegrateMean_kernel, resample.h:635-771, as main calls
 * them at heliosk.cu:1953-2000).  With x_j = fma(j, dnu, numin),
 * B_j the Planck function and dB_j its T derivative (B = dB = 0 at x = 0),
 * sums (4 doubles, device) = [sum_j B_j K_j, sum_{K_j != 0} dB_j / K_j,
 * sum_j B_j, sum_j dB_j]; the Planck mean is sums[0]/sums[2], the
 * Rosseland mean sums[3]/sums[1].  Plain sums, as the reference writes them
 * (its Simpson end weights are still zero when the means run). */
size_t hk_means_workspace_bytes(void);
int hk_means(const double *K, const hk_grid *grid, double T, double *sums,
             void *workspace, size_t workspace_bytes, hk_stream_t stream);

/* The same sums on explicit points x (per-bin grids: the reference's
 * useIndividualX branch, resample.h:700-716): trapezoid rule,
 * sum_j (f_j + f_j+1)/2 (x_j+1 - x_j); the Rosseland term only where both
 * K are nonzero. */
int hk_means_x(const double *K, const double *x, int64_t nx, double T, double *sums,
               void *workspace, size_t workspace_bytes, hk_stream_t stream);

/* Transmission function of each bin (Integrate_kernel, resample.h:541-614;
 * heliosk.cu:2400-2440).  Ksorted holds nbins bins of Nxb ascending values
 * (hk_ktable's Ksorted, or its reconstruction); nxmin (device, NULL = all 0)
 * the per-bin count of values at kmin.  For t = 0 .. nTr-1,
 * m = exp((t - nTr/2) dTr) and Tr[b*nTr + t] is the Simpson-corrected mean
 * of exp(-K m) over the bin, blended with exp(-kmin m) by Nxmin/Nxb. */
int hk_transmission(const double *Ksorted, int nbins, int Nxb, const int32_t *nxmin,
                    double kmin, int nTr, double dTr, double *Tr, hk_stream_t stream);

/* Adjoint of the Voigt sum: given gK_j = dL/dK_j on the grid, writes
 * gS[i] = dL/dS_i (S = the raw S_file), gvy[i] = dL/dvy_i and
 * gTP[0..1] = dL/dT, dL/dP, holding the partition function fixed.  The
 * profile is recomputed (FP64), not stored.  gTP is device memory. */
size_t hk_voigt_adjoint_workspace_bytes(int64_t nlines);
int hk_voigt_adjoint(const hk_lines *lines, const hk_tp *tp, const hk_cut *cut,
                     const hk_grid *grid, const double *gK, double *gS,
                     double *gvy, double *gTP, void *workspace,
                     size_t workspace_bytes, hk_stream_t stream);

/* Ascending sort of the bins [seg[b], seg[b+1]), b < nseg, of one
 * spectrum of nx points into Ksorted (the per-bin thrust sorts of
 * heliosk.cu:2046-2051 for irregular bins).  seg is a host array of nseg+1
 * indices; the call returns after the segment table is uploaded. */
size_t hk_sort_bins_workspace_bytes(int64_t nx, int nseg);
int hk_sort_bins(const double *K, int64_t nx, const int64_t *seg, int nseg, double *Ksorted,
                 void *workspace, size_t workspace_bytes, hk_stream_t stream);

/* Host helpers. */
/* First grid index of each bin by binKey semantics (voigt.h:779-806):
 * seg[b], b = 0..nbins+1 (seg[nbins+1] = nx). */
int hk_bin_segments(const hk_grid *grid, double numax, int nbins, int64_t *seg);
/* Seeded synthetic line list (counter-based SplitMix64; identical on
 * every host): fields nu, S, EL, gamma (= g_air = g_self), n.  kind 0:
 * nu U[0,30000), log10 S U[-30,-19], EL U[0,20000), gamma U[0.01,0.1],
 * n U[0.3,0.8]; kind 1: gamma log-uniform [1e-4, 1] and 5% strong lines
 * with log10 S U[-20,-17].  Unsorted. */
int hk_synth_lines(uint64_t seed, int species, int64_t n, int kind,
                   double *nu, double *S, double *EL, double *gamma,
                   double *nexp);

#ifdef __cplusplus
}
#endif
#endif
