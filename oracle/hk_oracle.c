/*
 * hk_oracle.c -- FP64 host oracle for the HELIOS-K line-by-line path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it as the checker / CPU restatement.
 * The product path never links or loads it.
 *
 * What each part restates and what pins it:
 *
 *  hko_erfcx, hko_w916, hko_wcf, hko_faddeeva
 *      Faddeeva function.  hko_w916 is Zaghloul & Ali, "Algorithm 916"
 *      (ACM TOMS 38(2), 2011), eqs. for Re w (as in voigt.h:165-216 and
 *      voigt.h:273-296 of the reference) plus the matching Im w series.
 *      hko_wcf is the Laplace continued fraction
 *      w(z) = (i/sqrt(pi)) / (z - (1/2)/(z - 1/(z - (3/2)/(z - ...)))).
 *      Pinned by 50-digit mpmath golden vectors (tests/golden) and by
 *      scipy.special.wofz, plus the closed forms w(iy) = erfcx(y),
 *      Re w(x,0) = exp(-x^2), and integral(Re w dx) = sqrt(pi).
 *  hko_line_params
 *      host.h:641-644 (readFile: vy, 1/alpha_D, S/Q) followed by
 *      voigt.h:338-356 (S2_kernel) and the cut of voigt.h:442-450
 *      (S3_kernel) expressed as a half-width in cm^-1.  Pinned by closed
 *      forms (S(T0) = S_file/Q*Ab, alpha_D ~ nu sqrt(T), alpha_L ~ P
 *      (T0/T)^n) checked in tests/test_oracle.py.
 *  hko_window
 *      The cut test |x_j - nu_i| < cut (voigt.h:1223 with vcut2 of
 *      voigt.h:442-450) evaluated in FP64 on x_j = fma(j, dnu, numin)
 *      (setX_kernel, voigt.h:751-759).  Pinned by brute force.
 *  hko_voigt_points
 *      K_j = kmin + sum_i S_i Re w(x_ij, y_i)/sqrt(pi): the sum the
 *      Line6fA/B/C kernels accumulate (voigt.h:1177-1281, 1437-1515,
 *      1616-1706) onto InitialK's kmin (voigt.h:713-722).  Pinned by
 *      linearity, translation, symmetry and the normalisation
 *      sum_j K_j dnu -> sum_i S_i for an uncut line.
 *  hko_profile_value, hko_profile_points
 *      Profiles 2-4 (Lorentz, Doppler, binned Gaussian; voigt.h:1235-1262
 *      with the S1 of S3_kernel, voigt.h:458-470) and the plinth removal
 *      of Plinth_kernel (voigt.h:1091-1164).  Pinned by closed forms: the
 *      binned Gaussian sums to S over the grid exactly (the erf differences
 *      telescope), the Doppler and Lorentz sums approach S (and the
 *      Lorentz tail's S y / (pi x)), the plinth-removed profile vanishes at
 *      the cut and equals the plain one minus a constant per line.
 *  hko_chi (and the chi argument of hko_profile_points)
 *      Sub-Lorentzian wing factor (sLChi, voigt.h:47-72, applied as
 *      chi x / nu at voigt.h:1345-1350).  Pinned by its properties:
 *      chi = 1 below sigma1, continuous at sigma1..3, log-slope -B_i in
 *      segment i, and the x / nu factor between points at equal |x - nu|.
 *  hko_voigt_grad
 *      Analytic gradient of sum_j gK_j K_j; pinned by central finite
 *      differences of hko_voigt_points.
 *  hko_bin_segments, hko_sort_segments
 *      binKey_kernel (voigt.h:779-806) + the two thrust sorts of
 *      heliosk.cu:2049-2050.  Pinned by plots/p004 and plots/p009.
 *  hko_qr, hko_ktable, hko_reconstruct
 *      resample.h:14-37 (Vandermonde), 73-227 (QR), 243-353 (least squares,
 *      NC-1 reflections), 362-388 (findCut), 398-437 (rescale),
 *      472-481 (lnK), 497-524 (expfx).  Pinned by plots/p004 and plots/p005
 *      (Out_r_cbin.dat, Out_r_bin.dat).
 *  hko_means
 *      Mean_kernel + IntegrateMean_kernel (resample.h:635-771): the four
 *      Planck / Rosseland sums, without the Simpson end terms (their weights
 *      are still zero in the reference when the means run, heliosk.cu:
 *      1953-2000 before SimpsonCoefficient at 2392).  Pinned by
 *      plots/p009/Out_p{1,3}_mean.dat and by closed forms (constant K gives
 *      both means = K; sum B dnu -> sigma T^4 / pi).
 *  hko_means_x
 *      The trapezoid branch of IntegrateMean_kernel (resample.h:700-716) for
 *      per-bin grids.  Pinned by closed forms (constant K; the trapezoid
 *      integral of B converges to sigma T^4 / pi).
 *  hko_transmission
 *      Integrate_kernel (resample.h:541-614) with SimpsonCoefficient's end
 *      weights (resample.h:618-626).  Pinned by plots/p006/Out_i_tr0000.dat
 *      and _tr0004.dat (from Out_i_bin0000/0004.dat) and by the closed form
 *      Tr(m) = exp(-K m) for a constant bin.
 *
 * Floating-point contraction: the reference was compiled by nvcc, whose
 * default --fmad=true turns a*b+c into one fused multiply-add.  Where the
 * reference writes such an expression and the rounding matters (grid
 * positions, rescale interpolation) this file writes fma() explicitly.
 */
#include "hk_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HKO_SQRTPI  1.77245385090551602730
#define HKO_ISQRTPI 0.56418958354775628695

/* ------------------------------------------------------------------ */
/* erfcx(y) = exp(y^2) erfc(y), y >= 0                                 */
/* ------------------------------------------------------------------ */
double hko_erfcx(double y)
{
    if (y < 0.0) return 2.0 * exp(y * y) - hko_erfcx(-y);
    if (y < 5.0) return exp(y * y) * erfc(y);
    /* Laplace continued fraction, evaluated backward. */
    double t = y;
    for (int k = 60; k >= 1; --k) t = y + 0.5 * k / t;
    return HKO_ISQRTPI / t;
}

static double sinc(double t) { return t == 0.0 ? 1.0 : sin(t) / t; }

/* ------------------------------------------------------------------ */
/* Algorithm 916 (Zaghloul & Ali 2011), FP64, real and imaginary part. */
/* Series truncation: all terms below exp(-45) relative are dropped.   */
/* ------------------------------------------------------------------ */
void hko_w916(double x, double y, double tol, double *re, double *im)
{
    double sgn = x < 0.0 ? -1.0 : 1.0;
    x = fabs(x);
    if (x == 0.0) { *re = hko_erfcx(y); *im = 0.0; return; }
    double a = M_PI / sqrt(-log(tol / 2.0));
    double yy = y * y, xx = x * x;
    double ex2 = exp(-xx);
    int N1 = (int)ceil(6.8 / a) + 1;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
    for (int n = 1; n <= N1; ++n) {
        double an = a * n, den = an * an + yy;
        double e1 = exp(-(an * an + xx));
        double e2 = exp(-(an + x) * (an + x));
        s1 += e1 / den;
        s2 += e2 / den;
        s5 += an * e2 / den;
    }
    long n0 = (long)ceil(x / a);
    long nlo = n0 - N1; if (nlo < 1) nlo = 1;
    long nhi = n0 + N1;
    for (long n = nlo; n <= nhi; ++n) {
        double an = a * (double)n, den = an * an + yy;
        double e3 = exp(-(an - x) * (an - x));
        s3 += e3 / den;
        s4 += an * e3 / den;
    }
    double xy = x * y, c2 = cos(2.0 * xy), s2xy = sin(2.0 * xy);
    double sxy = sin(xy);
    double ca = 2.0 * a / M_PI;
    double ex = hko_erfcx(y);
    double r, i;
    if (y == 0.0) {
        r = ex2;
    } else {
        r = ex2 * ex * c2 + ca * (sxy * x * sinc(xy) * ex2
                                  + y * (-c2 * s1 + 0.5 * (s2 + s3)));
    }
    i = -ex2 * ex * s2xy + ca * (x * ex2 * sinc(2.0 * xy) + y * s2xy * s1
                                 + 0.5 * s4 - 0.5 * s5);
    *re = r;
    *im = sgn * i;
}

/* ------------------------------------------------------------------ */
/* Laplace continued fraction, backward evaluation in real arithmetic. */
/* t_N = z; t_k = z - (k/2)/t_{k+1}; w = i / (sqrt(pi) t_1).           */
/* ------------------------------------------------------------------ */
void hko_wcf(double x, double y, int depth, double *re, double *im)
{
    double tr = x, ti = y;
    for (int k = depth; k >= 1; --k) {
        double m = 0.5 * k / (tr * tr + ti * ti);
        tr = x - m * tr;
        ti = y + m * ti;
    }
    double q = HKO_ISQRTPI / (tr * tr + ti * ti);
    *re = ti * q;
    *im = tr * q;
}

void hko_faddeeva(double x, double y, double *re, double *im)
{
    double r2 = x * x + y * y;
    if (r2 < 100.0) { hko_w916(x, y, 1.0e-17, re, im); return; }
    int depth = r2 >= 1.0e4 ? 8 : (r2 >= 1.0e3 ? 10 : (r2 >= 400.0 ? 12 : 20));
    hko_wcf(x, y, depth, re, im);
    if (y == 0.0) *re = exp(-x * x);
}

void hko_faddeeva_batch(int64_t n, const double *x, const double *y,
                        double *re, double *im)
{
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) hko_faddeeva(x[k], y[k], &re[k], &im[k]);
}

/* ------------------------------------------------------------------ */
/* Line parameters                                                     */
/* ------------------------------------------------------------------ */
void hko_line_params(int64_t nl,
                     const double *nu0, const double *S0, const double *EL,
                     const double *A, const double *delta, const double *vy0,
                     const double *nexp, const int32_t *iso,
                     const double *iso_mass, const double *iso_ab,
                     const double *iso_sscale, const double *iso_Q,
                     double T, double P, int cutMode, double cut, int form,
                     double *nu_o, double *S_o, double *ia_o, double *y_o,
                     double *hw_o)
{
    const double c = HKO_H * HKO_C / HKO_KB;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nl; ++i) {
        int k = iso ? iso[i] : 0;
        /* readFile, host.h:641-644; L_kernelExomol / L_kernelExomolSuper,
           voigt.h:484-531 (super-lines: no partition function) */
        double ialphaD_h = HKO_C * sqrt(iso_mass[k] / (2.0 * HKO_KB * T));
        double Sh = form == 2 ? S0[i] * iso_ab[k] * iso_sscale[k]
                              : S0[i] / iso_Q[k] * iso_ab[k] * iso_sscale[k];
        double nu, S, alphaL;
        double E = EL[i];
        if (form == 0) {
            /* S2_kernel, voigt.h:338-356 */
            nu = fma(delta[i], P, nu0[i]);
            double ia = ialphaD_h / nu;
            S = Sh * (exp(-E * c / T + E * c / HKO_T0) * (1.0 - exp(-c * nu / T))
                      / (1.0 - exp(-c * nu / HKO_T0)) * ia);
            alphaL = fma(vy0[i], (P / HKO_POATM) * pow(HKO_T0 / T, nexp[i]),
                         A[i] / (4.0 * M_PI * HKO_C));
        } else if (form == 1) {
            /* Sf_kernel, voigt.h:371-384, c1 = h c / (kB T), T1 = T0 / T */
            double c1 = HKO_H * HKO_C / (HKO_KB * T);
            nu = nu0[i];
            double ia = ialphaD_h / nu;
            S = Sh * (exp(-c1 * E) * (1.0 - exp(-c1 * nu)) * ia);
            alphaL = fma(vy0[i], (P / HKO_POBAR) * pow(HKO_T0 / T, nexp[i]),
                         A[i] / (4.0 * M_PI * HKO_C));
        } else {
            /* SfSuper_kernel, voigt.h:405-414 */
            nu = nu0[i];
            double ia = ialphaD_h / nu;
            S = Sh * ia;
            alphaL = vy0[i] * ((P / HKO_POBAR) * pow(HKO_T0 / T, nexp[i]));
        }
        double ialphaD = ialphaD_h / nu;
        double y = alphaL * ialphaD;
        double hw;
        if (cutMode == 1) hw = cut * alphaL;
        else if (cutMode == 2) hw = cut / ialphaD;
        else hw = cut;
        if (nu == 0.0) { S = 0.0; ialphaD = 0.0; y = 0.0; hw = -1.0; }
        nu_o[i] = nu; S_o[i] = S; ia_o[i] = ialphaD; y_o[i] = y; hw_o[i] = hw;
    }
}

/* ------------------------------------------------------------------ */
/* Window: j in [jlo, jhi)  <=>  |fma(j,dnu,numin) - nu| < hw           */
/* ------------------------------------------------------------------ */
static int64_t clamp_index(double e, int64_t Nx)
{
    if (!(e > 0.0)) return 0;             /* also NaN */
    if (e >= (double)Nx) return Nx;
    return (int64_t)floor(e);
}

void hko_window(double nu, double hw, double numin, double dnu, int64_t Nx,
                int64_t *jlo, int64_t *jhi)
{
    if (!(hw > 0.0)) { *jlo = 0; *jhi = 0; return; }
    int64_t lo = clamp_index((nu - hw - numin) / dnu, Nx);
    while (lo > 0 && fma((double)(lo - 1), dnu, numin) - nu > -hw) --lo;
    while (lo < Nx && !(fma((double)lo, dnu, numin) - nu > -hw)) ++lo;
    int64_t hi = clamp_index((nu + hw - numin) / dnu, Nx);
    while (hi > 0 && !(fma((double)(hi - 1), dnu, numin) - nu < hw)) --hi;
    while (hi < Nx && fma((double)hi, dnu, numin) - nu < hw) ++hi;
    if (hi < lo) hi = lo;
    *jlo = lo; *jhi = hi;
}

/* ------------------------------------------------------------------ */
/* Voigt sum (gather over points, Neumaier-compensated)                */
/* ------------------------------------------------------------------ */
static int64_t lower_bound(const double *v, int64_t n, double key)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t m = (lo + hi) / 2; if (v[m] < key) lo = m + 1; else hi = m; }
    return lo;
}

void hko_voigt_points(int64_t nl, const double *nu, const double *S,
                      const double *ialphaD, const double *y,
                      const double *hw, double R,
                      double numin, double dnu, int64_t Nx, double kmin,
                      int64_t npts, const int64_t *jidx, double *K,
                      int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < npts; ++p) {
        int64_t j = jidx ? jidx[p] : p;
        double xj = fma((double)j, dnu, numin);
        int64_t i0 = lower_bound(nu, nl, xj - R);
        int64_t i1 = lower_bound(nu, nl, xj + R);
        if (i1 < nl) ++i1;
        double sum = kmin, comp = 0.0;
        for (int64_t i = i0; i < i1; ++i) {
            double d = xj - nu[i];
            if (!(d < hw[i] && d > -hw[i])) continue;
            double re, im;
            hko_faddeeva(d * ialphaD[i], y[i], &re, &im);
            double t = S[i] * re * HKO_ISQRTPI;
            double s2 = sum + t;
            if (fabs(sum) >= fabs(t)) comp += (sum - s2) + t;
            else comp += (t - s2) + sum;
            sum = s2;
        }
        K[p] = sum + comp;
    }
}

/* ------------------------------------------------------------------ */
/* The reference's single-precision Voigt arithmetic ("compat").       */
/* S3_kernel (voigt.h:436-482) rounds the FP64 line state to FP32:     */
/*   S1 = (float)(S y / pi), Sf = (float)S, yf = (float)y,             */
/*   va = (float)((numin - nu) ialphaD), vb = (float)(dnu ialphaD),    */
/*   vcut2 = (float)(hw^2 ialphaD^2)  (cut mode 0: cut^2 ialphaD^2);   */
/* each pair: x = fmaf((float)j, vb, va), t1 = x^2, r2 = t1 + yf^2,    */
/* kept when t1 < vcut2, and                                           */
/*   r2 >= 1e6 (Line6fA_kernel, voigt.h:1177-1281):  S1 / r2           */
/*   100 <= r2 < 1e6 (Line6fB_kernel, voigt.h:1437-1515): S1 t4,       */
/*     t4 = (pi t2 (2 t2 + r2) - 12 x^2) / (3 r2 (pi t2^2 - 6 x^2)),    */
/*     t2 = r2 + 1.5                                                   */
/*   r2 < 100 (Line6fC_kernel, voigt.h:1616-1706; Sigmabf,             */
/*     voigt.h:217-268): Algorithm 916 in FP32, each series summed     */
/*     until all three change by less than def_TOLF = 2.48e-12f, with  */
/*     a = pi/sqrt(-ln(TOLF/2)), b = 1/sqrt(pi), c = 2a/pi             */
/*     (heliosk.cu:1656-1658), times Sf b;                             */
/* values added in FP64 in line order.  erfcxf(yf) is hko_erfcx(yf)    */
/* rounded to float.  Built with -ffp-contract=off, so every float     */
/* operation rounds as written.                                        */
/* ------------------------------------------------------------------ */
#define HKO_TOLF 2.48e-12f
#define HKO_PIF 3.14159265358979323846f

typedef struct { float a, b, c; } hko_cabc;

static hko_cabc hko_compat_consts(void)
{
    hko_cabc k;
    double a = (double)(float)(M_PI * sqrt(-1.0 / log((double)HKO_TOLF * 0.5)));
    k.a = (float)a;
    k.b = (float)(1.0 / sqrt(M_PI));
    k.c = (float)(2.0 * a / M_PI);
    return k;
}

static void hko_sigmabf(float x, float y, float a, float *s1o, float *s2o, float *s3o)
{
    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, o1 = 0.0f, o2 = 0.0f, o3 = 0.0f;
    float yy = y * y;
    x = fabsf(x);
    int n0 = (int)ceilf(x / a);
    int st1 = 0, st2 = 0, st3 = 0;
    for (int n = 1; n < 100; ++n) {
        int n3p = n0 + n - 1, n3n = n0 - n;
        float an = a * (float)n, an3p = a * (float)n3p, an3n = a * (float)n3n;
        float f = 1.0f / (an * an + yy);
        float f3p = 1.0f / (an3p * an3p + yy);
        float f3n = 1.0f / (an3n * an3n + yy);
        s1 += f * expf(-(an * an + x * x));
        s2 += f * expf(-(an + x) * (an + x));
        s3 += f3p * expf(-(an3p - x) * (an3p - x));
        if (n3n >= 1) s3 += f3n * expf(-(an3n - x) * (an3n - x));
        if (fabsf(s1 - o1) < HKO_TOLF) st1 = 1;
        if (fabsf(s2 - o2) < HKO_TOLF) st2 = 1;
        if (fabsf(s3 - o3) < HKO_TOLF) st3 = 1;
        if (st1 && st2 && st3) break;
        o1 = s1; o2 = s2; o3 = s3;
    }
    *s1o = s1; *s2o = s2; *s3o = s3;
}

/* One pair: v = the reference's K increment for strengths S1f, Sf. */
static float hko_compat_pair(float S1, float Sf, float y, float x, float vcut2, float ey,
                             hko_cabc k)
{
    float t1 = x * x;
    float xxyy = t1 + y * y;
    if (!(t1 < vcut2)) return 0.0f;
    if (xxyy >= 1.0e6f) return S1 / xxyy;
    if (xxyy >= 100.0f) {
        float t6 = t1 * 6.0f;
        float t2 = xxyy + 1.5f;
        float t3 = HKO_PIF * t2;
        float t4 = (t3 * (2.0f * t2 + xxyy) - 2.0f * t6) / (3.0f * xxyy * (t3 * t2 - t6));
        return S1 * t4;
    }
    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    float ex2 = expf(-x * x);
    if (x != 0.0f && y != 0.0f) hko_sigmabf(x, y, k.a, &s1, &s2, &s3);
    float xy = x * y;
    float cos2xy = cosf(2.0f * xy);
    float sinxy = sinf(xy);
    float u1 = ex2 * ey * cos2xy;
    float u2 = sinxy * ex2 * sinxy / y;
    float u3 = y * (-cos2xy * s1 + 0.5f * (s2 + s3));
    u1 += k.c * (u2 + u3);
    if (x == 0.0f) u1 = ey;
    if (y == 0.0f) u1 = ex2;
    return Sf * u1 * k.b;
}

/* The compat profile Re w(x, y)/sqrt(pi) for a unit line (S = 1, no cut):
 * the three region formulas as functions of (x, y) alone. */
float hko_compat_profile(float x, float y)
{
    hko_cabc k = hko_compat_consts();
    float S1 = (float)(1.0 * (double)y / M_PI);
    float ey = (float)hko_erfcx((double)y);
    return hko_compat_pair(S1, 1.0f, y, x, INFINITY, ey, k);
}

void hko_voigt_compat_points(int64_t nl, const double *nu, const double *S,
                             const double *ialphaD, const double *y,
                             const double *hw, double R,
                             double numin, double dnu, int64_t Nx, double kmin,
                             int64_t npts, const int64_t *jidx, double *K,
                             int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    hko_cabc k = hko_compat_consts();
    /* the FP32 cut test can admit a point a little beyond the FP64 window */
    double Rs = R + 2.0 * dnu + 1e-6 * R;
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < npts; ++p) {
        int64_t j = jidx ? jidx[p] : p;
        double xj = fma((double)j, dnu, numin);
        int64_t i0 = lower_bound(nu, nl, xj - Rs);
        int64_t i1 = lower_bound(nu, nl, xj + Rs);
        if (i1 < nl) ++i1;
        double sum = 0.0;
        for (int64_t i = i0; i < i1; ++i) {
            if (!(hw[i] > 0.0)) continue;
            double ia = ialphaD[i];
            float S1 = (float)(S[i] * y[i] / M_PI);
            float Sf = (float)S[i];
            float yf = (float)y[i];
            float va = (float)((numin - nu[i]) * ia);
            float vb = (float)(dnu * ia);
            float vcut2 = (float)(hw[i] * hw[i] * ia * ia);
            float x = fmaf((float)j, vb, va);
            if (!(x * x < vcut2)) continue;
            float ey = (float)hko_erfcx((double)yf);
            sum += (double)hko_compat_pair(S1, Sf, yf, x, vcut2, ey, k);
        }
        K[p] = kmin + sum;
    }
}

/* ------------------------------------------------------------------ */
/* Line profiles 1-4 (voigt.h:1235-1262 with S1 of S3_kernel,          */
/* voigt.h:458-470) and the plinth (Plinth_kernel, voigt.h:1091-1164):  */
/* per pair, in Doppler units x = (x_j - nu) / alpha_D:                 */
/*   1 Voigt        S Re w(x, y) / sqrt(pi)                             */
/*   2 Lorentz      (S y / pi) / (x^2 + y^2)                            */
/*   3 Doppler      (S / sqrt(pi)) exp(-x^2)                            */
/*   4 binned Gauss S (1/vb) (erf(x + vb/2) - erf(x - vb/2)) / 2,       */
/*                  vb = dnu / alpha_D                                  */
/* minus, with remove_plinth, the profile 1-3 at the cut x_c = hw /     */
/* alpha_D (the plinth of the truncated line).                          */
/* ------------------------------------------------------------------ */
double hko_profile_value(int profile, double S, double x, double y, double vb)
{
    double re, im;
    switch (profile) {
    case 1: hko_faddeeva(x, y, &re, &im); return S * re * HKO_ISQRTPI;
    case 2: return S * y / M_PI / (x * x + y * y);
    case 3: return S * HKO_ISQRTPI * exp(-x * x);
    default: {
        /* (erf(x + h) - erf(x - h)) / 2 from erfc on the side away from the
           centre (the same value, without the cancellation in the wings) */
        double h = 0.5 * vb;
        double d = x >= 0.0 ? erfc(x - h) - erfc(x + h) : erfc(-x - h) - erfc(-x + h);
        return S / vb * 0.5 * d;
    }
    }
}

/* Perrin & Hartmann (1989) sub-Lorentzian factor chi(|x - nu|) (sLChi,
 * voigt.h:47-72): 1 below sigma1, then exp(-B1 (d - sigma1)), continued with
 * slopes B2 above sigma2 and B3 above sigma3.  chi = {sigma1..3, B1..3,
 * range_lo, range_hi}. */
double hko_chi(const double *chi, double d)
{
    const double s1 = chi[0], s2 = chi[1], s3 = chi[2], B1 = chi[3], B2 = chi[4], B3 = chi[5];
    if (d < s1) return 1.0;
    if (d < s2) return exp(-B1 * (d - s1));
    if (d < s3) return exp(-B1 * (s2 - s1) - B2 * (d - s2));
    return exp(-B1 * (s2 - s1) - B2 * (s3 - s2) - B3 * (d - s3));
}

void hko_profile_points(int64_t nl, const double *nu, const double *S,
                        const double *ialphaD, const double *y,
                        const double *hw, double R,
                        double numin, double dnu, int64_t Nx, double kmin,
                        int profile, int remove_plinth, const double *chi, double *K,
                        int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < Nx; ++j) {
        double xj = fma((double)j, dnu, numin);
        int64_t i0 = lower_bound(nu, nl, xj - R);
        int64_t i1 = lower_bound(nu, nl, xj + R);
        if (i1 < nl) ++i1;
        double sum = kmin, comp = 0.0;
        for (int64_t i = i0; i < i1; ++i) {
            double d = xj - nu[i];
            if (!(d < hw[i] && d > -hw[i])) continue;
            double vb = dnu * ialphaD[i];
            double t = hko_profile_value(profile, S[i], d * ialphaD[i], y[i], vb);
            /* sub-Lorentzian wings: chi(|x - nu|) x / nu inside the range
               (voigt.h:1345-1350) */
            if (chi && xj > chi[6] && xj < chi[7]) t *= hko_chi(chi, fabs(d)) * xj / nu[i];
            if (remove_plinth)
                t -= hko_profile_value(profile, S[i], hw[i] * ialphaD[i], y[i], vb);
            double s2 = sum + t;
            if (fabs(sum) >= fabs(t)) comp += (sum - s2) + t;
            else comp += (t - s2) + sum;
            sum = s2;
        }
        K[j] = sum + comp;
    }
}

/* ------------------------------------------------------------------ */
/* dRe w/dx = Re w'(z) and dRe w/dy = -Im w'(z) for |z|^2 >= 1e4 from   */
/* the asymptotic series w'(z) = -(i/sqrt(pi)) u sum_k (2k+1)!!/2^k u^k, */
/* u = 1/z^2 (six terms: the next is below 1e-20 relative).  Forming    */
/* them from w' = -2zw + 2i/sqrt(pi) instead cancels to ~eps |z|^2.     */
/* ------------------------------------------------------------------ */
static void drew_far(double x, double y, double r2, double *ddx, double *ddy)
{
    double inv = 1.0 / r2;
    double ur = (x - y) * (x + y) * inv * inv, ui = -2.0 * (x * inv) * (y * inv);
    static const double ck[6] = {1.0, 1.5, 3.75, 13.125, 59.0625, 324.84375};
    double pr = ck[5], pi = 0.0;
    for (int k = 4; k >= 0; --k) {
        double tr = pr * ur - pi * ui, ti = pr * ui + pi * ur;
        pr = tr + ck[k]; pi = ti;
    }
    double sr = ur * pr - ui * pi, si = ur * pi + ui * pr;
    *ddx = si * HKO_ISQRTPI;
    *ddy = sr * HKO_ISQRTPI;
}

/* ------------------------------------------------------------------ */
/* Analytic gradient of L = sum_j gK_j K_j                             */
/* with respect to T, P (gTP[0], gTP[1]) and per-line S0, vy0.         */
/* The partition function Q is held fixed (constant ratio).            */
/* ------------------------------------------------------------------ */
void hko_voigt_grad(int64_t nl,
                    const double *nu0, const double *S0, const double *EL,
                    const double *A, const double *delta, const double *vy0,
                    const double *nexp, const int32_t *iso,
                    const double *iso_mass, const double *iso_ab,
                    const double *iso_sscale, const double *iso_Q,
                    double T, double P, int cutMode, double cut, int form,
                    double numin, double dnu, int64_t Nx, const double *gK,
                    double *gS0, double *gvy0, double *gTP, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double c = HKO_H * HKO_C / HKO_KB;
    long double gT = 0.0L, gP = 0.0L;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:gT, gP)
    for (int64_t i = 0; i < nl; ++i) {
        int k = iso ? iso[i] : 0;
        double ialphaD_h = HKO_C * sqrt(iso_mass[k] / (2.0 * HKO_KB * T));
        double F0 = form == 2 ? iso_ab[k] * iso_sscale[k]
                              : 1.0 / iso_Q[k] * iso_ab[k] * iso_sscale[k];
        double dlt = form == 0 ? delta[i] : 0.0;
        double nu = fma(dlt, P, nu0[i]);
        if (nu == 0.0) { gS0[i] = 0.0; gvy0[i] = 0.0; continue; }
        double ia = ialphaD_h / nu;
        double E = EL[i];
        double eT = exp(-c * nu / T), e0 = exp(-c * nu / HKO_T0);
        double B = form == 0 ? exp(-E * c / T + E * c / HKO_T0) : (form == 1 ? exp(-c / T * E) : 1.0);
        double Rr = form == 0 ? (1.0 - eT) / (1.0 - e0) : (form == 1 ? 1.0 - eT : 1.0);
        double F = F0 * B * Rr * ia;            /* S = S0 * F */
        double S = S0[i] * F;
        double pw = pow(HKO_T0 / T, nexp[i]);
        double Pscale = form == 0 ? 1.0 / HKO_POATM : 1.0 / HKO_POBAR;
        double aLp = vy0[i] * P * Pscale * pw;  /* pressure part of alpha_L */
        double aL = aLp + (form == 2 ? 0.0 : A[i] / (4.0 * M_PI * HKO_C));
        double y = aL * ia;
        double hw = cutMode == 1 ? cut * aL : (cutMode == 2 ? cut / ia : cut);
        /* log-derivatives */
        double dlnR_dT = form == 2 ? 0.0 : -eT * (c * nu / (T * T)) / (1.0 - eT);
        double dlnR_dnu = eT * (c / T) / (1.0 - eT) - e0 * (c / HKO_T0) / (1.0 - e0);
        double dlnS_dT = (form == 2 ? 0.0 : c * E / (T * T)) + dlnR_dT - 0.5 / T;
        double dlnS_dP = dlt * (dlnR_dnu - 1.0 / nu);
        double dia_dT = -0.5 * ia / T, dia_dP = -ia * dlt / nu;
        double daL_dT = -nexp[i] / T * aLp, daL_dP = vy0[i] * Pscale * pw;
        double dy_dT = daL_dT * ia + aL * dia_dT;
        double dy_dP = daL_dP * ia + aL * dia_dP;
        double dy_dvy = P * Pscale * pw * ia;
        int64_t jlo, jhi;
        hko_window(nu, hw, numin, dnu, Nx, &jlo, &jhi);
        double sre = 0.0, sdy = 0.0, sT = 0.0, sP = 0.0;
        for (int64_t j = jlo; j < jhi; ++j) {
            double g = gK[j];
            if (g == 0.0) continue;
            double d = fma((double)j, dnu, numin) - nu;
            double x = d * ia;
            double re, im, dRe_dx, dRe_dy;
            hko_faddeeva(x, y, &re, &im);
            double r2 = x * x + y * y;
            if (r2 >= 1e4) {
                drew_far(x, y, r2, &dRe_dx, &dRe_dy);
            } else {
                dRe_dx = -2.0 * (x * re - y * im);
                dRe_dy = 2.0 * (x * im + y * re) - 2.0 * HKO_ISQRTPI;
            }
            double dx_dT = d * dia_dT;
            double dx_dP = -dlt * ia + d * dia_dP;
            sre += g * re;
            sdy += g * dRe_dy;
            sT += g * (re * dlnS_dT + dRe_dx * dx_dT + dRe_dy * dy_dT);
            sP += g * (re * dlnS_dP + dRe_dx * dx_dP + dRe_dy * dy_dP);
        }
        gS0[i] = F * sre * HKO_ISQRTPI;
        gvy0[i] = S * sdy * dy_dvy * HKO_ISQRTPI;
        gT += (long double)(S * sT * HKO_ISQRTPI);
        gP += (long double)(S * sP * HKO_ISQRTPI);
    }
    gTP[0] = (double)gT;
    gTP[1] = (double)gP;
}

/* ------------------------------------------------------------------ */
/* Bins (binKey_kernel semantics) and per-bin sort                     */
/* ------------------------------------------------------------------ */
static int bin_key(double xj, const double *bnd, int nbins, double numax)
{
    int key = 0;
    for (int i = 0; i < nbins; ++i)
        if (bnd[i] <= xj && xj < bnd[i + 1]) { key = i; break; }
    if (xj >= numax) key = nbins;
    return key;
}

/* seg[b] = first grid index whose bin key is >= b, b = 0..nbins+1 */
void hko_bin_segments(double numin, double numax, double dnu, int64_t Nx,
                      int nbins, int64_t *seg)
{
    double *bnd = (double *)malloc((nbins + 1) * sizeof(double));
    for (int i = 0; i < nbins; ++i)
        bnd[i] = numin + i * (numax - numin) / ((double)nbins);   /* heliosk.cu:438-441 */
    bnd[nbins] = numax;
    for (int b = 0; b <= nbins + 1; ++b) {
        int64_t lo = 0, hi = Nx;      /* keys are nondecreasing in j */
        while (lo < hi) {
            int64_t m = (lo + hi) / 2;
            if (bin_key(fma((double)m, dnu, numin), bnd, nbins, numax) < b) lo = m + 1;
            else hi = m;
        }
        seg[b] = lo;
    }
    free(bnd);
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

void hko_sort_segments(double *K, int nseg, const int64_t *seg)
{
    #pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nseg; ++b)
        qsort(K + seg[b], (size_t)(seg[b + 1] - seg[b]), sizeof(double), cmp_double);
}

/* ------------------------------------------------------------------ */
/* Vandermonde + Householder QR (resample.h:14-37, 73-227)             */
/* V is column-major Nxb x nC.                                         */
/* ------------------------------------------------------------------ */
void hko_qr(int NL, int NC, double *V, double *C, double *D)
{
    double nl = (double)NL;
    for (int id = 0; id < NL; ++id) {
        double x = id * 2.0 / (nl - 1.0) - 1.0;
        V[id] = 1.0;
        if (NC > 1) V[id + (size_t)NL] = x;
        double t = x, t0 = 1.0;
        for (int i = 2; i < NC; ++i) {
            double tt = t;
            t = fma(2.0 * x, t, -t0);
            t0 = tt;
            V[id + (size_t)NL * i] = t;
        }
    }
    for (int i = 0; i < NC; ++i) {
        double *vi = V + (size_t)NL * i;
        double scale = 0.0;
        for (int k = i; k < NL; ++k) scale = fmax(scale, fabs(vi[k]));
        double ss = 0.0;
        for (int k = i; k < NL; ++k) { vi[k] /= scale; ss += vi[k] * vi[k]; }
        double s = sqrt(ss), v = vi[i];
        double sigma = v < 0.0 ? -s : s;
        vi[i] += sigma;
        C[i] = sigma * (v + sigma);
        D[i] = -scale * sigma;
        for (int j = i + 1; j < NC; ++j) {
            double *vj = V + (size_t)NL * j;
            double dot = 0.0;
            for (int k = i; k < NL; ++k) dot += vi[k] * vj[k];
            double tq = dot / C[i];
            for (int k = i; k < NL; ++k) vj[k] -= tq * vi[k];
        }
    }
}

/* ------------------------------------------------------------------ */
/* k-table: findCut, rescale(+1), copyK2, lnK, leastSquare             */
/* K (length Nx = nbins*Nxb) is the per-bin sorted cross-section; it is */
/* overwritten with ln K (as the reference does with K_d).             */
/* coeff[b*nC + c], Nxmin[b].                                          */
/* ------------------------------------------------------------------ */
void hko_ktable(double *K, int64_t Nx, int Nxb, int nbins, int nC,
                double kmin, double unitScale, double *coeff, int *Nxmin)
{
    /* findCut_kernel, resample.h:362-388 (rule order = statement order) */
    for (int b = 0; b < nbins; ++b) Nxmin[b] = 0;
    int *r1 = (int *)malloc(nbins * sizeof(int));
    int *r2 = (int *)malloc(nbins * sizeof(int));
    int *r3 = (int *)malloc(nbins * sizeof(int));
    for (int b = 0; b < nbins; ++b) { r1[b] = -1; r2[b] = -1; r3[b] = -1; }
    for (int64_t id = 0; id < Nx - 1; ++id) {
        int ib = (int)(id / Nxb);
        if (ib >= nbins) continue;
        double k0 = K[id], k1 = K[id + 1];
        if (k0 <= kmin && k1 > kmin) r1[ib] = (int)(id - (int64_t)ib * Nxb + 1);
        if (k0 <= kmin && id % Nxb == Nxb - 1) r2[ib] = Nxb;
        if (k0 <= kmin && id == Nx - 3) r3[ib] = Nxb;
    }
    for (int b = 0; b < nbins; ++b) {
        if (r1[b] >= 0) Nxmin[b] = r1[b];
        if (r2[b] >= 0) Nxmin[b] = r2[b];
        if (r3[b] >= 0) Nxmin[b] = r3[b];
    }
    free(r1); free(r2); free(r3);

    /* rescale_kernel f=+1 and copyK2_kernel, resample.h:398-466 */
    double *K2 = (double *)malloc((size_t)Nxb * sizeof(double));
    for (int b = 0; b < nbins; ++b) {
        int n = Nxmin[b];
        if (n <= 0) continue;
        double *Kb = K + (size_t)b * Nxb;
        double s = 1.0 - n / ((double)(Nxb - 1));
        for (int k = 0; k < Nxb; ++k) {
            double ii = fma(s, (double)k, (double)n);
            if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
            int il = (int)(ii / ((double)Nxb) * Nxb);
            double Kl = Kb[il], Kr = Kb[il + 1];
            K2[k] = fma(Kr - Kl, ii - il, Kl);
        }
        memcpy(Kb, K2, (size_t)Nxb * sizeof(double));
    }
    free(K2);

    /* lnK_kernel, resample.h:472-481 */
    for (int64_t id = 0; id < Nx; ++id) K[id] = log(K[id] * unitScale);

    /* leastSquare_kernel, resample.h:243-353 */
    double *V = (double *)malloc((size_t)Nxb * nC * sizeof(double));
    double *C = (double *)malloc(nC * sizeof(double));
    double *D = (double *)malloc(nC * sizeof(double));
    hko_qr(Nxb, nC, V, C, D);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbins; ++b) {
        double *bb = K + (size_t)b * Nxb;
        for (int j = 0; j < nC - 1; ++j) {
            const double *vj = V + (size_t)Nxb * j;
            double dot = 0.0;
            for (int k = j; k < Nxb; ++k) dot += vj[k] * bb[k];
            double t = dot / C[j];
            for (int k = j; k < Nxb; ++k) bb[k] -= t * vj[k];
        }
        bb[nC - 1] /= D[nC - 1];
        for (int i = nC - 2; i >= 0; --i) {
            double a = 0.0;
            for (int j = i + 1; j < nC; ++j) a += V[i + (size_t)Nxb * j] * bb[j];
            bb[i] = (bb[i] - a) / D[i];
        }
        for (int c = 0; c < nC; ++c) coeff[(size_t)b * nC + c] = bb[c];
    }
    free(V); free(C); free(D);
}

/* ------------------------------------------------------------------ */
/* leastSquare_kernel alone (resample.h:243-353) on an arbitrary b of  */
/* length Nxb: NC-1 reflections then back substitution; c[0..nC).      */
/* ------------------------------------------------------------------ */
void hko_lsq(int Nxb, int nC, const double *b_in, double *c)
{
    double *V = (double *)malloc((size_t)Nxb * nC * sizeof(double));
    double *C = (double *)malloc(nC * sizeof(double));
    double *D = (double *)malloc(nC * sizeof(double));
    double *bb = (double *)malloc((size_t)Nxb * sizeof(double));
    hko_qr(Nxb, nC, V, C, D);
    memcpy(bb, b_in, (size_t)Nxb * sizeof(double));
    for (int j = 0; j < nC - 1; ++j) {
        const double *vj = V + (size_t)Nxb * j;
        double dot = 0.0;
        for (int k = j; k < Nxb; ++k) dot += vj[k] * bb[k];
        double t = dot / C[j];
        for (int k = j; k < Nxb; ++k) bb[k] -= t * vj[k];
    }
    bb[nC - 1] /= D[nC - 1];
    for (int i = nC - 2; i >= 0; --i) {
        double a = 0.0;
        for (int j = i + 1; j < nC; ++j) a += V[i + (size_t)Nxb * j] * bb[j];
        bb[i] = (bb[i] - a) / D[i];
    }
    memcpy(c, bb, (size_t)nC * sizeof(double));
    free(V); free(C); free(D); free(bb);
}

/* ------------------------------------------------------------------ */
/* Reconstruction K(y): expfx_kernel + rescale(f=-1) (resample.h:497-524, */
/* 398-437); Ky has nbins*Nxb entries.                                 */
/* ------------------------------------------------------------------ */
void hko_reconstruct(const double *coeff, const int *Nxmin, int Nxb,
                     int nbins, int nC, double kmin, double unitScale,
                     double *Ky)
{
    double *F = (double *)malloc((size_t)Nxb * sizeof(double));
    for (int b = 0; b < nbins; ++b) {
        const double *cb = coeff + (size_t)b * nC;
        for (int k = 0; k < Nxb; ++k) {
            double x = -1.0 + k * 2.0 / ((double)Nxb);
            double d1 = 0.0, d2 = 0.0;
            for (int i = nC - 1; i >= 1; --i) {
                double t = d1;
                d1 = fma(2.0 * x, d1, -d2) + cb[i];
                d2 = t;
            }
            double f = fma(x, d1, -d2) + 0.5 * cb[0] + 0.5 * cb[0];
            F[k] = exp(f) / unitScale;
        }
        double *out = Ky + (size_t)b * Nxb;
        int n = Nxmin[b];
        if (n > 0) {
            for (int k = 0; k < Nxb; ++k) {
                double Ki = kmin;
                if (k >= n) {
                    double ii = (k - n) / ((double)(Nxb - n - 1)) * (Nxb - 1);
                    if (ii >= Nxb - 1) ii = 0.999999 * (Nxb - 1);
                    /* Nxmin == Nxb-1 gives 0/0; the device conversion of NaN is 0 */
                    int il = ii == ii ? (int)(ii / ((double)Nxb) * Nxb) : 0;
                    double Kl = F[il], Kr = F[il + 1];
                    Ki = fma(Kr - Kl, ii - il, Kl);
                }
                out[k] = Ki;
            }
        } else {
            memcpy(out, F, (size_t)Nxb * sizeof(double));
        }
    }
    free(F);
}


/* ---------------------------------------------------------------- means */
/* resample.h:635-664 (Mean_kernel) and 669-771 (IntegrateMean_kernel). */
/* Neumaier-compensated running sum: the oracle's sums carry O(eps) error
 * whatever the order and dynamic range of the terms. */
static void hko_nsum(double *s, double *c, double v) {
    double t = *s + v;
    if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
    else *c += (v - t) + *s;
    *s = t;
}

void hko_means(const double *K, double numin, double dnu, int64_t Nx, double T,
               double *sums) {
    double s[4] = {0.0, 0.0, 0.0, 0.0}, cmp[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = 0; j < Nx; ++j) {
        double x = fma((double)j, dnu, numin);
        double t1 = 2.0 * HKO_H * x * x * x * HKO_C * HKO_C;
        double t2 = HKO_H * x * HKO_C / (HKO_KB * T);
        double e = exp(t2);
        double B = t1 / (e - 1.0);
        double dB = t1 * t2 / (T * (e - 2.0 + 1.0 / e));
        if (x == 0.0) { B = 0.0; dB = 0.0; }
        hko_nsum(&s[0], &cmp[0], B * K[j]);
        if (K[j] != 0.0) hko_nsum(&s[1], &cmp[1], dB / K[j]);
        hko_nsum(&s[2], &cmp[2], B);
        hko_nsum(&s[3], &cmp[3], dB);
    }
    for (int q = 0; q < 4; ++q) sums[q] = s[q] + cmp[q];
}

/* resample.h:700-716 (IntegrateMean_kernel, useIndividualX = 1): trapezoid
 * rule on explicit points. */
void hko_means_x(const double *K, const double *x, int64_t Nx, double T, double *sums) {
    double s[4] = {0.0, 0.0, 0.0, 0.0}, cmp[4] = {0.0, 0.0, 0.0, 0.0};
    double Bp = 0.0, dBp = 0.0;
    for (int64_t j = 0; j < Nx; ++j) {
        double xx = x[j];
        double t1 = 2.0 * HKO_H * xx * xx * xx * HKO_C * HKO_C;
        double t2 = HKO_H * xx * HKO_C / (HKO_KB * T);
        double e = exp(t2);
        double B = t1 / (e - 1.0);
        double dB = t1 * t2 / (T * (e - 2.0 + 1.0 / e));
        if (xx == 0.0) { B = 0.0; dB = 0.0; }
        if (j > 0) {
            double w = x[j] - x[j - 1], k0 = K[j - 1], k1 = K[j];
            hko_nsum(&s[0], &cmp[0], 0.5 * (Bp * k0 + B * k1) * w);
            if (k0 != 0.0 && k1 != 0.0) hko_nsum(&s[1], &cmp[1], 0.5 * (dBp / k0 + dB / k1) * w);
            hko_nsum(&s[2], &cmp[2], 0.5 * (Bp + B) * w);
            hko_nsum(&s[3], &cmp[3], 0.5 * (dBp + dB) * w);
        }
        Bp = B; dBp = dB;
    }
    for (int q = 0; q < 4; ++q) sums[q] = s[q] + cmp[q];
}

/* --------------------------------------------------------- transmission */
/* resample.h:541-614; end weights resample.h:618-626. */
void hko_transmission(const double *Ks, int nbins, int Nxb, const int32_t *nxmin,
                      double kmin, int nTr, double dTr, double *Tr) {
    static const double w[3] = {-5.0 / 8.0, 1.0 / 6.0, -1.0 / 24.0};
    for (int b = 0; b < nbins; ++b) {
        const double *k = Ks + (int64_t)b * Nxb;
        double nmin = nxmin ? (double)nxmin[b] : 0.0;
        for (int t = 0; t < nTr; ++t) {
            double m = exp((double)(t - nTr / 2) * dTr);
            double a = 0.0, ca = 0.0;
            for (int i = 0; i < Nxb; ++i) hko_nsum(&a, &ca, exp(-k[i] * m));
            for (int i = 0; i < 3; ++i) {
                hko_nsum(&a, &ca, w[i] * exp(-k[i] * m));
                hko_nsum(&a, &ca, w[2 - i] * exp(-k[Nxb - 1 - i] * m));
            }
            a += ca;
            Tr[(int64_t)b * nTr + t] = a / ((double)(Nxb - 1)) * (Nxb - nmin) / ((double)Nxb)
                                       + exp(-kmin * m) * nmin / ((double)Nxb);
        }
    }
}
