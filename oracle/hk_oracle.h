/*
 * hk_oracle.h -- FP64 host oracle for the HELIOS-K line-by-line path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * heliosk driver) links, loads or calls this code.  It is used by tests/,
 * by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline
 * leg as the timed CPU restatement.
 *
 * Every function restates a part of the reference (exoclime/HELIOS-K 2.08)
 * and cites the file:line it follows.  How each function is pinned
 * (golden vectors, fixtures, closed forms) is listed in the header of
 * hk_oracle.c.
 */
#ifndef HK_ORACLE_H
#define HK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants, restated from define.h:36-43. */
#define HKO_T0     296.0
#define HKO_POBAR  0.986923
#define HKO_POATM  1.0
#define HKO_KB     1.3806489e-16
#define HKO_H      6.62606957e-27
#define HKO_C      2.99792458e10
#define HKO_NA     6.0221412927e23

/* ---- Faddeeva function w(z) = exp(-z^2) erfc(-iz), z = x + iy, y >= 0 ---- */
double hko_erfcx(double y);
void   hko_faddeeva(double x, double y, double *re, double *im);
void   hko_w916(double x, double y, double tol, double *re, double *im);
void   hko_wcf(double x, double y, int depth, double *re, double *im);
void   hko_faddeeva_batch(int64_t n, const double *x, const double *y,
                          double *re, double *im);

/* ---- line parameters at (T, P) (readFile + S2_kernel semantics) ---- */
void hko_line_params(int64_t nl,
                     const double *nu0, const double *S0, const double *EL,
                     const double *A, const double *delta, const double *vy0,
                     const double *nexp, const int32_t *iso,
                     const double *iso_mass, const double *iso_ab,
                     const double *iso_sscale, const double *iso_Q,
                     double T, double P, int cutMode, double cut, int form,
                     double *nu, double *S, double *ialphaD, double *y,
                     double *hw);

/* ---- integer grid window of one line ---- */
void hko_window(double nu, double hw, double numin, double dnu, int64_t Nx,
                int64_t *jlo, int64_t *jhi);

/* ---- Voigt sum K_j = kmin + sum_i S_i Re w(x_ij, y_i) / sqrt(pi) ---- */
void hko_voigt_points(int64_t nl, const double *nu, const double *S,
                      const double *ialphaD, const double *y,
                      const double *hw, double search_radius,
                      double numin, double dnu, int64_t Nx, double kmin,
                      int64_t npts, const int64_t *jidx, double *K,
                      int nthreads);

/* ---- the reference's own FP32 Voigt arithmetic (compat form) ---- */
float hko_compat_profile(float x, float y);
void hko_voigt_compat_points(int64_t nl, const double *nu, const double *S,
                             const double *ialphaD, const double *y,
                             const double *hw, double search_radius,
                             double numin, double dnu, int64_t Nx, double kmin,
                             int64_t npts, const int64_t *jidx, double *K,
                             int nthreads);

/* ---- profiles 1-4 with optional plinth removal (all grid points) ---- */
double hko_profile_value(int profile, double S, double x, double y, double vb);
double hko_chi(const double *chi, double d);
void hko_profile_points(int64_t nl, const double *nu, const double *S,
                        const double *ialphaD, const double *y,
                        const double *hw, double search_radius,
                        double numin, double dnu, int64_t Nx, double kmin,
                        int profile, int remove_plinth, const double *chi, double *K,
                        int nthreads);

/* ---- gradient of L = sum_j gK_j K_j w.r.t. T, P, S0_i, vy0_i ---- */
void hko_voigt_grad(int64_t nl,
                    const double *nu0, const double *S0, const double *EL,
                    const double *A, const double *delta, const double *vy0,
                    const double *nexp, const int32_t *iso,
                    const double *iso_mass, const double *iso_ab,
                    const double *iso_sscale, const double *iso_Q,
                    double T, double P, int cutMode, double cut, int form,
                    double numin, double dnu, int64_t Nx, const double *gK,
                    double *gS0, double *gvy0, double *gTP, int nthreads);

/* ---- k-table (resample.h) ---- */
void hko_bin_segments(double numin, double numax, double dnu, int64_t Nx,
                      int nbins, int64_t *seg);
void hko_sort_segments(double *K, int nseg, const int64_t *seg);
void hko_qr(int Nxb, int nC, double *V, double *C, double *D);
void hko_ktable(double *K, int64_t Nx, int Nxb, int nbins, int nC,
                double kmin, double unitScale, double *coeff, int *Nxmin);
void hko_lsq(int Nxb, int nC, const double *b, double *c);
void hko_reconstruct(const double *coeff, const int *Nxmin, int Nxb,
                     int nbins, int nC, double kmin, double unitScale,
                     double *Ky);

/* ---- consumers of K (resample.h) ---- */
/* Planck / Rosseland mean sums [sum B K, sum_{K!=0} dB/K, sum B, sum dB] on
 * x_j = fma(j, dnu, numin) (resample.h:635-771, plain sums). */
void hko_means(const double *K, double numin, double dnu, int64_t Nx, double T,
               double *sums);
/* The same sums on explicit points x, trapezoid rule (resample.h:700-716). */
void hko_means_x(const double *K, const double *x, int64_t Nx, double T, double *sums);
/* Transmission of each of nbins sorted bins of Nxb values
 * (resample.h:541-614); nxmin may be NULL (all 0). */
void hko_transmission(const double *Ks, int nbins, int Nxb, const int32_t *nxmin,
                      double kmin, int nTr, double dTr, double *Tr);

#ifdef __cplusplus
}
#endif
#endif
