// hk_device.h -- device helpers shared by the HIP kernels of the line-by-line
// path: physical constants, the FP64 cut window, and the Faddeeva pieces of
// the FP64 and FP32 Voigt forms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hk {

// Physical constants of define.h:36-43.
constexpr double kT0 = 296.0;
constexpr double kPOatm = 1.0;
constexpr double kPObar = 0.986923;          // define.h:37
constexpr double kKB = 1.3806489e-16;
constexpr double kH = 6.62606957e-27;
constexpr double kC = 2.99792458e10;

constexpr double kPi = 3.14159265358979323846;
constexpr double kISqrtPi = 0.56418958354775628695;   // 1/sqrt(pi)

constexpr int kBlockLines = 256;   // lines per blk_max entry

// --------------------------------------------------------------------------
// Cut window of one line on the grid x_j = fma(j, dnu, numin):
// j in [lo, hi)  <=>  |x_j - nu| < hw, evaluated in FP64.  The estimate is
// corrected by stepping until both predicates flip, so the result is the
// exact set defined by the predicates whatever the estimate.
// --------------------------------------------------------------------------
__device__ __forceinline__ int64_t clamp_index(double e, int64_t nx) {
    if (!(e > 0.0)) return 0;
    if (e >= (double)nx) return nx;
    return (int64_t)floor(e);
}

__device__ __forceinline__ void cut_window(double nu, double hw, double numin, double dnu,
                                           int64_t nx, int64_t &lo, int64_t &hi) {
    if (!(hw > 0.0)) { lo = 0; hi = 0; return; }
    int64_t l = clamp_index((nu - hw - numin) / dnu, nx);
    while (l > 0 && fma((double)(l - 1), dnu, numin) - nu > -hw) --l;
    while (l < nx && !(fma((double)l, dnu, numin) - nu > -hw)) ++l;
    int64_t h = clamp_index((nu + hw - numin) / dnu, nx);
    while (h > 0 && !(fma((double)(h - 1), dnu, numin) - nu < hw)) --h;
    while (h < nx && fma((double)h, dnu, numin) - nu < hw) ++h;
    if (h < l) h = l;
    lo = l; hi = h;
}

// --------------------------------------------------------------------------
// Fast reciprocals.  v_rcp_f32 is accurate to 1 ulp; the FP64 one is refined
// by two Newton steps to full double precision.
// --------------------------------------------------------------------------
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// --------------------------------------------------------------------------
// Voigt regions.  r2 = x^2 + y^2.
//  FP64: core r2 < 100 (Algorithm 916), 100..1e4 Gauss-Hermite 8 nodes,
//        1e4..1e6 Gauss-Hermite 4 nodes, >= 1e6 three-term asymptotic.
// Gauss-Hermite: Re w = (y/pi) sum_k w_k / ((x - t_k)^2 + y^2); each
// symmetric pair folds to 2 w u / D with u = r2 + t^2, D = u^2 - 4 t^2 x^2.
// --------------------------------------------------------------------------
constexpr double kRB64 = 100.0, kRM64 = 1.0e4, kRF64 = 1.0e6;
// FP32: core r2 < 25 (Algorithm 916), 25..100 8-node Gauss-Hermite plus the
// Doppler term e^{-x^2} for near-Doppler lines (y < kY0), 100..4096 the
// four-term asymptotic series, >= 4096 the two-term one (far4_32, far32).
constexpr float kRB32 = 25.0f, kRM32 = 100.0f, kRF32 = 4096.0f;
constexpr float kY0 = 0.1f;

// Hermite nodes/weights (positive half), n = 4 and n = 8.
constexpr double kGH4t1 = 0.5246476232752904, kGH4t2 = 1.6506801238857842;
constexpr double kGH4w1 = 0.804914090005513, kGH4w2 = 0.08131283544724505;
constexpr double kGH8t1 = 0.3811869902073221, kGH8t2 = 1.1571937124467802;
constexpr double kGH8t3 = 1.9816567566958427, kGH8t4 = 2.930637420257244;
constexpr double kGH8w1 = 0.6611470125582415, kGH8w2 = 0.2078023258148917;
constexpr double kGH8w3 = 0.01707798300741341, kGH8w4 = 0.0001996040722113662;

// Sum_k w_k u_k / D_k for the 4-node rule, FP64 (multiply by 2y/pi).
__device__ __forceinline__ double gh4_sum64(double x2, double r2) {
    const double a1 = kGH4t1 * kGH4t1, a2 = kGH4t2 * kGH4t2;
    double u1 = r2 + a1, u2 = r2 + a2;
    double D1 = fma(u1, u1, -4.0 * a1 * x2);
    double D2 = fma(u2, u2, -4.0 * a2 * x2);
    double num = fma(kGH4w1 * u1, D2, kGH4w2 * u2 * D1);
    return num * rcp64(D1 * D2);
}

__device__ __forceinline__ double gh8_sum64(double x2, double r2) {
    const double a1 = kGH8t1 * kGH8t1, a2 = kGH8t2 * kGH8t2;
    const double a3 = kGH8t3 * kGH8t3, a4 = kGH8t4 * kGH8t4;
    double u1 = r2 + a1, u2 = r2 + a2, u3 = r2 + a3, u4 = r2 + a4;
    double D1 = fma(u1, u1, -4.0 * a1 * x2);
    double D2 = fma(u2, u2, -4.0 * a2 * x2);
    double D3 = fma(u3, u3, -4.0 * a3 * x2);
    double D4 = fma(u4, u4, -4.0 * a4 * x2);
    double D12 = D1 * D2, D34 = D3 * D4;
    double n12 = fma(kGH8w1 * u1, D2, kGH8w2 * u2 * D1);
    double n34 = fma(kGH8w3 * u3, D4, kGH8w4 * u4 * D3);
    double num = fma(n12, D34, n34 * D12);
    return num * rcp64(D12 * D34);
}

__device__ __forceinline__ float gh8_sum32(float x2, float r2) {
    const float a1 = (float)(kGH8t1 * kGH8t1), a2 = (float)(kGH8t2 * kGH8t2);
    const float a3 = (float)(kGH8t3 * kGH8t3), a4 = (float)(kGH8t4 * kGH8t4);
    const float u1 = r2 + a1, u2 = r2 + a2, u3 = r2 + a3, u4 = r2 + a4;
    const float D1 = fmaf(u1, u1, -4.0f * a1 * x2);
    const float D2 = fmaf(u2, u2, -4.0f * a2 * x2);
    const float D3 = fmaf(u3, u3, -4.0f * a3 * x2);
    const float D4 = fmaf(u4, u4, -4.0f * a4 * x2);
    const float D12 = D1 * D2, D34 = D3 * D4;
    const float n12 = fmaf((float)kGH8w1 * u1, D2, (float)kGH8w2 * u2 * D1);
    const float n34 = fmaf((float)kGH8w3 * u3, D4, (float)kGH8w4 * u4 * D3);
    return fmaf(n12, D34, n34 * D12) * rcp32(D12 * D34);
}

// --------------------------------------------------------------------------
// Algorithm 916 (Zaghloul & Ali, ACM TOMS 38(2), 2011), real part,
// x >= 0, y >= 0, for the core region.  The reference evaluates the same
// series for its region C (voigt.h:217-268, 1659-1686):
//   Re w = e^{-x^2} erfcx(y) cos 2xy + (2a/pi) [sin^2(xy)/y e^{-x^2}
//          + y (-cos(2xy) s1 + (s2 + s3)/2)]
//   s1 = sum_n e^{-(a^2n^2 + x^2)} d_n,  s2 = sum_n e^{-(an+x)^2} d_n,
//   s3 = sum_n e^{-(an-x)^2} d_n,        d_n = 1/(a^2 n^2 + y^2),
// a = pi/sqrt(-ln(tol/2)).  y is constant along a line, so erfcx(y) and
// q1 = sum_n e^{-a^2 n^2} d_n (s1 = e^{-x^2} q1) are computed once per line
// and tile (core_line_consts*).  The series keep the terms whose exponent is
// above -(a N)^2: FP64 tol 1e-15, N = 13 (e^-47); FP32 tol 1e-7, N = 5
// (e^-14.7; the first omitted term is below e^-21 = 6.7e-10 of the leading
// one).  Both cores build the exponentials of each series from two exp()
// calls by the ratio recurrences e^{-(a(n+1)+x)^2} = e^{-(an+x)^2} e^{-2ax}
// e^{-a^2(2n+1)}.
// --------------------------------------------------------------------------
constexpr double kA64 = 0.529275382982092;          // tol 1e-15
constexpr int kN64 = 13;
constexpr double kEA64 = 0.755683658862919;         // e^{-a^2}
constexpr double kE2A64 = 0.5710577922724486;       // e^{-2a^2}
constexpr double kE3A64 = 0.4315390418866247;       // e^{-3a^2}
constexpr double kA32 = 0.7662137865078357;         // tol 1e-7
constexpr int kN32 = 5;

__device__ __forceinline__ double core_q1(double y, double a, int N) {
    double s = 0.0, yy = y * y;
    for (int n = 1; n <= N; ++n) {
        const double an = a * n;
        s += exp(-an * an) / fma(an, an, yy);
    }
    return s;
}

__device__ __forceinline__ double rew_core64(double x, double y, double ey, double q1) {
    constexpr double a = kA64;
    x = fabs(x);
    if (x == 0.0) return ey;
    const double xx = x * x, ex2 = exp(-xx);
    if (y == 0.0) return ex2;
    const double yy = y * y;
    const double t = exp(-2.0 * a * x);
    double tn = ex2 * kEA64 * t, g = kE3A64 * t, s2 = 0.0;
    for (int n = 1; n <= kN64; ++n) {
        const double an = a * n;
        s2 = fma(tn, rcp64(fma(an, an, yy)), s2);
        tn *= g;
        g *= kE2A64;
    }
    const int n0 = (int)ceil(x / a);
    const double dl = a * n0 - x;
    const double E0 = exp(-dl * dl), e2 = exp(-2.0 * a * dl);
    double s3 = 0.0;
    {
        double E = E0, rho = e2 * kEA64;
        for (int k = 0; k <= kN64; ++k) {
            const double an = a * (n0 + k);
            s3 = fma(E, rcp64(fma(an, an, yy)), s3);
            E *= rho;
            rho *= kE2A64;
        }
    }
    {
        double E = E0, sg = kEA64 * rcp64(e2);
        for (int k = 1; k <= kN64; ++k) {
            const int n = n0 - k;
            if (n < 1) break;
            E *= sg;
            sg *= kE2A64;
            const double an = a * n;
            s3 = fma(E, rcp64(fma(an, an, yy)), s3);
        }
    }
    const double xy = x * y;
    double sxy, cxy;
    sincos(xy, &sxy, &cxy);
    const double c2 = fma(cxy, cxy, -sxy * sxy);
    const double sxy2_over_y = sxy * x * (xy < 1e-8 ? 1.0 : sxy / xy);
    constexpr double ca = 2.0 * a / kPi;
    return ex2 * ey * c2 + ca * (sxy2_over_y * ex2 + y * (-c2 * (ex2 * q1) + 0.5 * (s2 + s3)));
}

// FP32 core for r2 < 25: n0 = ceil(x/a) <= 7, so every series term uses
// n <= kNT32; dtab[n] = 1/(a^2 n^2 + y^2) is filled once per line and tile,
// with dtab[0] = 0 so that the downward terms below n = 1 vanish without a
// branch.
// As in the FP64 core, each series is built from two hardware exponentials by
// the ratio recurrences (4 exponentials per pair instead of 19); the
// recurrences add < 7 roundings per term (< 5e-7 relative).
constexpr int kNT32 = 13;
constexpr float kEA32 = 5.559463031e-01f;       // e^{-a^2},  a = kA32
constexpr float kE2A32 = 3.090762919e-01f;      // e^{-2a^2}
constexpr float kE3A32 = 1.718298218e-01f;      // e^{-3a^2}

__device__ __forceinline__ float rew_core32(float x, float y, float ey, float q1,
                                            const float *dtab) {
    constexpr float a = (float)kA32;
    x = fabsf(x);
    if (x == 0.0f) return ey;
    const float xx = x * x, ex2 = __expf(-xx);
    if (y == 0.0f) return ex2;
    // s2 = sum_n e^{-(an+x)^2} d_n: the n-th term times e^{-a^2(2n+1)} e^{-2ax}
    // gives the next
    const float t = __expf(-2.0f * a * x);
    float tn = ex2 * kEA32 * t, g = kE3A32 * t, s2 = 0.0f;
#pragma unroll
    for (int n = 1; n <= kN32; ++n) {
        s2 = fmaf(tn, dtab[n], s2);
        tn *= g;
        g *= kE2A32;
    }
    // s3 = sum_{n>=1} e^{-(an-x)^2} d_n, from its largest term n0 = ceil(x/a)
    // upwards (k <= kN32: exponents (ka + dl)^2, dl = a n0 - x in [0, a)) and
    // downwards (k <= kN32 + 1: exponents (ka - dl)^2 fall one step later)
    const int n0 = (int)ceilf(x * (float)(1.0 / kA32));
    const float dl = fmaf(a, (float)n0, -x);
    const float E0 = __expf(-dl * dl), e2 = __expf(-2.0f * a * dl);
    float s3 = 0.0f;
    {
        float E = E0, rho = e2 * kEA32;
#pragma unroll
        for (int k = 0; k <= kN32; ++k) {
            s3 = fmaf(E, dtab[n0 + k], s3);
            E *= rho;
            rho *= kE2A32;
        }
    }
    {
        float E = E0, sg = kEA32 * rcp32(e2);
#pragma unroll
        for (int k = 1; k <= kN32 + 1; ++k) {
            E *= sg;
            sg *= kE2A32;
            s3 = fmaf(E, dtab[max(n0 - k, 0)], s3);
        }
    }
    const float xy = x * y;
    float sxy, cxy;
    sincosf(xy, &sxy, &cxy);
    const float c2 = fmaf(cxy, cxy, -sxy * sxy);
    const float sxy2_over_y = sxy * x * (xy < 1e-4f ? 1.0f : sxy * rcp32(xy));
    constexpr float ca = (float)(2.0 * kA32 / kPi);
    return ex2 * ey * c2 + ca * (sxy2_over_y * ex2 + y * fmaf(-c2, ex2 * q1, 0.5f * (s2 + s3)));
}

// Reference-form core (per-term exponentials), kept for rew64.
__device__ __noinline__ double rew916_64(double x, double y) {
    return rew_core64(x, y, erfcx(y), core_q1(y, kA64, kN64));
}

// Re w for any r2 in FP64 (used by the adjoint and by tests of the pieces).
__device__ __forceinline__ double rew64(double x, double y) {
    double x2 = x * x, r2 = fma(y, y, x2);
    if (r2 >= kRF64) {
        double inv = rcp64(r2), q = inv * inv, y2 = y * y;
        double c = 1.0 + fma(3.0, x2, -y2) * 0.5 * q
                 + 0.75 * fma(fma(5.0, x2, -10.0 * y2), x2, y2 * y2) * q * q;
        return y * kISqrtPi * inv * c;
    }
    if (r2 >= kRM64) return (2.0 / kPi) * y * gh4_sum64(x2, r2);
    if (r2 >= kRB64) return (2.0 / kPi) * y * gh8_sum64(x2, r2);
    return rew916_64(x, y);
}

// --------------------------------------------------------------------------
// Complex w in FP64 (real and imaginary part), same regions as rew64.  The
// adjoint needs Im w for w'(z) = -2 z w(z) + 2i/sqrt(pi).
//  asymptotic: w ~ (i/sqrt(pi)) (1/z + 1/(2 z^3) + 3/(4 z^5))
//  Gauss-Hermite pair: Im = (2x/pi) w_k (r2 - t_k^2) / D_k
//  Algorithm 916: Im w = -e^{-x^2} erfcx(y) sin 2xy + (2a/pi) [x e^{-x^2}
//    sinc(2xy) + y sin(2xy) s1 + s4/2 - s5/2],  s4 = sum a n e^{-(an-x)^2}/
//    (a^2n^2+y^2), s5 = sum a n e^{-(an+x)^2}/(a^2n^2+y^2).
// --------------------------------------------------------------------------
__device__ __noinline__ void w916_64(double x, double y, double &re, double &im) {
    constexpr double a = 0.529275382982092;
    constexpr int N1 = 13;
    double sgn = x < 0.0 ? -1.0 : 1.0;
    x = fabs(x);
    if (x == 0.0) { re = erfcx(y); im = 0.0; return; }
    double xx = x * x, yy = y * y;
    double ex2 = exp(-xx);
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0;
    for (int n = 1; n <= N1; ++n) {
        double an = a * n;
        double inv = 1.0 / (an * an + yy);
        double e2 = exp(-(an + x) * (an + x)) * inv;
        s1 += exp(-(an * an + xx)) * inv;
        s2 += e2;
        s5 += an * e2;
    }
    int n0 = (int)ceil(x / a);
    for (int k = -N1; k <= N1; ++k) {
        int n = n0 + k;
        if (n < 1) continue;
        double an = a * n, d = an - x;
        double e3 = exp(-d * d) / (an * an + yy);
        s3 += e3;
        s4 += an * e3;
    }
    double xy = x * y, sxy, cxy;
    sincos(xy, &sxy, &cxy);
    double c2 = fma(cxy, cxy, -sxy * sxy), s2xy = 2.0 * sxy * cxy;
    double sinc1 = xy < 1e-8 ? 1.0 : sxy / xy;
    double sinc2 = xy < 1e-8 ? 1.0 : s2xy / (2.0 * xy);
    constexpr double ca = 2.0 * a / kPi;
    double ex = erfcx(y);
    re = y == 0.0 ? ex2 : ex2 * ex * c2 + ca * (sxy * x * sinc1 * ex2 + y * (-c2 * s1 + 0.5 * (s2 + s3)));
    im = sgn * (-ex2 * ex * s2xy + ca * (x * ex2 * sinc2 + y * s2xy * s1 + 0.5 * s4 - 0.5 * s5));
}

__device__ __forceinline__ void w64(double x, double y, double &re, double &im) {
    double x2 = x * x, r2 = fma(y, y, x2), y2 = y * y;
    if (r2 >= kRF64) {
        double inv = rcp64(r2), q = inv * inv;
        double cr = 1.0 + fma(3.0, x2, -y2) * 0.5 * q
                  + 0.75 * fma(fma(5.0, x2, -10.0 * y2), x2, y2 * y2) * q * q;
        double ci = 1.0 + fma(-3.0, y2, x2) * 0.5 * q
                  + 0.75 * fma(fma(-10.0, y2, x2), x2, 5.0 * y2 * y2) * q * q;
        re = y * kISqrtPi * inv * cr;
        im = x * kISqrtPi * inv * ci;
        return;
    }
    if (r2 >= kRB64) {
        double sr = 0.0, si = 0.0;
        if (r2 >= kRM64) {
            const double t[2] = {kGH4t1, kGH4t2}, w[2] = {kGH4w1, kGH4w2};
            for (int k = 0; k < 2; ++k) {
                double a = t[k] * t[k], u = r2 + a;
                double iD = rcp64(fma(u, u, -4.0 * a * x2));
                sr = fma(w[k] * u, iD, sr);
                si = fma(w[k] * (r2 - a), iD, si);
            }
        } else {
            const double t[4] = {kGH8t1, kGH8t2, kGH8t3, kGH8t4};
            const double w[4] = {kGH8w1, kGH8w2, kGH8w3, kGH8w4};
            for (int k = 0; k < 4; ++k) {
                double a = t[k] * t[k], u = r2 + a;
                double iD = rcp64(fma(u, u, -4.0 * a * x2));
                sr = fma(w[k] * u, iD, sr);
                si = fma(w[k] * (r2 - a), iD, si);
            }
        }
        re = (2.0 / kPi) * y * sr;
        im = (2.0 / kPi) * x * si;
        return;
    }
    w916_64(x, y, re, im);
}

// dRe w/dx = Re w'(z) and dRe w/dy = -Im w'(z) for |z|^2 >= 1e4 from the
// asymptotic series w'(z) = -(i/sqrt(pi)) u (1 + 3/2 u + 15/4 u^2 + 105/8 u^3),
// u = 1/z^2 (next term below 6e-15 relative).  The identity
// w' = -2zw + 2i/sqrt(pi) cancels to ~eps |z|^2 there.
__device__ __forceinline__ void drew_far64(double x, double y, double r2, double &ddx,
                                           double &ddy) {
    const double inv = rcp64(r2);
    const double ur = (x - y) * (x + y) * inv * inv, ui = -2.0 * (x * inv) * (y * inv);
    double pr = 13.125 * ur + 3.75, pi = 13.125 * ui;
    double tr = pr * ur - pi * ui, ti = pr * ui + pi * ur;
    pr = tr + 1.5; pi = ti;
    tr = pr * ur - pi * ui; ti = pr * ui + pi * ur;
    pr = tr + 1.0; pi = ti;
    ddx = (ur * pi + ui * pr) * kISqrtPi;
    ddy = (ur * pr - ui * pi) * kISqrtPi;
}

}  // namespace hk
