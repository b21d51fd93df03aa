// hk_lines.hip -- line parameters at one (T, P).
//
// One thread per line computes, in FP64 and in the reference's order of
// operations, the per-line part of readFile (host.h:641-644) followed by
// S2_kernel (voigt.h:338-356) -- or, for ExoMol-family and super-line lists,
// Sf_kernel / SfSuper_kernel (voigt.h:365-420; see HK_FORM_* in heliosk.h):
//   1/alpha_D = c sqrt(m / (2 kB T)) / nu,    nu = nu0 + delta P
//   S = S_file/Q*Ab*Sscale * exp(-c2 E/T + c2 E/T0)
//       * (1 - exp(-c2 nu/T)) / (1 - exp(-c2 nu/T0)) * (1/alpha_D)
//   alpha_L = vy (P/1 atm) (T0/T)^n + A/(4 pi c),  y = alpha_L / alpha_D
// and the cut half width of S3_kernel (voigt.h:442-450) in cm^-1, plus
// erfcx(y) for the line-core series of the Voigt kernels.  Each
// workgroup also writes the block maxima the Voigt kernels use for their
// line-range search radius and FP32 scaling.
#include <math.h>

#include "hk_device.h"
#include "hk_internal.h"

namespace {

struct IsoTab {
    double mass[HK_MAX_ISO], ab[HK_MAX_ISO], ss[HK_MAX_ISO], Q[HK_MAX_ISO];
};

__global__ __launch_bounds__(256) void line_params_kernel(
    const double *__restrict__ nu0, const double *__restrict__ S0,
    const double *__restrict__ EL, const double *__restrict__ A,
    const double *__restrict__ delta, const double *__restrict__ vy0,
    const double *__restrict__ nexp, const int32_t *__restrict__ iso, IsoTab tab, int form,
    int64_t n, double T, double P, int cut_mode, double cut,
    double *__restrict__ o_nu, double *__restrict__ o_S, double *__restrict__ o_ia,
    double *__restrict__ o_y, double *__restrict__ o_hw, double *__restrict__ o_ey,
    double *__restrict__ blk_max,
    double *__restrict__ blk_rad, double *__restrict__ blk_core) {
    using namespace hk;
    __shared__ double smax[256], srad[256], score[256];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 256 + t;
    double smx = 0.0, rad = 0.0, crad = 0.0;
    if (i < n) {
        const double c = kH * kC / kKB;
        int k = iso ? min(max(iso[i], 0), HK_MAX_ISO - 1) : 0;
        double ialphaD_h = kC * sqrt(tab.mass[k] / (2.0 * kKB * T));
        double Sh = form == HK_FORM_SUPERLINE ? S0[i] * tab.ab[k] * tab.ss[k]
                                              : S0[i] / tab.Q[k] * tab.ab[k] * tab.ss[k];
        double key = nu0[i];
        double nu = form == HK_FORM_HITRAN ? fma(delta[i], P, key) : key;
        double ia = ialphaD_h / nu;
        double E = EL[i];
        double S;
        if (form == HK_FORM_HITRAN)
            S = Sh * (exp(-E * c / T + E * c / kT0) * (1.0 - exp(-c * nu / T))
                      / (1.0 - exp(-c * nu / kT0)) * ia);
        else if (form == HK_FORM_EXOMOL)
            S = Sh * (exp(-c / T * E) * (1.0 - exp(-c / T * nu)) * ia);   // voigt.h:379
        else
            S = Sh * ia;                                                  // voigt.h:407
        double Pn = form == HK_FORM_HITRAN ? P / kPOatm : P / kPObar;
        double An = form == HK_FORM_SUPERLINE ? 0.0 : A[i] / (4.0 * kPi * kC);
        double aL = fma(vy0[i], Pn * pow(kT0 / T, nexp[i]), An);
        double y = aL * ia;
        double hw = cut_mode == 1 ? cut * aL : (cut_mode == 2 ? cut / ia : cut);
        if (nu == 0.0) { S = 0.0; ia = 0.0; y = 0.0; hw = -1.0; }
        o_nu[i] = nu; o_S[i] = S; o_ia[i] = ia; o_y[i] = y; o_hw[i] = hw;
        o_ey[i] = erfcx(y);
        smx = fabs(S);
        rad = hw > 0.0 ? hw + fabs(nu - key) : 0.0;
        // the line core x^2 + y^2 < 100 lies within 10 alpha_D of the centre
        // and is empty for y >= 10 (10.5: margin for rounding in x)
        crad = (hw > 0.0 && y < 10.5) ? fmin(hw, 10.5 / ia) + fabs(nu - key) : 0.0;
    }
    smax[t] = smx; srad[t] = rad; score[t] = crad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            smax[t] = fmax(smax[t], smax[t + s]);
            srad[t] = fmax(srad[t], srad[t + s]);
            score[t] = fmax(score[t], score[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        blk_max[blockIdx.x] = smax[0];
        blk_rad[blockIdx.x] = srad[0];
        blk_core[blockIdx.x] = score[0];
    }
}

// out[r] = max over the n values v[r * n ..] (workgroup r; fixed order:
// deterministic).
__global__ __launch_bounds__(1024) void reduce_max_kernel(const double *__restrict__ v,
                                                          int64_t n, double *__restrict__ out) {
    __shared__ double s[1024];
    v += (int64_t)blockIdx.x * n;
    out += blockIdx.x;
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, v[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

}  // namespace

extern "C" size_t hk_line_state_bytes(int64_t nlines) {
    if (nlines < 0) return 0;
    int64_t nb = (nlines + hk::kBlockLines - 1) / hk::kBlockLines;
    // nu, S, ialphaD, y, hw, ey  +  blk_max  +  radius[0..1] and per-block radii
    return (size_t)(6 * nlines + nb + 2 + 2 * nb) * sizeof(double) + 256;
}

extern "C" int hk_line_state_init(hk_line_state *st, int64_t nlines, void *mem, size_t bytes) {
    HK_CHECK_ARG(st && mem, "hk_line_state_init: null argument");
    HK_CHECK_ARG(nlines >= 0, "hk_line_state_init: nlines < 0");
    HK_CHECK_ARG(bytes >= hk_line_state_bytes(nlines),
                 "hk_line_state_init: %zu bytes given, %zu needed", bytes,
                 hk_line_state_bytes(nlines));
    double *p = (double *)hk::align_up((size_t)mem, 256);
    int64_t nb = (nlines + hk::kBlockLines - 1) / hk::kBlockLines;
    st->nlines = nlines;
    st->nu = p; p += nlines;
    st->S = p; p += nlines;
    st->ialphaD = p; p += nlines;
    st->y = p; p += nlines;
    st->hw = p; p += nlines;
    st->ey = p; p += nlines;
    st->blk_max = p; p += nb;
    // radius[0]: cut reach, radius[1]: line-core reach; then the nb per-block
    // cut reaches and the nb per-block core reaches
    st->radius = p;
    st->key = nullptr;
    return HK_OK;
}

extern "C" int hk_line_params(const hk_lines *L, const hk_tp *tp, const hk_cut *cut,
                              hk_line_state *st, hk_stream_t stream) {
    HK_CHECK_ARG(L && tp && cut && st, "hk_line_params: null argument");
    HK_CHECK_ARG(st->nlines == L->nlines, "hk_line_params: state holds %lld lines, list %lld",
                 (long long)st->nlines, (long long)L->nlines);
    HK_CHECK_ARG(L->niso >= 1 && L->niso <= HK_MAX_ISO, "hk_line_params: niso %d out of 1..%d",
                 L->niso, HK_MAX_ISO);
    HK_CHECK_ARG(L->form >= HK_FORM_HITRAN && L->form <= HK_FORM_SUPERLINE,
                 "hk_line_params: form %d", L->form);
    HK_CHECK_ARG(tp->T > 0.0, "hk_line_params: T must be > 0");
    HK_CHECK_ARG(cut->cut_mode >= 0 && cut->cut_mode <= 2, "hk_line_params: cut_mode %d",
                 cut->cut_mode);
    const int64_t n = L->nlines;
    const int64_t nb = (n + hk::kBlockLines - 1) / hk::kBlockLines;
    st->key = L->nu;
    double cutv = cut->cut == 0.0 ? 1.0e30 : cut->cut;    // heliosk.cu:941
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(reduce_max_kernel, dim3(2), dim3(1024), 0, s, st->radius + 2,
                           (int64_t)0, st->radius);
        HK_LAUNCH_CHECK();
        return HK_OK;
    }
    IsoTab tab;
    for (int k = 0; k < HK_MAX_ISO; ++k) {
        bool used = k < L->niso;
        tab.mass[k] = used ? L->iso_mass[k] : 1.0;
        tab.ab[k] = used ? L->iso_ab[k] : 0.0;
        tab.ss[k] = used ? L->iso_sscale[k] : 0.0;
        tab.Q[k] = used ? tp->Q[k] : 1.0;
    }
    hipLaunchKernelGGL(line_params_kernel, dim3((unsigned)nb), dim3(256), 0, s, L->nu, L->S,
                       L->EL, L->A, L->delta, L->vy, L->n, L->iso, tab, L->form, n, tp->T, tp->P,
                       cut->cut_mode, cutv, st->nu, st->S, st->ialphaD, st->y, st->hw, st->ey,
                       st->blk_max, st->radius + 2, st->radius + 2 + nb);
    HK_LAUNCH_CHECK();
    // the two reaches (per-block values at radius + 2 and radius + 2 + nb)
    hipLaunchKernelGGL(reduce_max_kernel, dim3(2), dim3(1024), 0, s, st->radius + 2, nb,
                       st->radius);
    HK_LAUNCH_CHECK();
    return HK_OK;
}
