// fk_gpp_common.h -- device code shared by the GPP kernels: the pole factor
// and SX factor of one (my_igp, ig) element (BGW-Kernels/gppKer.cpp:19-65,
// :394-411) and the reductions around them.
#pragma once

#include "fk_complex.h"

namespace fk {
namespace gpp {

constexpr int kMaxFreq = 16;
constexpr int kXcds = 8;       // MI355X: consecutive workgroups go to different XCDs

struct PoleConst {
    double wx[kMaxFreq];
    double iwx2[kMaxFreq];     // 1 / wx^2: conj(cden) / |cden|^2 of the mask branch, cden = wx^2 real (:41-44)
};

// Scalars of the per-element epilogue.
struct PoleParams {
    double limittwo, limitone, to1;
    int mask_on;
    int nw;
    PoleConst pc;
};

inline PoleParams make_pole_params(const double *wx, int nfreq, double limittwo, double limitone, double to1,
                                   bool mask_on)
{
    PoleParams pp;
    pp.limittwo = limittwo;
    pp.limitone = limitone;
    pp.to1 = to1;
    pp.mask_on = mask_on ? 1 : 0;
    pp.nw = nfreq;
    for (int iw = 0; iw < kMaxFreq; ++iw) {
        const double w = iw < nfreq ? wx[iw] : 0.0;
        pp.pc.wx[iw] = w;
        pp.pc.iwx2[iw] = iw < nfreq ? 1.0 / (w * w) : 0.0;
    }
    return pp;
}

__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2((double)v.x, (double)v.y); }

// 1/x from v_rcp_f64 refined by STEPS Newton steps.  Measured on MI355X
// over 4M inputs spanning 2^-40 .. 2^40: v_rcp_f64 alone is within 4.2e8
// units of 2^-53 (about 2^-24 relative), one step within 20, two steps
// equal the IEEE quotient.
template <int STEPS = 2>
__device__ __forceinline__ double rcp_fast(double x)
{
    double r = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int k = 0; k < STEPS; ++k) {
        const double e = __builtin_fma(-x, r, 1.0);
        r = __builtin_fma(r, e, r);
    }
    return r;
}

// delw = wtilde * conj(wdiff) * rden and |delw|^2 in the reference's FP64
// evaluation order (BGW-Kernels/gppKer.cpp:30-36, :396-402).  Contraction is
// off so, given the same rden, they round exactly as the reference's.
__device__ __forceinline__ void pole_delw(double2 wt, double wdr, double rden, double2 &delw, double &delwr)
{
#pragma clang fp contract(off)
    const double wdi = -wt.y;
    delw = make_double2((wt.x * wdr - wt.y * (-wdi)) * rden, (wt.x * (-wdi) + wt.y * wdr) * rden);
    delwr = delw.x * delw.x + delw.y * delw.y;
}

__device__ __forceinline__ double pole_wdiffr(double wdr, double wdi)
{
#pragma clang fp contract(off)
    return wdr * wdr + wdi * wdi;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One (my_igp, ig) element for the NWC frequencies w0 .. w0+NWC-1: adds
// B * D to sach[q] and Bo * ssx to sasx[q], where D = mask * delw * I_eps
// and ssx is the three-way SX factor (:37-57).  `ok` = a real, kept
// element; a frequency past nw repeats the last one and adds nothing.
// The SX factor of the mask branch (:41-44) is Omega2 / wx^2 (cden = wx^2
// is real); that of the near-pole branch (:49-52) is
// -I_eps delw conj(z) / (4 |z|^2), z = delw + 0.5 (Omega2 conj(cden) /
// |cden|^2 with cden = 4 wtilde^2 z and Omega2 = wtilde^2 I_eps loses its
// wtilde^2).
// HOIST = false evaluates the products per frequency, as the MFMA kernels'
// epilogues need (fewer live registers); HOIST = true forms B I_eps,
// Bo Omega2 and Bo I_eps once per element, for the stand-alone pole kernel:
//   B D = (B I_eps) delw,  Bo ssx = (Bo Omega2) / wx^2  or
//   -(Bo I_eps) delw conj(z) / (4 |z|^2).
// Both equal the reference's expressions up to rounding.
template <int NWC, bool HOIST = false>
__device__ __forceinline__ void pole_element(const PoleParams &pp, int w0, bool ok, double2 wt, double2 ie,
                                             double2 B, double2 Bo, bool want_ssx, double2 (&sach)[NWC],
                                             double2 (&sasx)[NWC])
{
    // HOIST (the mixed-precision pole kernel): one Newton step, 2e-15
    // relative, far inside the 1e-13 band that takes the IEEE quotient
    constexpr int kSteps = HOIST ? 1 : 2;
    const double tol1 = 1e-13 * pp.limitone, tol2 = 1e-13 * pp.to1;
    const double2 Om = cmul(cmul(wt, wt), ie);
    double2 Bie = B, BoOm = Om, Boie = ie;
    if constexpr (HOIST) {
        Bie = cmul(B, ie);
        BoOm = cmul(Bo, Om);
        Boie = cmul(Bo, ie);
    }
#pragma unroll
    for (int q = 0; q < NWC; ++q) {
        const int iw = min(w0 + q, pp.nw - 1);
        const bool live = ok && (w0 + q < pp.nw);
        const double wdr = pp.pc.wx[iw] - wt.x;
        const double wdiffr = pole_wdiffr(wdr, -wt.y);
        double2 delw;
        double delwr;
        pole_delw(wt, wdr, rcp_fast<kSteps>(wdiffr), delw, delwr);
        // points within 1e-13 of a branch threshold take the IEEE quotient,
        // so every branch decision is the reference's
        const bool near = fabs(delwr - pp.limitone) <= tol1 || fabs(delwr - pp.to1) <= tol2;
        if (__ballot(near)) {
            if (near) pole_delw(wt, wdr, 1.0 / wdiffr, delw, delwr);
        }
        const bool m = (wdiffr > pp.limittwo) && (delwr < pp.limitone);
        if constexpr (HOIST) {
            const double2 t = (live && (m || !pp.mask_on)) ? cmul(Bie, delw) : make_double2(0.0, 0.0);
            sach[q] = cadd(sach[q], t);
        } else {
            const double2 D = (m || !pp.mask_on) ? cmul(delw, ie) : make_double2(0.0, 0.0);
            const double2 t = cmul(B, D);
            if (live) sach[q] = cadd(sach[q], t);
        }
        if (want_ssx) {
            double2 S = m ? make_double2(BoOm.x * pp.pc.iwx2[iw], BoOm.y * pp.pc.iwx2[iw]) : make_double2(0.0, 0.0);
            const bool need2 = !m && delwr > pp.to1;
            if (__ballot(need2)) {
                if (need2) {
                    const double2 z = make_double2(delw.x + 0.5, delw.y);
                    const double sc = -0.25 * rcp_fast<kSteps>(z.x * z.x + z.y * z.y);
                    const double2 v = cmul(Boie, cmul(delw, cconj(z)));
                    S = make_double2(sc * v.x, sc * v.y);
                }
            }
            if constexpr (HOIST) {
                if (live) sasx[q] = cadd(sasx[q], S);
            } else {
                const double2 u = cmul(Bo, S);
                if (live) sasx[q] = cadd(sasx[q], u);
            }
        }
    }
}

}  // namespace gpp
}  // namespace fk
