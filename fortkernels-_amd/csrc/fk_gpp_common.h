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
    double wx2[kMaxFreq];      // cden = pow(wxt,2) of the mask branch (:41)
    double rden1[kMaxFreq];    // 1 / |cden|^2 (:42-43)
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
        const double w2 = w * w;
        pp.pc.wx[iw] = w;
        pp.pc.wx2[iw] = w2;
        pp.pc.rden1[iw] = 1.0 / (w2 * w2);
    }
    return pp;
}

__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2((double)v.x, (double)v.y); }

// 1/x from v_rcp_f64 refined by two Newton steps (within an ulp or two of
// the IEEE quotient).
__device__ __forceinline__ double rcp_fast(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
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

// ssx of the near-pole branch (:46-53): -Omega2 conj(cden)/|cden|^2 delw,
// cden = 4 wtilde^2 (delw + 0.5).
__device__ __forceinline__ double2 ssx_branch2(double2 wt2, double2 Om, double2 delw)
{
    double2 cden = cmul(make_double2(4.0 * wt2.x, 4.0 * wt2.y), make_double2(delw.x + 0.5, delw.y));
    double rd = 1.0 / (cden.x * cden.x + cden.y * cden.y);
    double2 t = cmul(make_double2(-Om.x, -Om.y), cconj(cden));
    return cmul(make_double2(t.x * rd, t.y * rd), delw);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One (my_igp, ig) element for the NWC frequencies w0 .. w0+NWC-1: adds
// B * D to sach[q] and Bo * ssx to sasx[q], where D = mask * delw * I_eps
// and ssx is the three-way SX factor.  `ok` = a real, kept element; a
// frequency past nw repeats the last one and adds nothing.
template <int NWC>
__device__ __forceinline__ void pole_element(const PoleParams &pp, int w0, bool ok, double2 wt, double2 ie,
                                             double2 B, double2 Bo, bool want_ssx, double2 (&sach)[NWC],
                                             double2 (&sasx)[NWC])
{
    const double tol1 = 1e-13 * pp.limitone, tol2 = 1e-13 * pp.to1;
    const double2 wt2 = cmul(wt, wt);
    const double2 Om = cmul(wt2, ie);
#pragma unroll
    for (int q = 0; q < NWC; ++q) {
        const int iw = min(w0 + q, pp.nw - 1);
        const bool live = ok && (w0 + q < pp.nw);
        const double wdr = pp.pc.wx[iw] - wt.x;
        const double wdiffr = pole_wdiffr(wdr, -wt.y);
        double2 delw;
        double delwr;
        pole_delw(wt, wdr, rcp_fast(wdiffr), delw, delwr);
        // points within 1e-13 of a branch threshold take the IEEE quotient,
        // so every branch decision is the reference's
        const bool near = fabs(delwr - pp.limitone) <= tol1 || fabs(delwr - pp.to1) <= tol2;
        if (__ballot(near)) {
            if (near) pole_delw(wt, wdr, 1.0 / wdiffr, delw, delwr);
        }
        const bool m = (wdiffr > pp.limittwo) && (delwr < pp.limitone);
        const double2 D = (m || !pp.mask_on) ? cmul(delw, ie) : make_double2(0.0, 0.0);
        const double2 t = cmul(B, D);
        if (live) sach[q] = cadd(sach[q], t);
        if (want_ssx) {
            // mask branch (:41-44): Omega2 * conj(wx^2) / |wx^2|^2
            double2 S = m ? make_double2(Om.x * pp.pc.wx2[iw] * pp.pc.rden1[iw], Om.y * pp.pc.wx2[iw] * pp.pc.rden1[iw])
                          : make_double2(0.0, 0.0);
            const bool need2 = !m && delwr > pp.to1;
            if (__ballot(need2)) {
                if (need2) S = ssx_branch2(wt2, Om, delw);
            }
            const double2 u = cmul(Bo, S);
            if (live) sasx[q] = cadd(sasx[q], u);
        }
    }
}

}  // namespace gpp
}  // namespace fk
