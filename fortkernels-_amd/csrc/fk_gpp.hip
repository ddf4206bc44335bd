// fk_gpp.hip -- GPP band sum on MI355X (gfx950).
//
// The reference evaluates, per band n1, column my_igp, G-vector ig and
// frequency iw (BGW-Kernels/gppKer.cpp:303-439), one term
//     0.5 vcoul[igp] conj(aqsm[n1][igp]) aqsn[n1][ig] * D(my_igp, ig, iw)
// with the pole factor D = mask * delw * I_eps[my_igp][ig] (:394-411).
// D does not depend on n1, so the band sum factorises:
//     achtemp[iw] = sum_{my_igp, ig} B[my_igp][ig] * D(my_igp, ig, iw)
//     B[my_igp][ig] = sum_n1 c[n1][my_igp] aqsn[n1][ig],
//     c[n1][my_igp] = 0.5 vcoul[igp] conj(aqsm[n1][igp]),
// a complex GEMM (M = ngpown, N = ncouls, K = number_bands) whose epilogue
// evaluates D and the SX factor ssx (:19-65) per (my_igp, ig, iw) and reduces.
// The occupied-band SX sum (:337-347, :416-419) is the same contraction over
// n1 < nvband with ssx in place of D.
//
// Kernels (one call = three launches on the caller's stream):
//   gpp_coef_kernel   c[Kpad][ldc] from aqsm / vcoul / the column map
//   gpp_fused_kernel  MFMA GEMM tile (LDS double-buffered) + pole/SX
//                     epilogue + wave/workgroup reduction -> per-tile partials
//   gpp_reduce_kernel fixed-order sum of the per-tile partials -> outputs
// Every sum is taken in a fixed order, so results are bitwise reproducible.
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>
#include <cmath>

namespace fk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMaxFreq = 16;
constexpr int kChunk = 4;      // frequencies per epilogue pass
constexpr int kXcds = 8;       // MI355X: consecutive workgroups go to different XCDs

struct PoleConst {
    double wx[kMaxFreq];
    double wx2[kMaxFreq];      // cden = pow(wxt,2) of the mask branch (:41)
    double rden1[kMaxFreq];    // 1 / |cden|^2 (:42-43)
};

template <class E>
struct GppArgs {
    const E *c;            // [Kpad][ldc] zero-padded coefficient matrix
    const E *aqsn;         // [nb][C]
    const E *I_eps;        // [G][C]
    const E *wtilde;       // [G][C]
    const int *igp;        // [G], -1 = skipped column
    double2 *partials;     // [tiles][nw][2] (achtemp part, asxtemp part)
    int nb, nv, C, G, nw, ldc, Kpad;
    int tiles_m, tiles_n;
    double limittwo, limitone, to1;
    int mask_on;
    PoleConst pc;
};

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class E> struct Acc;
template <> struct Acc<double2> { using type = d4; };
template <> struct Acc<float2> { using type = f4; };

__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2((double)v.x, (double)v.y); }

// Row of the 16x16 MFMA result held in accumulator register r of lane l
// (column l % 16): FP64 16x16x4 interleaves rows across registers, FP32
// 16x16x4 keeps four consecutive rows per lane group.
template <class E> __device__ __forceinline__ int acc_row(int lane, int r);
template <> __device__ __forceinline__ int acc_row<double2>(int lane, int r) { return 4 * r + (lane >> 4); }
template <> __device__ __forceinline__ int acc_row<float2>(int lane, int r) { return 4 * (lane >> 4) + r; }

template <class E> __device__ __forceinline__ E zero_elem();
template <> __device__ __forceinline__ double2 zero_elem<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero_elem<float2>() { return make_float2(0.f, 0.f); }

// Workgroup b -> logical tile, so that the tiles one XCD runs back to back
// share an aqsn column block (tm varies fastest within an XCD's range).
__device__ __forceinline__ void tile_coords(int b, int tiles_m, int total, int &tm, int &tn)
{
    int q = total / kXcds, r = total % kXcds, x = b % kXcds, i = b / kXcds;
    int t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    tm = t % tiles_m;
    tn = t / tiles_m;
}

// Pole factor D (achtemp) and SX factor ssx (asxtemp) of one (my_igp, ig, iw)
// point, in the reference's FP64 evaluation order (ssxt_scht_solver,
// BGW-Kernels/gppKer.cpp:19-65; unoccupied loop :396-406).  Contraction is
// off so |wdiff|^2, delw and |delw|^2 round exactly as the reference's and
// the mask/branch decisions are identical to it.
__device__ __forceinline__ void pole_point(double wx, double wx2, double rden1, double2 wt, double2 ie, double2 wt2,
                                           double2 Om, double limittwo, double limitone, double to1, bool mask_on,
                                           bool want_ssx, double2 &D, double2 &S)
{
#pragma clang fp contract(off)
    double wdr = wx - wt.x, wdi = -wt.y;                       // wdiff = wxt - wtilde
    double wdiffr = wdr * wdr + wdi * wdi;                     // real(wdiff*conj(wdiff))
    double rden = 1.0 / wdiffr;
    double2 delw = make_double2((wt.x * wdr - wt.y * (-wdi)) * rden,   // wtilde*conj(cden)*rden
                                (wt.x * (-wdi) + wt.y * wdr) * rden);
    double delwr = delw.x * delw.x + delw.y * delw.y;
    bool m = (wdiffr > limittwo) && (delwr < limitone);
    D = (m || !mask_on) ? cmul(delw, ie) : make_double2(0.0, 0.0);
    S = make_double2(0.0, 0.0);
    if (want_ssx) {
        if (m) {                                               // :41-44
            S = make_double2(Om.x * wx2 * rden1, Om.y * wx2 * rden1);
        } else if (delwr > to1) {                              // :49-52
            double2 cden = cmul(make_double2(4.0 * wt2.x, 4.0 * wt2.y), make_double2(delw.x + 0.5, delw.y));
            double rd = 1.0 / (cden.x * cden.x + cden.y * cden.y);
            double2 t = cmul(make_double2(-Om.x, -Om.y), cconj(cden));
            S = cmul(make_double2(t.x * rd, t.y * rd), delw);
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---------------------------------------------------------------------------
// c[k][m] = 0.5 vcoul[igp] conj(aqsm[k][igp]) for k < nb, m < G and a kept
// column; zero elsewhere (padding to Kpad x ldc).  The resolved column map
// travels in the kernel arguments (a chunk of columns per launch), so the
// call needs no host-to-device copy; the launch also stores it for the
// fused kernel.
// ---------------------------------------------------------------------------
constexpr int kColChunk = 896;

template <class E>
struct CoefArgs {
    const E *aqsm;
    const double *vcoul;
    E *c;
    int *igp_out;
    int nb, C, G, Kpad, ldc, m_begin, m_count;
    int igp[kColChunk];
};

template <class E>
__global__ void gpp_coef_kernel(CoefArgs<E> a)
{
    int64_t n = (int64_t)a.Kpad * a.m_count;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        int k = (int)(idx / a.m_count), mm = (int)(idx % a.m_count), m = a.m_begin + mm;
        int g = m < a.G ? a.igp[mm] : -1;
        if (k == 0 && m < a.G) a.igp_out[m] = g;
        double2 v = make_double2(0.0, 0.0);
        if (k < a.nb && g >= 0) {
            double2 am = widen(a.aqsm[(size_t)k * a.C + g]);
            double s = 0.5 * a.vcoul[g];
            v = make_double2(am.x * s, -am.y * s);
        }
        E *dst = a.c + (size_t)k * a.ldc + m;
        if constexpr (sizeof(E) == sizeof(double2)) *dst = v;
        else *dst = make_float2((float)v.x, (float)v.y);
    }
}

// ---------------------------------------------------------------------------
// Fused GEMM tile + epilogue.  256 threads = 4 waves in a 2x2 arrangement;
// workgroup tile BM (my_igp) x BN (ig); wave tile BM/2 x BN/2 made of 16x16
// MFMA blocks.  MFMA 16x16x4 operand layout: lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; accumulator register r of lane l is C[acc_row(l,r)][l%16].
// Complex products use four real MFMAs (re += ar br - ai bi, im += ar bi + ai br).
// ---------------------------------------------------------------------------
template <class E, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gpp_fused_kernel(GppArgs<E> a)
{
    using A4 = typename Acc<E>::type;
    constexpr int WM = BM / 2, WN = BN / 2, MB = WM / 16, NB = WN / 16;
    constexpr int LA = BK * BM / 256, LB = BK * BN / 256;
    static_assert(BK % 4 == 0 && LA * 256 == BK * BM && LB * 256 == BK * BN, "tile shape");

    __shared__ E As[2][BK * BM];
    __shared__ E Bs[2][BK * BN];
    __shared__ double red[4][4 * kChunk];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_coords(blockIdx.x, a.tiles_m, a.tiles_m * a.tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    A4 acc_re[MB][NB], acc_im[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            acc_re[i][j] = A4{0, 0, 0, 0};
            acc_im[i][j] = A4{0, 0, 0, 0};
        }

    E ra[LA], rb[LB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            int idx = tid + 256 * i, kk = idx / BM, mm = idx % BM;
            ra[i] = a.c[(size_t)(k0 + kk) * a.ldc + m0 + mm];
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            int idx = tid + 256 * i, k = k0 + idx / BN, n = n0 + idx % BN;
            rb[i] = (k < a.nb && n < a.C) ? a.aqsn[(size_t)k * a.C + n] : zero_elem<E>();
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LA; ++i) As[buf][tid + 256 * i] = ra[i];
#pragma unroll
        for (int i = 0; i < LB; ++i) Bs[buf][tid + 256 * i] = rb[i];
    };

    const int nk = a.Kpad / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            E af[MB], bf[NB];
#pragma unroll
            for (int i = 0; i < MB; ++i) af[i] = As[buf][kr * BM + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NB; ++j) bf[j] = Bs[buf][kr * BN + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc_re[i][j] = mfma(af[i].x, bf[j].x, acc_re[i][j]);
                    acc_im[i][j] = mfma(af[i].x, bf[j].y, acc_im[i][j]);
                }
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc_re[i][j] = mfma(-af[i].y, bf[j].y, acc_re[i][j]);
                    acc_im[i][j] = mfma(af[i].y, bf[j].x, acc_im[i][j]);
                }
        }
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }

    // ---- epilogue: pole / SX factors and the reduction over the tile ----
    const bool want_ssx = a.nv > 0;
    const size_t tile = (size_t)tm + (size_t)tn * a.tiles_m;
    for (int w0 = 0; w0 < a.nw; w0 += kChunk) {
        double2 sach[kChunk], sasx[kChunk];
#pragma unroll
        for (int q = 0; q < kChunk; ++q) { sach[q] = make_double2(0.0, 0.0); sasx[q] = make_double2(0.0, 0.0); }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WM + i * 16 + acc_row<E>(lane, r);
                if (row >= a.G || a.igp[row] < 0) continue;
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const int col = n0 + wn * WN + j * 16 + (lane & 15);
                    if (col >= a.C) continue;
                    const double2 B = make_double2((double)acc_re[i][j][r], (double)acc_im[i][j][r]);
                    const size_t e = (size_t)row * a.C + col;
                    const double2 wt = widen(a.wtilde[e]);
                    const double2 ie = widen(a.I_eps[e]);
                    double2 Bocc = make_double2(0.0, 0.0), wt2 = Bocc, Om = Bocc;
                    if (want_ssx) {
                        for (int n1 = 0; n1 < a.nv; ++n1)
                            Bocc = cfma(Bocc, widen(a.c[(size_t)n1 * a.ldc + row]), widen(a.aqsn[(size_t)n1 * a.C + col]));
                        wt2 = cmul(wt, wt);
                        Om = cmul(wt2, ie);
                    }
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) {
                        const int iw = w0 + q;
                        if (iw >= a.nw) break;
                        double2 D, S;
                        pole_point(a.pc.wx[iw], a.pc.wx2[iw], a.pc.rden1[iw], wt, ie, wt2, Om, a.limittwo,
                                   a.limitone, a.to1, a.mask_on != 0, want_ssx, D, S);
                        sach[q] = cfma(sach[q], B, D);
                        sasx[q] = cfma(sasx[q], Bocc, S);
                    }
                }
            }
        // wave reduction, then a fixed-order sum over the four waves
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
            double v0 = wave_sum(sach[q].x), v1 = wave_sum(sach[q].y);
            double v2 = wave_sum(sasx[q].x), v3 = wave_sum(sasx[q].y);
            if (lane == 0) {
                red[wave][4 * q + 0] = v0;
                red[wave][4 * q + 1] = v1;
                red[wave][4 * q + 2] = v2;
                red[wave][4 * q + 3] = v3;
            }
        }
        __syncthreads();
        if (tid < 4 * kChunk) {
            const int q = tid >> 2, comp = tid & 3;
            const int iw = w0 + q;
            if (iw < a.nw) {
                double s = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
                double *dst = reinterpret_cast<double *>(a.partials + (tile * a.nw + iw) * 2);
                dst[comp] = s;   // comp 0,1: achtemp re,im; 2,3: asxtemp re,im
            }
        }
        __syncthreads();
    }
}

// One workgroup per output value: sums the per-tile partials in a fixed order.
__global__ __launch_bounds__(256) void gpp_reduce_kernel(const double2 *partials, int tiles, int nw, double asx_scale,
                                                          double2 *achtemp, double2 *asxtemp)
{
    __shared__ double s_re[256], s_im[256];
    const int o = blockIdx.x;          // o = 2*iw + which
    const int iw = o >> 1, which = o & 1;
    double re = 0.0, im = 0.0;
    for (int t = threadIdx.x; t < tiles; t += 256) {
        double2 v = partials[((size_t)t * nw + iw) * 2 + which];
        re += v.x;
        im += v.y;
    }
    s_re[threadIdx.x] = re;
    s_im[threadIdx.x] = im;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_re[threadIdx.x] += s_re[threadIdx.x + w];
            s_im[threadIdx.x] += s_im[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (which == 0) achtemp[iw] = make_double2(s_re[0], s_im[0]);
        else asxtemp[iw] = make_double2(s_re[0] * asx_scale, s_im[0] * asx_scale);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct TileCfg { int bm, bn, bk; };
constexpr TileCfg kCfgs[] = {{64, 128, 16}, {32, 64, 16}};
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

int choose_cfg(const fk_gpp_problem *p, const fk_options *o)
{
    if (o->config > 0) return o->config - 1;
    // the big tile unless it leaves the 256 CUs with under two tiles each
    long tiles = (long)div_up(p->ngpown, kCfgs[0].bm) * div_up(p->ncouls, kCfgs[0].bn);
    return tiles >= 512 ? 0 : 1;
}

struct Layout {
    int cfg, Kpad, ldc, tiles_m, tiles_n;
    size_t off_c, off_igp, off_part;
};

Layout plan_layout(const fk_gpp_problem *p, const fk_options *o, Scratch &ws, size_t elem)
{
    Layout L;
    L.cfg = choose_cfg(p, o);
    const TileCfg &t = kCfgs[L.cfg];
    L.Kpad = div_up(std::max(p->number_bands, 1), t.bk) * t.bk;
    L.ldc = div_up(p->ngpown, t.bm) * t.bm;
    L.tiles_m = L.ldc / t.bm;
    L.tiles_n = div_up(p->ncouls, t.bn);
    L.off_c = ws.plan((size_t)L.Kpad * L.ldc * elem);
    L.off_igp = ws.plan(sizeof(int) * (size_t)p->ngpown);
    L.off_part = ws.plan(sizeof(double2) * 2 * (size_t)p->nfreq * L.tiles_m * L.tiles_n);
    return L;
}

int check_problem(const fk_gpp_problem *p, const fk_options *o)
{
    if (!p || !o) { set_error("gpp_kernel: NULL problem or options"); return FK_ERR_ARG; }
    if (p->number_bands < 0 || p->nvband < 0 || p->nvband > p->number_bands || p->ncouls <= 0 ||
        p->ngpown <= 0 || p->nfreq < 1 || p->nfreq > kMaxFreq) {
        set_error("gpp_kernel: bad sizes (nb=%d nv=%d ncouls=%d ngpown=%d nfreq=%d; need 0<=nv<=nb, "
                  "ncouls>0, ngpown>0, 1<=nfreq<=%d)", p->number_bands, p->nvband, p->ncouls, p->ngpown,
                  p->nfreq, kMaxFreq);
        return FK_ERR_ARG;
    }
    if (!p->wx || !p->inv_igp_index || !p->indinv || !p->vcoul || !p->aqsmtemp || !p->aqsntemp ||
        !p->I_eps_array || !p->wtilde_array) {
        set_error("gpp_kernel: NULL input array");
        return FK_ERR_ARG;
    }
    if (o->memory != FK_MEM_HOST && o->memory != FK_MEM_DEVICE) { set_error("gpp_kernel: bad memory kind"); return FK_ERR_ARG; }
    if (o->precision != FK_PREC_FP64 && o->precision != FK_PREC_MIXED) { set_error("gpp_kernel: bad precision"); return FK_ERR_ARG; }
    if (o->config < 0 || o->config > kNumCfgs) { set_error("gpp_kernel: config %d out of range", o->config); return FK_ERR_ARG; }
    if ((int64_t)p->ngpown * p->ncouls > (int64_t)INT32_MAX * 8) { set_error("gpp_kernel: problem too large"); return FK_ERR_ARG; }
    return FK_OK;
}

template <class E>
int launch_gpp(const fk_gpp_problem *p, const fk_options *o, const Layout &L, Scratch &ws,
               const std::vector<int32_t> &cols, const E *aqsm, const E *aqsn, const E *ieps, const E *wt,
               const double *vcoul, double2 *d_ach, double2 *d_asx, hipStream_t s)
{
    const int C = p->ncouls, G = p->ngpown;
    E *c = static_cast<E *>(ws.at(L.off_c));
    int *igp = static_cast<int *>(ws.at(L.off_igp));
    double2 *part = static_cast<double2 *>(ws.at(L.off_part));

    for (int m_begin = 0; m_begin < L.ldc; m_begin += kColChunk) {
        CoefArgs<E> ca;
        ca.aqsm = aqsm; ca.vcoul = vcoul; ca.c = c; ca.igp_out = igp;
        ca.nb = p->number_bands; ca.C = C; ca.G = G; ca.Kpad = L.Kpad; ca.ldc = L.ldc;
        ca.m_begin = m_begin;
        ca.m_count = std::min(kColChunk, L.ldc - m_begin);
        for (int mm = 0; mm < kColChunk; ++mm) {
            int m = m_begin + mm;
            ca.igp[mm] = (mm < ca.m_count && m < G) ? cols[m] : -1;
        }
        int64_t n = (int64_t)L.Kpad * ca.m_count;
        int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(gpp_coef_kernel<E>, dim3(blocks), dim3(256), 0, s, ca);
        FK_HIP(hipGetLastError());
    }

    GppArgs<E> a;
    a.c = c; a.aqsn = aqsn; a.I_eps = ieps; a.wtilde = wt; a.igp = igp; a.partials = part;
    a.nb = p->number_bands; a.nv = p->nvband; a.C = C; a.G = G; a.nw = p->nfreq; a.ldc = L.ldc; a.Kpad = L.Kpad;
    a.tiles_m = L.tiles_m; a.tiles_n = L.tiles_n;
    a.limittwo = p->limittwo; a.limitone = p->limitone; a.to1 = p->to1; a.mask_on = o->mask ? 1 : 0;
    for (int iw = 0; iw < kMaxFreq; ++iw) {
        double w = iw < p->nfreq ? p->wx[iw] : 0.0;
        double w2 = w * w;
        a.pc.wx[iw] = w;
        a.pc.wx2[iw] = w2;
        a.pc.rden1[iw] = 1.0 / (w2 * w2);
    }
    const int tiles = L.tiles_m * L.tiles_n;
    KernelTimer timer(o->kernel_ms, s);
    FK_TRY(timer.start());
    switch (L.cfg) {
    case 0: hipLaunchKernelGGL((gpp_fused_kernel<E, 64, 128, 16>), dim3(tiles), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((gpp_fused_kernel<E, 32, 64, 16>), dim3(tiles), dim3(256), 0, s, a); break;
    }
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());
    hipLaunchKernelGGL(gpp_reduce_kernel, dim3(2 * p->nfreq), dim3(256), 0, s, part, tiles, p->nfreq,
                       2.0 * p->occ, d_ach, d_asx);
    FK_HIP(hipGetLastError());
    return timer.finish();
}

// Host-memory path: upload, run, download (synchronous).
template <class E>
int run_host(const fk_gpp_problem *p, const fk_options *o, const Layout &L, Scratch &ws,
             const std::vector<int32_t> &cols, double2 *ach, double2 *asx, hipStream_t s)
{
    const size_t nbC = (size_t)p->number_bands * p->ncouls, GC = (size_t)p->ngpown * p->ncouls;
    E *dm = nullptr, *dn = nullptr, *die = nullptr, *dwt = nullptr;
    double *dv = nullptr;
    double2 *dout = nullptr;
    int rc = FK_OK;
    auto alloc = [&](void **ptr, size_t bytes) {
        if (rc != FK_OK) return;
        if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) { set_error("hipMalloc(%zu) failed", bytes); rc = FK_ERR_NOMEM; }
    };
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        if (rc != FK_OK || bytes == 0) return;
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) { set_error("upload: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
    };
    alloc((void **)&dm, nbC * sizeof(E));
    alloc((void **)&dn, nbC * sizeof(E));
    alloc((void **)&die, GC * sizeof(E));
    alloc((void **)&dwt, GC * sizeof(E));
    alloc((void **)&dv, sizeof(double) * p->ncouls);
    alloc((void **)&dout, sizeof(double2) * 2 * p->nfreq);
    copy(dm, p->aqsmtemp, nbC * sizeof(E));
    copy(dn, p->aqsntemp, nbC * sizeof(E));
    copy(die, p->I_eps_array, GC * sizeof(E));
    copy(dwt, p->wtilde_array, GC * sizeof(E));
    copy(dv, p->vcoul, sizeof(double) * p->ncouls);
    if (rc == FK_OK) rc = launch_gpp<E>(p, o, L, ws, cols, dm, dn, die, dwt, dv, dout, dout + p->nfreq, s);
    if (rc == FK_OK) {
        hipError_t e = hipMemcpyAsync(ach, dout, sizeof(double2) * p->nfreq, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(asx, dout + p->nfreq, sizeof(double2) * p->nfreq, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error("download: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
    } else {
        (void)hipStreamSynchronize(s);
    }
    for (void *ptr : {(void *)dm, (void *)dn, (void *)die, (void *)dwt, (void *)dv, (void *)dout})
        if (ptr) (void)hipFree(ptr);
    return rc;
}

}  // namespace
}  // namespace fk

using namespace fk;

extern "C" {

size_t gpp_kernel_workspace_size(const fk_gpp_problem *p, const fk_options *o)
{
    fk_options d = fk_default_options();
    if (!o) o = &d;
    if (check_problem(p, o) != FK_OK) return 0;
    Scratch ws;
    plan_layout(p, o, ws, o->precision == FK_PREC_MIXED ? sizeof(float2) : sizeof(double2));
    return ws.total();
}

int gpp_kernel(const fk_gpp_problem *p, const fk_options *o, fk_cdouble *achtemp, fk_cdouble *asxtemp,
               fk_cdouble *acht_n1_loc)
{
    clear_error();
    FK_TRY(check_problem(p, o));
    if (!achtemp || !asxtemp) { set_error("gpp_kernel: NULL output"); return FK_ERR_ARG; }
    if (acht_n1_loc) {
        set_error("gpp_kernel: acht_n1_loc is not available in this build");
        return FK_ERR_UNSUPPORTED;
    }
    std::vector<int32_t> igp;
    FK_TRY(resolve_columns(p->ncouls, p->ngpown, p->inv_igp_index, p->indinv, igp));

    int prev_dev = -1;
    if (o->memory == FK_MEM_HOST && o->device >= 0) {
        FK_HIP(hipGetDevice(&prev_dev));
        FK_HIP(hipSetDevice(o->device));
    }
    hipStream_t s = static_cast<hipStream_t>(o->stream);
    const bool mixed = o->precision == FK_PREC_MIXED;
    int rc;
    {
        Scratch ws;
        Layout L = plan_layout(p, o, ws, mixed ? sizeof(float2) : sizeof(double2));
        rc = ws.acquire(o->workspace, o->workspace_bytes, s);
        double2 *ach = reinterpret_cast<double2 *>(achtemp), *asx = reinterpret_cast<double2 *>(asxtemp);
        if (rc == FK_OK) {
            if (o->memory == FK_MEM_DEVICE) {
                if (mixed)
                    rc = launch_gpp<float2>(p, o, L, ws, igp, (const float2 *)p->aqsmtemp, (const float2 *)p->aqsntemp,
                                            (const float2 *)p->I_eps_array, (const float2 *)p->wtilde_array, p->vcoul,
                                            ach, asx, s);
                else
                    rc = launch_gpp<double2>(p, o, L, ws, igp, (const double2 *)p->aqsmtemp, (const double2 *)p->aqsntemp,
                                             (const double2 *)p->I_eps_array, (const double2 *)p->wtilde_array,
                                             p->vcoul, ach, asx, s);
            } else {
                rc = mixed ? run_host<float2>(p, o, L, ws, igp, ach, asx, s)
                       : run_host<double2>(p, o, L, ws, igp, ach, asx, s);
            }
        }
    }
    if (prev_dev >= 0) (void)hipSetDevice(prev_dev);
    return rc;
}

}  // extern "C"
