// fk_ff.hip -- FF frequency contraction on MI355X (gfx950).
//
// The Fortran original contracts (ffkernel.deslippe.best.f90.back1:402-418)
//     tempz(my_igp, ifreq, n1) = sum_ig I_epsR(ig, my_igp, ifreq) aqsntemp(ig, n1)
//         (zgemm('t','n', ngpown*nfreq, number_bands, ncouls, ...), :404)
//     schdt_matrix(ifreq, n1) += sum_my_igp tempz(my_igp, ifreq, n1) conj(aqsmtemp(igp, n1))
// and the static CH sum (ffkernel.deslippe.best.f90:251-267; C++
// FFKer_cori/ffKernel.cpp:107-131) is the vcoul-weighted contraction at
// one frequency.  In C order, with X = I_eps viewed as [nfreq*ngpown][ncouls]:
//     T[n1][m] = sum_ig aqsn[n1][ig] X[m][ig]            (GEMM, K = ncouls)
//     schDt[n1][f] = sum_my_igp cf[n1][my_igp] T[n1][f*G + my_igp]
//     schDv[n1][f] = sum_my_igp vg[my_igp] cf[n1][my_igp] T[n1][f*G + my_igp]
//     achsDtemp   = -0.5 sum_n1 schDv[n1][achs_freq]
// cf[n1][my_igp] = conj(aqsm[n1][igp]) and vg[my_igp] = vcoul[igp] (zero for a
// skipped column).  tempz is never stored.
//
// Kernels: ff_coef_kernel (cf, vg), ff_fused_kernel (MFMA GEMM tile with
// fragments loaded straight from global memory + row contraction of the
// tile -> per-tile partials), ff_reduce_kernel (fixed-order sums).
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>

namespace fk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int kXcds = 8;

struct FfArgs {
    const double2 *aqsn;    // [nb][C]
    const double2 *X;       // [nfreq*G][C]
    const double2 *cf;      // [nb][G]  conj(aqsm[n1][igp])
    const double *vg;       // [G]      vcoul[igp]
    double2 *partials;      // [nfreq][mtiles][nb][2] (schDt part, schDv part)
    int nb, C, G, nfreq, nks;
    int tiles_m, tiles_n, mtiles;   // tiles_n = nfreq * mtiles
    int stagger_lo, stagger_hi, stagger_ticks;
    int ksplit;             // ig ranges per column block (1: the whole range)
    int kper;               // ff_fused_kernel with KSPLIT: k-steps per range (a multiple of P)
};

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4 mfma_nega(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

// Workgroup b -> tile (tm over bands fastest within an XCD's range, so the
// band tiles that share one block of X run back to back on one XCD).
__device__ __forceinline__ void tile_coords(int b, int tiles_m, int total, int &tm, int &tn)
{
    int q = total / kXcds, r = total % kXcds, x = b % kXcds, i = b / kXcds;
    int t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    tm = t % tiles_m;
    tn = t / tiles_m;
}

__global__ void ff_coef_kernel(const double2 *aqsm, const double *vcoul, const int *igp_in, int nb, int C, int G,
                               double2 *cf, double *vg)
{
    int64_t n = (int64_t)nb * G;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        int n1 = (int)(idx / G), m = (int)(idx % G);
        int g = igp_in[m];
        double2 v = make_double2(0.0, 0.0);
        if (g >= 0) {
            double2 am = aqsm[(size_t)n1 * C + g];
            v = make_double2(am.x, -am.y);
        }
        cf[idx] = v;
        if (n1 == 0) vg[m] = g >= 0 ? vcoul[g] : 0.0;
    }
}

template <int MB, int NB>
struct Frags {
    double2 a[MB], b[NB];
};

// 16x16x4 FP64 MFMA: lane l supplies A[l%16][l/16] (here aqsn[n1 = row l%16]
// [ig = l/16]) and B[l/16][l%16] (X[m = col l%16][ig = l/16]); accumulator
// register r of lane l holds T[4r + l/16][l%16].
// KSPLIT: the ig range is split into ksplit parts of kper k-steps, one per
// workgroup (column tile index = column block * ksplit + part), for problems
// with too few tiles to fill the GPU; each part writes its own partial.
template <int MB, int NB, int P, int WGM, int WGN, int MINW, bool KSPLIT>
__global__ __launch_bounds__(64 * WGM * WGN, MINW) void ff_fused_kernel(FfArgs a)
{
    constexpr int NWAVES = WGM * WGN;
    constexpr int WM = MB * 16, WN = NB * 16, BM = WM * WGM, BN = WN * WGN;
    __shared__ double2 red[NWAVES][2 * WM];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    if ((int)blockIdx.x >= a.stagger_lo && (int)blockIdx.x < a.stagger_hi) {
        const uint64_t t0 = wall_clock64();
        while (wall_clock64() - t0 < (uint64_t)a.stagger_ticks) __builtin_amdgcn_s_sleep(32);
    }
    int tm, tn;
    tile_coords(blockIdx.x, a.tiles_m, a.tiles_m * a.tiles_n, tm, tn);
    const int f = tn / a.mtiles, mt = tn % a.mtiles;
    const int kpart = KSPLIT ? mt % a.ksplit : 0;
    const int s_beg = KSPLIT ? kpart * a.kper : 0;
    const int s_end = KSPLIT ? min(a.nks, s_beg + a.kper) : a.nks;
    const int rowbase = tm * BM + wm * WM;            // band n1
    const int colbase = (KSPLIT ? mt / a.ksplit : mt) * BN + wn * WN;   // my_igp within frequency f
    const int kl = lane >> 4, cl = lane & 15;

    // the complex product as three real GEMMs (the 3M method; Higham, SIAM
    // J. Matrix Anal. Appl. 13(3), 1992): T1 = (ar + ai) xr, T2 = ar (xi - xr),
    // T3 = ai (xr + xi); re T = T1 - T3, im T = T1 + T2 (acc_re, acc_im,
    // acc_t3 hold T1, T2, T3): 12 MFMAs per k-step where four real products
    // take 16
    d4 acc_re[MB][NB], acc_im[MB][NB], acc_t3[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            acc_re[i][j] = d4{0, 0, 0, 0};
            acc_im[i][j] = d4{0, 0, 0, 0};
            acc_t3[i][j] = d4{0, 0, 0, 0};
        }

    // row / column addresses clamped into the arrays; rows past nb and
    // columns past G are dropped in the epilogue, k past C is zeroed
    const double2 *pa[MB];
    const double2 *pb[NB];
#pragma unroll
    for (int i = 0; i < MB; ++i) pa[i] = a.aqsn + (size_t)min(rowbase + i * 16 + cl, a.nb - 1) * a.C;
#pragma unroll
    for (int j = 0; j < NB; ++j)
        pb[j] = a.X + ((size_t)f * a.G + min(colbase + j * 16 + cl, a.G - 1)) * a.C;
    auto load = [&](Frags<MB, NB> &fr, int s) {
        s = min(s, s_end - 1);
        const int k = 4 * s + kl;
        const int kc = min(k, a.C - 1);
        const bool kok = k < a.C;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            double2 v = pa[i][kc];
            fr.a[i] = kok ? v : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) fr.b[j] = pb[j][kc];
    };
    auto mma = [&](const Frags<MB, NB> &fr) {
        double as[MB], bd[NB], bs[NB];
#pragma unroll
        for (int i = 0; i < MB; ++i) as[i] = fr.a[i].x + fr.a[i].y;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            bd[j] = fr.b[j].y - fr.b[j].x;
            bs[j] = fr.b[j].x + fr.b[j].y;
        }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                acc_re[i][j] = mfma(as[i], fr.b[j].x, acc_re[i][j]);
                acc_im[i][j] = mfma(fr.a[i].x, bd[j], acc_im[i][j]);
                acc_t3[i][j] = mfma(fr.a[i].y, bs[j], acc_t3[i][j]);
            }
    };
    Frags<MB, NB> ring[P];
#pragma unroll
    for (int q = 0; q < P; ++q) load(ring[q], s_beg + q);
    for (int s0 = s_beg; s0 < s_end; s0 += P) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            mma(ring[q]);
            load(ring[q], s0 + q + P);
        }
    }

    // ---- epilogue: contract each row with cf (and vg cf) over the tile's columns
    double2 rt[MB][4], rv[MB][4];     // per row (i, r): schDt and schDv partials
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) { rt[i][r] = make_double2(0.0, 0.0); rv[i][r] = make_double2(0.0, 0.0); }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int col = colbase + j * 16 + cl;
        const bool cok = col < a.G;
        const double vgc = a.vg[min(col, a.G - 1)];
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rowbase + i * 16 + 4 * r + kl;
                const double2 c = a.cf[(size_t)min(row, a.nb - 1) * a.G + min(col, a.G - 1)];
                const double2 t = make_double2(acc_re[i][j][r] - acc_t3[i][j][r], acc_re[i][j][r] + acc_im[i][j][r]);
                const double2 p = cmul(t, c);
                if (cok) {
                    rt[i][r] = cadd(rt[i][r], p);
                    rv[i][r] = cadd(rv[i][r], make_double2(p.x * vgc, p.y * vgc));
                }
            }
    }
    // sum over the 16 lanes that hold one row (fixed xor order), then over
    // the waves of the workgroup that share the rows (fixed order)
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                rt[i][r].x += __shfl_xor(rt[i][r].x, off, 64);
                rt[i][r].y += __shfl_xor(rt[i][r].y, off, 64);
                rv[i][r].x += __shfl_xor(rv[i][r].x, off, 64);
                rv[i][r].y += __shfl_xor(rv[i][r].y, off, 64);
            }
        }
    for (int w = 0; w < WGN; ++w) {
        if (wn == w && cl == 0) {
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int lr = i * 16 + 4 * r + kl;    // row within the wave tile
                    double2 *dt = &red[wm][2 * lr];
                    if (w == 0) { dt[0] = rt[i][r]; dt[1] = rv[i][r]; }
                    else { dt[0] = cadd(dt[0], rt[i][r]); dt[1] = cadd(dt[1], rv[i][r]); }
                }
        }
        __syncthreads();
    }
    // one row per thread: write this tile's partial for (f, mt, n1)
    for (int t = tid; t < BM; t += 64 * NWAVES) {
        const int w = t / WM, lr = t % WM;
        const int n1 = tm * BM + t;
        if (n1 < a.nb) {
            double2 *dst = a.partials + (((size_t)f * a.mtiles + mt) * a.nb + n1) * 2;
            dst[0] = red[w][2 * lr];
            dst[1] = red[w][2 * lr + 1];
        }
    }
}

// Few bands (at most 16 rows per tile): the contraction is bound by reading
// X once, so X and aqsn pass through LDS in chunks of 16 ig, loaded with
// full-width coalesced reads (one 256-byte row segment per 16 lanes), double
// buffered, and the MFMA fragments come from LDS (row pitch 17 double2, so
// the 16 rows of one 16-lane read pass fall in distinct banks).  Workgroup:
// 16 bands x 64 columns of one frequency, four waves of one 16x16 block.
// The ig range may be split over ksplit workgroups (more workgroups, even
// load over the CUs); each writes its own partial (column tile index
// mt = column block * ksplit + part), which the reduction sums.
// Same partials as ff_fused_kernel.
constexpr int kFbBM = 16, kFbBN = 64, kFbKC = 16, kFbPitch = kFbKC + 1;

__global__ __launch_bounds__(256, 2) void ff_fewband_kernel(FfArgs a)
{
    __shared__ double2 xs[2][kFbBN][kFbPitch];
    __shared__ double2 as[2][kFbBM][kFbPitch];
    __shared__ double2 red[4][2 * kFbBM];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int kl = lane >> 4, cl = lane & 15;
    const int tm = blockIdx.x % a.tiles_m, tn = blockIdx.x / a.tiles_m;
    const int f = tn / a.mtiles, mt = tn % a.mtiles;
    const int kpart = mt % a.ksplit;
    const int rowbase = tm * kFbBM, colbase = (mt / a.ksplit) * kFbBN;
    const int nall = a.nks / 4;
    const int c0 = (int)((int64_t)kpart * nall / a.ksplit), c1 = (int)((int64_t)(kpart + 1) * nall / a.ksplit);
    const int nchunks = c1 - c0;
    // this thread's global loads of a chunk: X element e = 256 k + tid
    // (row e / 16, ig e % 16), k = 0..3; aqsn element tid (row tid / 16)
    const double2 *xrow[4];
    int xr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = 256 * k + tid;
        xr[k] = e >> 4;
        xrow[k] = a.X + ((size_t)f * a.G + min(colbase + xr[k], a.G - 1)) * a.C;
    }
    const int ar = tid >> 4, ai = tid & 15;
    const double2 *arow = a.aqsn + (size_t)min(rowbase + ar, a.nb - 1) * a.C;
    // two chunks in flight in registers (sets R0, R1), one in each LDS buffer
    struct Regs { double2 x[4], a; };
    auto fetch = [&](int c, Regs &r) {
        const int ig = (c0 + c) * kFbKC + ai;
        const bool ok = ig < a.C;
        const int igc = min(ig, a.C - 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double2 v = xrow[k][igc];
            r.x[k] = ok ? v : make_double2(0.0, 0.0);
        }
        const double2 v = arow[igc];
        r.a = ok ? v : make_double2(0.0, 0.0);
    };
    auto stash = [&](int b, const Regs &r) {
#pragma unroll
        for (int k = 0; k < 4; ++k) xs[b][xr[k]][ai] = r.x[k];
        as[b][ar][ai] = r.a;
    };
    d4 acc_re = d4{0, 0, 0, 0}, acc_im = d4{0, 0, 0, 0};
    auto compute = [&](int b) {
#pragma unroll
        for (int s = 0; s < kFbKC / 4; ++s) {
            const double2 av = as[b][cl][4 * s + kl];
            const double2 bv = xs[b][w * 16 + cl][4 * s + kl];
            acc_re = mfma(av.x, bv.x, acc_re);
            acc_im = mfma(av.x, bv.y, acc_im);
            acc_re = mfma_nega(av.y, bv.y, acc_re);
            acc_im = mfma(av.y, bv.x, acc_im);
        }
    };
    Regs r0, r1;
    if (nchunks < 1) {   // an empty ig range still writes its (zero) partial
        r0.a = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 4; ++k) r0.x[k] = make_double2(0.0, 0.0);
    } else {
        fetch(0, r0);
    }
    if (1 < nchunks) fetch(1, r1);
    stash(0, r0);
    __syncthreads();
    for (int c = 0; c < nchunks; c += 2) {
        // chunk c in buffer 0, chunk c+1 in r1
        if (c + 2 < nchunks) fetch(c + 2, r0);
        compute(0);
        if (c + 1 < nchunks) stash(1, r1);
        __syncthreads();
        if (c + 1 >= nchunks) break;
        // chunk c+1 in buffer 1, chunk c+2 in r0
        if (c + 3 < nchunks) fetch(c + 3, r1);
        compute(1);
        if (c + 2 < nchunks) stash(0, r0);
        __syncthreads();
    }
    // contract each band with cf (and vg cf) over the wave's 16 columns:
    // register r of lane l holds T[band 4 r + l / 16][column l % 16]
    const int col = colbase + w * 16 + cl;
    const bool cok = col < a.G;
    const double vgc = a.vg[min(col, a.G - 1)];
    double2 rt[4], rv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = rowbase + 4 * r + kl;
        const double2 cfv = a.cf[(size_t)min(row, a.nb - 1) * a.G + min(col, a.G - 1)];
        const double2 p = cmul(make_double2(acc_re[r], acc_im[r]), cfv);
        rt[r] = cok ? p : make_double2(0.0, 0.0);
        rv[r] = cok ? make_double2(p.x * vgc, p.y * vgc) : make_double2(0.0, 0.0);
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            rt[r].x += __shfl_xor(rt[r].x, off, 64);
            rt[r].y += __shfl_xor(rt[r].y, off, 64);
            rv[r].x += __shfl_xor(rv[r].x, off, 64);
            rv[r].y += __shfl_xor(rv[r].y, off, 64);
        }
    }
    if (cl == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            red[w][2 * (4 * r + kl)] = rt[r];
            red[w][2 * (4 * r + kl) + 1] = rv[r];
        }
    }
    __syncthreads();
    if (tid < 2 * kFbBM) {
        const int lr = tid >> 1, which = tid & 1;
        double2 sum = red[0][tid];
#pragma unroll
        for (int q = 1; q < 4; ++q) sum = cadd(sum, red[q][tid]);
        const int n1 = rowbase + lr;
        if (n1 < a.nb) a.partials[(((size_t)f * a.mtiles + mt) * a.nb + n1) * 2 + which] = sum;
    }
}

// schDt[n1][f], schDv[n1][f]: sum over the column tiles in order.  Thread
// (f, n1), n1 fastest, so a warp's partials are contiguous; the row sums of
// frequency achs_freq (unscaled schDv) also go to achs_rows[n1].
__global__ void ff_reduce_kernel(const double2 *partials, int nb, int nfreq, int mtiles, double vscale,
                                 double2 *schDt, double2 *schDv, int achs_freq, double2 *achs_rows)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * nfreq) return;
    const int f = idx / nb, n1 = idx % nb;
    double2 st = make_double2(0.0, 0.0), sv = st;
    const double2 *p = partials + ((size_t)f * mtiles * nb + n1) * 2;
    const size_t step = (size_t)nb * 2;
#pragma unroll 8
    for (int mt = 0; mt < mtiles; ++mt) {
        st = cadd(st, p[mt * step]);
        sv = cadd(sv, p[mt * step + 1]);
    }
    const int o = n1 * nfreq + f;
    if (schDt) schDt[o] = st;
    if (schDv) schDv[o] = make_double2(vscale * sv.x, vscale * sv.y);
    if (achs_rows && f == achs_freq) achs_rows[n1] = sv;
}

// achsDtemp = -0.5 sum_n1 schDv[n1][achs_freq], one workgroup, fixed order
__global__ __launch_bounds__(256) void ff_achs_kernel(const double2 *achs_rows, int nb, double2 *achs)
{
    __shared__ double sre[256], sim[256];
    double re = 0.0, im = 0.0;
    for (int n1 = threadIdx.x; n1 < nb; n1 += 256) {
        re += achs_rows[n1].x;
        im += achs_rows[n1].y;
    }
    sre[threadIdx.x] = re;
    sim[threadIdx.x] = im;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sre[threadIdx.x] += sre[threadIdx.x + w];
            sim[threadIdx.x] += sim[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *achs = make_double2(-0.5 * sre[0], -0.5 * sim[0]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct FfCfg { int bm, bn, p, minw; };
// config ids (fk_options.config = id + 1)
//   0: 32x32 waves, 2x2 waves (64x64 tile), ring 3, 2 workgroups/CU
//   1: 32x64 waves, 2x2 waves (64x128 tile), ring 3, 1 workgroup/CU (default:
//      15.4 ms against 16.2 ms for id 0 at 256/4096/4096/32 on MI355X)
//   2: ff_fewband_kernel, 16x64 tile through LDS (chunks of 16 ig = 4
//      k-steps): for 16 bands or fewer, where the product is bound by
//      reading X and a 64-row tile would spend 4x the MFMA work
//      (kFfConfigFewBands)
constexpr FfCfg kFfCfgs[] = {{64, 64, 3, 2}, {64, 128, 3, 1}, {kFbBM, kFbBN, 4, 2}};
static_assert(kFfConfigFewBands == 3, "kFfConfigFewBands names the 16-row tile");
constexpr int kFfDefaultCfg = 1;
constexpr int kNumFfCfgs = sizeof(kFfCfgs) / sizeof(kFfCfgs[0]);

void launch_ff_fused(int cfg, int tiles, hipStream_t s, const FfArgs &a)
{
    const bool ks = a.ksplit > 1;
    switch (cfg) {
    case 0:
        if (ks) hipLaunchKernelGGL((ff_fused_kernel<2, 2, 3, 2, 2, 2, true>), dim3(tiles), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((ff_fused_kernel<2, 2, 3, 2, 2, 2, false>), dim3(tiles), dim3(256), 0, s, a);
        break;
    case 2: hipLaunchKernelGGL(ff_fewband_kernel, dim3(tiles), dim3(256), 0, s, a); break;
    default:
        if (ks) hipLaunchKernelGGL((ff_fused_kernel<2, 4, 3, 2, 2, 1, true>), dim3(tiles), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((ff_fused_kernel<2, 4, 3, 2, 2, 1, false>), dim3(tiles), dim3(256), 0, s, a);
        break;
    }
}

}  // namespace

int ff_fp64_configs() { return kNumFfCfgs; }

FfLayout ff_plan(const fk_ff_problem *p, const fk_options *o, Scratch &ws)
{
    FfLayout L;
    L.split = o->precision == FK_PREC_MIXED;
    L.cfg = o->config > 0 ? o->config - 1 : kFfDefaultCfg;
    if (L.split) {
        ff_split_plan(p->number_bands, p->ncouls, p->ngpown, p->nfreq, o->config, ws, L.sp);
        L.cfg = 0;
        L.tiles_m = L.sp.nmt;
        L.mtiles = L.sp.nkc * L.sp.ngt;
        L.nks = L.sp.nstages;
    } else {
        const FfCfg &c = kFfCfgs[L.cfg];
        L.tiles_m = div_up(std::max(p->number_bands, 1), c.bm);
        L.mtiles = div_up(p->ngpown, c.bn);
        L.nks = div_up(p->ncouls, 4 * c.p) * c.p;
        if (L.cfg == kFfConfigFewBands - 1) {
            // split the ig range until there are about four workgroups per
            // CU, keeping at least two chunks per part
            const int tiles = L.tiles_m * p->nfreq * L.mtiles, chunks = L.nks / 4;
            while (tiles * L.ksplit < 1024 && L.ksplit < 8 && chunks / (2 * L.ksplit) >= 2) L.ksplit *= 2;
            L.mtiles *= L.ksplit;
        } else {
            // too few tiles for two rounds over the CUs (small problems,
            // the per-band row contraction): split the ig range, each part
            // at least 16 ring passes, no part empty
            const int tiles = L.tiles_m * p->nfreq * L.mtiles, rounds = L.nks / c.p;
            int ks = 1;
            while (tiles * ks < 512 * c.minw && ks < 8 && rounds >= 2 * ks * 16) ks *= 2;
            if (ks > 1) {
                const int per = div_up(rounds, ks);
                L.ksplit = div_up(rounds, per);
                L.kper = per * c.p;
                L.mtiles *= L.ksplit;
            }
        }
    }
    L.off_cf = ws.plan(sizeof(double2) * (size_t)std::max(p->number_bands, 1) * p->ngpown);
    L.off_vg = ws.plan(sizeof(double) * (size_t)p->ngpown);
    L.off_igp = ws.plan(sizeof(int) * (size_t)p->ngpown);
    L.off_part = ws.plan(sizeof(double2) * 2 * (size_t)p->nfreq * L.mtiles * std::max(p->number_bands, 1));
    L.off_achs = ws.plan(sizeof(double2) * (size_t)std::max(p->number_bands, 1));
    return L;
}

int ff_check(const fk_ff_problem *p, const fk_options *o)
{
    if (!p || !o) { set_error("ff_kernel: NULL problem or options"); return FK_ERR_ARG; }
    if (p->number_bands < 1 || p->ncouls < 1 || p->ngpown < 1 || p->nfreq < 1) {
        set_error("ff_kernel: bad sizes (nb=%d ncouls=%d ngpown=%d nfreq=%d; all must be >= 1)", p->number_bands,
                  p->ncouls, p->ngpown, p->nfreq);
        return FK_ERR_ARG;
    }
    if (p->ngpown > p->ncouls) {   // FFKer_cori/ffKernel.cpp:320-325
        set_error("ff_kernel: ngpown (%d) cannot be greater than ncouls (%d)", p->ngpown, p->ncouls);
        return FK_ERR_ARG;
    }
    if (!p->aqsmtemp || !p->aqsntemp || !p->I_eps_array || !p->vcoul || !p->inv_igp_index || !p->indinv) {
        set_error("ff_kernel: NULL input array");
        return FK_ERR_ARG;
    }
    if (p->achs_freq < 0 || p->achs_freq >= p->nfreq) {
        set_error("ff_kernel: achs_freq %d outside [0, nfreq=%d)", p->achs_freq, p->nfreq);
        return FK_ERR_ARG;
    }
    if (o->memory != FK_MEM_HOST && o->memory != FK_MEM_DEVICE) { set_error("ff_kernel: bad memory kind"); return FK_ERR_ARG; }
    if (o->precision != FK_PREC_FP64 && o->precision != FK_PREC_MIXED) {
        set_error("ff_kernel: precision must be FK_PREC_FP64 or FK_PREC_MIXED");
        return FK_ERR_ARG;
    }
    const int ncfg = o->precision == FK_PREC_MIXED ? ff_split_configs() : kNumFfCfgs;
    if (o->config < 0 || o->config > ncfg) { set_error("ff_kernel: config %d out of range [0, %d]", o->config, ncfg); return FK_ERR_ARG; }
    return FK_OK;
}

// The FP64 MFMA GEMM with its row contraction -> partials.
static int ff_fp64_gemm(const FfLayout &L, const double2 *aqsn, const double2 *X, const double2 *cf, const double *vg,
                        double2 *part, int nb, int C, int G, int nf, KernelTimer &timer, hipStream_t s)
{
    FfArgs a;
    a.aqsn = aqsn; a.X = X; a.cf = cf; a.vg = vg; a.partials = part;
    a.nb = nb; a.C = C; a.G = G; a.nfreq = nf; a.nks = L.nks;
    a.tiles_m = L.tiles_m; a.mtiles = L.mtiles; a.tiles_n = nf * L.mtiles;
    a.ksplit = L.ksplit;
    a.kper = L.kper;
    const int tiles = a.tiles_m * a.tiles_n;
    a.stagger_lo = a.stagger_hi = a.stagger_ticks = 0;
    if (kFfCfgs[L.cfg].minw == 2 && L.cfg != kFfConfigFewBands - 1 && L.ksplit == 1) {   // MFMA-bound tiles only
        int dev = 0, ncu = 0;
        FK_HIP(hipGetDevice(&dev));
        FK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (tiles > 2 * ncu) {
            const double tile_cycles = 8.0 * kFfCfgs[L.cfg].bm * kFfCfgs[L.cfg].bn * 4.0 * L.nks / 64.0;
            a.stagger_lo = ncu;
            a.stagger_hi = 2 * ncu;
            a.stagger_ticks = (int)std::min(0.5 * tile_cycles / 24.0, 1.0e6);
        }
    }
    FK_TRY(timer.start());
    launch_ff_fused(L.cfg, tiles, s, a);
    FK_HIP(hipGetLastError());
    return timer.stop();
}

int ff_launch(const fk_ff_problem *p, const fk_options *o, const FfLayout &L, Scratch &ws,
              const std::vector<int32_t> &cols, const double2 *aqsm, const double2 *aqsn, const double2 *X,
              const double *vcoul, double2 *schDt, double2 *schDv, double2 *achs, double vscale, hipStream_t s,
              bool upload_cols)
{
    const int nb = p->number_bands, C = p->ncouls, G = p->ngpown, nf = p->nfreq;
    double2 *cf = static_cast<double2 *>(ws.at(L.off_cf));
    double *vg = static_cast<double *>(ws.at(L.off_vg));
    int *igp = static_cast<int *>(ws.at(L.off_igp));
    double2 *part = static_cast<double2 *>(ws.at(L.off_part));
    if (upload_cols) FK_HIP(hipMemcpyAsync(igp, cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice, s));
    {
        int64_t n = (int64_t)nb * G;
        int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(ff_coef_kernel, dim3(blocks), dim3(256), 0, s, aqsm, vcoul, igp, nb, C, G, cf, vg);
        FK_HIP(hipGetLastError());
    }
    KernelTimer timer(o->kernel_ms, s);
    if (L.split)
        FK_TRY(ff_split_launch(L.sp, ws, aqsn, X, cf, vg, part, nb, C, G, nf, timer, s));
    else
        FK_TRY(ff_fp64_gemm(L, aqsn, X, cf, vg, part, nb, C, G, nf, timer, s));
    double2 *achs_rows = achs ? static_cast<double2 *>(ws.at(L.off_achs)) : nullptr;
    hipLaunchKernelGGL(ff_reduce_kernel, dim3(div_up(nb * nf, 256)), dim3(256), 0, s, part, nb, nf, L.mtiles, vscale,
                       schDt, schDv, p->achs_freq, achs_rows);
    FK_HIP(hipGetLastError());
    if (achs) {
        hipLaunchKernelGGL(ff_achs_kernel, dim3(1), dim3(256), 0, s, achs_rows, nb, achs);
        FK_HIP(hipGetLastError());
    }
    return timer.finish();
}

namespace {
fk_ff_problem rows_problem(int nb, int C, int G, const double2 *aqsm, const double2 *aqsn, const double2 *X,
                           const double *vcoul)
{
    fk_ff_problem p;
    p.number_bands = nb; p.ncouls = C; p.ngpown = G; p.nfreq = 1;
    p.aqsmtemp = aqsm; p.aqsntemp = aqsn; p.I_eps_array = X; p.vcoul = vcoul;
    p.inv_igp_index = nullptr; p.indinv = nullptr; p.achs_freq = 0;
    return p;
}
}  // namespace

size_t ff_rows_weighted_bytes(int nb, int C, int G)
{
    const fk_ff_problem p = rows_problem(nb, C, G, nullptr, nullptr, nullptr, nullptr);
    fk_options o = fk_default_options();
    o.memory = FK_MEM_DEVICE;
    Scratch ws;
    ff_plan(&p, &o, ws);
    return ws.total();
}

int ff_rows_weighted(int nb, int C, int G, const double2 *aqsm, const double2 *aqsn, const double2 *X,
                     const double *vcoul, const std::vector<int32_t> &cols, double vscale, double2 *out,
                     void *wsp, size_t ws_bytes, hipStream_t s)
{
    const fk_ff_problem p = rows_problem(nb, C, G, aqsm, aqsn, X, vcoul);
    fk_options o = fk_default_options();
    o.memory = FK_MEM_DEVICE;
    Scratch ws;
    FfLayout L = ff_plan(&p, &o, ws);
    FK_TRY(ws.acquire(wsp, ws_bytes, s));
    return ff_launch(&p, &o, L, ws, cols, aqsm, aqsn, X, vcoul, nullptr, out, nullptr, vscale, s);
}

}  // namespace fk

using namespace fk;

extern "C" {

size_t ff_kernel_workspace_size(const fk_ff_problem *p, const fk_options *o)
{
    fk_options d = fk_default_options();
    if (!o) o = &d;
    if (ff_check(p, o) != FK_OK) return 0;
    Scratch ws;
    ff_plan(p, o, ws);
    return ws.total();
}

int ff_kernel(const fk_ff_problem *p, const fk_options *o, fk_cdouble *schDt, fk_cdouble *schDv,
              fk_cdouble *achsDtemp)
{
    clear_error();
    FK_TRY(ff_check(p, o));
    if (!schDt) { set_error("ff_kernel: NULL schDt"); return FK_ERR_ARG; }
    std::vector<int32_t> cols;
    FK_TRY(resolve_columns(p->ncouls, p->ngpown, p->inv_igp_index, p->indinv, cols));
    int prev_dev = -1;
    if (o->memory == FK_MEM_HOST && o->device >= 0) {
        FK_HIP(hipGetDevice(&prev_dev));
        FK_HIP(hipSetDevice(o->device));
    }
    hipStream_t s = static_cast<hipStream_t>(o->stream);
    const size_t nbC = (size_t)p->number_bands * p->ncouls;
    const size_t XC = (size_t)p->nfreq * p->ngpown * p->ncouls;
    const size_t nout = (size_t)p->number_bands * p->nfreq;
    int rc = FK_OK;
    {
        Scratch ws;
        FfLayout L = ff_plan(p, o, ws);
        rc = ws.acquire(o->workspace, o->workspace_bytes, s);
        if (rc == FK_OK && o->memory == FK_MEM_DEVICE) {
            rc = ff_launch(p, o, L, ws, cols, (const double2 *)p->aqsmtemp, (const double2 *)p->aqsntemp,
                           (const double2 *)p->I_eps_array, p->vcoul, (double2 *)schDt, (double2 *)schDv,
                           (double2 *)achsDtemp, 1.0, s);
        } else if (rc == FK_OK) {
            double2 *dm = nullptr, *dn = nullptr, *dx = nullptr, *dout = nullptr;
            double *dv = nullptr;
            auto alloc = [&](void **ptr, size_t bytes) {
                if (rc != FK_OK) return;
                if (hipMalloc(ptr, bytes) != hipSuccess) { set_error("hipMalloc(%zu) failed", bytes); rc = FK_ERR_NOMEM; }
            };
            auto up = [&](void *dst, const void *src, size_t bytes) {
                if (rc != FK_OK) return;
                hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
                if (e != hipSuccess) { set_error("upload: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
            };
            alloc((void **)&dm, nbC * sizeof(double2));
            alloc((void **)&dn, nbC * sizeof(double2));
            alloc((void **)&dx, XC * sizeof(double2));
            alloc((void **)&dv, sizeof(double) * p->ncouls);
            alloc((void **)&dout, sizeof(double2) * (2 * nout + 1));
            up(dm, p->aqsmtemp, nbC * sizeof(double2));
            up(dn, p->aqsntemp, nbC * sizeof(double2));
            up(dx, p->I_eps_array, XC * sizeof(double2));
            up(dv, p->vcoul, sizeof(double) * p->ncouls);
            if (rc == FK_OK)
                rc = ff_launch(p, o, L, ws, cols, dm, dn, dx, dv, dout, dout + nout, dout + 2 * nout, 1.0, s);
            if (rc == FK_OK) {
                hipError_t e = hipMemcpyAsync(schDt, dout, sizeof(double2) * nout, hipMemcpyDeviceToHost, s);
                if (e == hipSuccess && schDv)
                    e = hipMemcpyAsync(schDv, dout + nout, sizeof(double2) * nout, hipMemcpyDeviceToHost, s);
                if (e == hipSuccess && achsDtemp)
                    e = hipMemcpyAsync(achsDtemp, dout + 2 * nout, sizeof(double2), hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) { set_error("download: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
            } else {
                (void)hipStreamSynchronize(s);
            }
            for (void *ptr : {(void *)dm, (void *)dn, (void *)dx, (void *)dv, (void *)dout})
                if (ptr) (void)hipFree(ptr);
        }
    }
    if (prev_dev >= 0) (void)hipSetDevice(prev_dev);
    return rc;
}

}  // extern "C"
