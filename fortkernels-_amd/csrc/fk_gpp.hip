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
//   gpp_coef_kernel   c[Kpad][ldc] from aqsm / vcoul / the column map, with
//                     the occupied bands first, zero-padded to a k-step
//   gpp_fused_kernel  MFMA GEMM tile with operand fragments loaded straight
//                     from global memory into a register ring (no LDS, no
//                     barriers), the occupied-band partial sum taken as a
//                     register snapshot, then the pole/SX epilogue and a
//                     wave/workgroup reduction -> per-tile partials
//   gpp_reduce_kernel fixed-order sum of the per-tile partials -> outputs
// Every sum is taken in a fixed order, so results are bitwise reproducible.
#include "fk_complex.h"
#include "fk_gpp_common.h"
#include "fk_internal.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <utility>

namespace fk {
namespace {

using namespace gpp;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class E>
struct GppArgs {
    const E *c;            // [Kpad][ldc] zero-padded coefficient matrix
    const E *aqsn;         // [nb][C]
    const E *I_eps;        // [G][C]
    const E *wtilde;       // [G][C]
    const int *igp;        // [G], -1 = skipped column
    double2 *partials;     // [tiles][nw][2] (achtemp part, asxtemp part)
    double2 *achs_part;    // [tiles] static-CH remainder part (ACHS tiles only)
    int nb, nv, C, G, nw, ldc, Kpad;
    int nvpad;             // occupied bands rounded up to a k-step (4)
    int nks;               // k-steps (Kpad / 4)
    int tiles_m, tiles_n, group_m;
    int stagger_lo, stagger_hi;   // workgroups that start late ...
    int stagger_ticks;            // ... by this many wall-clock ticks (100 MHz)
    PoleParams pp;
};

// Loads and stores through a pointer into global memory, as global (not
// flat) instructions.  The single-problem kernel's pointers are kernel
// arguments and known to be global; the batch kernel reads them from its
// problem table, which leaves them generic, and flat instructions take
// 64-bit addresses per lane and also count against the LDS counter.  The
// element is moved as a plain vector: the HIP vector types' copy operations
// would turn the access back into a generic one.
template <class T> struct RawOf { typedef T type; };
template <> struct RawOf<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct RawOf<float2> { typedef float type __attribute__((ext_vector_type(2))); };

// GLOBAL = false: a plain access (the single-problem kernel keeps the code
// the compiler makes for its kernel-argument pointers).
template <bool GLOBAL, class T>
__device__ __forceinline__ T gld(const T *p, size_t i)
{
    if constexpr (!GLOBAL) {
        return p[i];
    } else {
        typedef typename RawOf<T>::type R;
        const R v = *(const __attribute__((address_space(1))) R *)(p + i);
        if constexpr (__is_same(R, T)) {
            return v;
        } else {
            T e;
            e.x = v.x;
            e.y = v.y;
            return e;
        }
    }
}

template <bool GLOBAL>
__device__ __forceinline__ void gst(double *p, size_t i, double v)
{
    if constexpr (!GLOBAL) p[i] = v;
    else *(__attribute__((address_space(1))) double *)(p + i) = v;
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// -a * b + c: the FP64 MFMA's neg modifier on A (the blgp field of the
// builtin); FP32 negates in a VALU op.
__device__ __forceinline__ d4 mfma_nega(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}
__device__ __forceinline__ f4 mfma_nega(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(-a, b, c, 0, 0, 0);
}

template <class E> struct Acc;
template <> struct Acc<double2> { using type = d4; };
template <> struct Acc<float2> { using type = f4; };

// Row of the 16x16 MFMA result held in accumulator register r of lane l
// (column l % 16): FP64 16x16x4 interleaves rows across registers, FP32
// 16x16x4 keeps four consecutive rows per lane group.
template <class E> __device__ __forceinline__ int acc_row(int lane, int r);
template <> __device__ __forceinline__ int acc_row<double2>(int lane, int r) { return 4 * r + (lane >> 4); }
template <> __device__ __forceinline__ int acc_row<float2>(int lane, int r) { return 4 * (lane >> 4) + r; }

template <class E> __device__ __forceinline__ E zero_elem();
template <> __device__ __forceinline__ double2 zero_elem<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero_elem<float2>() { return make_float2(0.f, 0.f); }

// Workgroup b -> logical tile.  Each XCD (every kXcds-th workgroup) runs a
// contiguous range of the logical order, and the logical order walks bands
// of group_m row tiles column by column: the tiles an XCD runs together
// form a block of group_m row tiles x (its workgroups / group_m) column
// tiles, so both the coefficient panels and the aqsn panels they read are
// shared in that XCD's L2.  group_m = tiles_m walks whole columns.
__device__ __forceinline__ void tile_coords(int b, int tiles_m, int tiles_n, int group_m, int &tm, int &tn)
{
    const int total = tiles_m * tiles_n;
    const int q = total / kXcds, r = total % kXcds, x = b % kXcds, i = b / kXcds;
    const int t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    const int band = t / (group_m * tiles_n);
    const int first = band * group_m;
    const int gm = min(tiles_m - first, group_m);
    const int u = t - band * group_m * tiles_n;
    tm = first + u % gm;
    tn = u / gm;
}


// ---------------------------------------------------------------------------
// c[kr][m] = 0.5 vcoul[igp] conj(aqsm[band][igp]) for a kept column m < G,
// zero elsewhere.  Row kr holds band band_of(kr): the occupied bands first
// (kr < nv), zero rows up to nvpad (a multiple of 4), then the unoccupied
// bands; rows past the last band are zero (padding to Kpad).  The resolved
// column map travels in the kernel arguments (a chunk of columns per
// launch), so the call needs no host-to-device copy; the launch also stores
// it for the fused kernel.
// ---------------------------------------------------------------------------
constexpr int kColChunk = 896;

// band of coefficient row kr, or -1 for a zero row
__device__ __forceinline__ int band_of(int kr, int nb, int nv, int nvpad)
{
    if (kr < nv) return kr;
    if (kr < nvpad) return -1;
    int b = kr - nvpad + nv;
    return b < nb ? b : -1;
}

template <class E>
struct CoefArgs {
    const E *aqsm;
    const double *vcoul;
    E *c;
    int *igp_out;
    int nb, nv, nvpad, C, G, Kpad, ldc, m_begin, m_count;
    int igp[kColChunk];
};

template <class E>
__global__ void gpp_coef_kernel(CoefArgs<E> a)
{
    int64_t n = (int64_t)a.Kpad * a.m_count;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        int kr = (int)(idx / a.m_count), mm = (int)(idx % a.m_count), m = a.m_begin + mm;
        int g = m < a.G ? a.igp[mm] : -1;
        if (kr == 0 && m < a.G) a.igp_out[m] = g;
        double2 v = make_double2(0.0, 0.0);
        int band = band_of(kr, a.nb, a.nv, a.nvpad);
        if (band >= 0 && g >= 0) {
            double2 am = widen(a.aqsm[(size_t)band * a.C + g]);
            double s = 0.5 * a.vcoul[g];
            v = make_double2(am.x * s, -am.y * s);
        }
        E *dst = a.c + (size_t)kr * a.ldc + m;
        if constexpr (sizeof(E) == sizeof(double2)) *dst = v;
        else *dst = make_float2((float)v.x, (float)v.y);
    }
}

// ---------------------------------------------------------------------------
// Fused GEMM tile + epilogue.  WGM x WGN waves per workgroup, each wave a
// WM x WN = (16 MB) x (16 NB) tile of (my_igp, ig); MINW workgroups per CU;
// NWC frequencies per epilogue pass.
//
// MFMA 16x16x4 operand layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l is C[acc_row(l,r)][l%16].
// With A = c[k][m] and B = aqsn[k][n] both row-major in k, a lane's fragment
// is one 16-byte element, and the 64 lanes of a wave read four 256-byte row
// segments: fragments are loaded straight from global memory (L1/L2) into a
// ring of P k-steps, so the loop has no LDS traffic and no barrier.
// Complex products use four real MFMAs (re += ar br - ai bi, im += ar bi + ai br).
// ---------------------------------------------------------------------------
template <class E, int MB, int NB>
struct Frags {
    E a[MB], b[NB];
};

// LDS of one workgroup: per-wave reduction slots and the occupied-band
// partial sums, one slot per (block, register, lane)
template <class E, int MB, int NB, int NWAVES, int NWC>
struct TileLds {
    double red[NWAVES][4 * NWC];
    E occ[NWAVES][MB * NB * 4 * 64];
};

// One (tm, tn) tile of one problem: GEMM, epilogue, partials.  Ends with
// a workgroup barrier, so the LDS can be reused for the next tile.
// ACHS: also the static-CH remainder of the tile, -sum B I_eps over its kept
// (my_igp, ig) (gppkernel.deslippe.09.30.2014.f90:192-230; B carries the
// 0.5 vcoul of the coefficients), into achs_part[tile].
// TABLE: the arguments come from the batch kernel's problem table (global
// instructions through gld / gst).
template <class E, int MB, int NB, int P, int WGM, int WGN, int NWC, bool ACHS, bool TABLE = false>
__device__ __forceinline__ void gpp_tile(const GppArgs<E> &a, int tm, int tn, TileLds<E, MB, NB, WGM * WGN, NWC> &sh)
{
    using A4 = typename Acc<E>::type;
    constexpr int NWAVES = WGM * WGN;
    constexpr int WM = MB * 16, WN = NB * 16, BM = WM * WGM, BN = WN * WGN;
    auto &red = sh.red;
    auto &occ_lds = sh.occ;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int m0 = tm * BM, n0 = tn * BN;
    const int rowbase = m0 + wm * WM, colbase = n0 + wn * WN;
    const int nmax = a.C - 1;
    const int kl = lane >> 4, cl = lane & 15;

    // FP64: the complex product as three real GEMMs (the 3M method; Higham,
    // SIAM J. Matrix Anal. Appl. 13(3), 1992):
    //   T1 = (cr + ci) ar,  T2 = cr (ai - ar),  T3 = ci (ar + ai),
    //   re B = T1 - T3,     im B = T1 + T2
    // (acc_re, acc_im, acc_t3 hold T1, T2, T3), 12 MFMAs per k-step where
    // four real products take 16.  FP32 (the mixed tiles) keeps the four.
    constexpr bool K3 = sizeof(E) == sizeof(double2);
    A4 acc_re[MB][NB], acc_im[MB][NB], acc_t3[K3 ? MB : 1][K3 ? NB : 1];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            acc_re[i][j] = A4{0, 0, 0, 0};
            acc_im[i][j] = A4{0, 0, 0, 0};
            if constexpr (K3) acc_t3[i][j] = A4{0, 0, 0, 0};
        }
    // B of accumulator register r of block (i, j)
    auto acc_value = [&](int i, int j, int r) -> double2 {
        if constexpr (K3)
            return make_double2(acc_re[i][j][r] - acc_t3[i][j][r], acc_re[i][j][r] + acc_im[i][j][r]);
        else
            return make_double2((double)acc_re[i][j][r], (double)acc_im[i][j][r]);
    };
    E *occ = &occ_lds[wave][0];
    auto occ_store = [&](bool zero) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double2 b = acc_value(i, j, r);
                    E v;
                    v.x = zero ? 0 : b.x;
                    v.y = zero ? 0 : b.y;
                    occ[((i * NB + j) * 4 + r) * 64 + lane] = v;
                }
    };

    // A fragment of k-step s, block i: c[4s + kl][rowbase + 16 i + cl]
    // (c is zero-padded to Kpad x ldc: always in bounds).  B fragment of
    // k-step s, block j: aqsn[band_of(4s + kl)][colbase + 16 j + cl], with the
    // band and the column clamped into the array (a zero row of c makes the
    // clamped value irrelevant).
    const E *pa = a.c + (size_t)kl * a.ldc + rowbase + cl;
    int colj[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) colj[j] = min(colbase + j * 16 + cl, nmax);
    auto load = [&](Frags<E, MB, NB> &f, int s) {
        s = min(s, a.nks - 1);
        const E *qa = pa + (size_t)(4 * s) * a.ldc;
#pragma unroll
        for (int i = 0; i < MB; ++i) f.a[i] = gld<TABLE>(qa, i * 16);
        const int kr = 4 * s + kl;
        int band = kr < a.nv ? kr : kr - a.nvpad + a.nv;
        band = min(max(band, 0), a.nb - 1);
        const E *qb = a.aqsn + (size_t)band * a.C;
#pragma unroll
        for (int j = 0; j < NB; ++j) f.b[j] = gld<TABLE>(qb, colj[j]);
    };
    auto mma = [&](const Frags<E, MB, NB> &f) {
        if constexpr (K3) {
            double as[MB], bd[NB], bs[NB];
#pragma unroll
            for (int i = 0; i < MB; ++i) as[i] = f.a[i].x + f.a[i].y;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                bd[j] = f.b[j].y - f.b[j].x;
                bs[j] = f.b[j].x + f.b[j].y;
            }
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc_re[i][j] = mfma(as[i], f.b[j].x, acc_re[i][j]);
                    acc_im[i][j] = mfma(f.a[i].x, bd[j], acc_im[i][j]);
                    acc_t3[i][j] = mfma(f.a[i].y, bs[j], acc_t3[i][j]);
                }
        } else {
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc_re[i][j] = mfma(f.a[i].x, f.b[j].x, acc_re[i][j]);
                    acc_im[i][j] = mfma(f.a[i].x, f.b[j].y, acc_im[i][j]);
                }
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc_re[i][j] = mfma_nega(f.a[i].y, f.b[j].y, acc_re[i][j]);
                    acc_im[i][j] = mfma(f.a[i].y, f.b[j].x, acc_im[i][j]);
                }
        }
    };

    // k-steps of the occupied bands come first; after the last of them the
    // accumulators hold the occupied-band contraction (asxtemp's factor)
    const int s_occ = a.nvpad / 4;
    if (s_occ == 0) occ_store(true);
    Frags<E, MB, NB> ring[P];
#pragma unroll
    for (int q = 0; q < P; ++q) load(ring[q], q);
    for (int s0 = 0; s0 < a.nks; s0 += P) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            mma(ring[q]);
            if (s0 + q + 1 == s_occ) occ_store(false);
            load(ring[q], s0 + q + P);
        }
    }

    // ---- epilogue ----------------------------------------------------------
    // Per (my_igp, ig) of the tile: B and Bocc from the registers, then per
    // frequency the pole factor D and the SX factor ssx, accumulated per lane
    // and reduced over the tile.
    const bool want_ssx = a.nv > 0;

    // rows of this lane that are real, kept columns
    unsigned rowok = 0;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rowbase + i * 16 + acc_row<E>(lane, r);
            const bool ok = row < a.G && gld<TABLE>(a.igp, min(row, a.G - 1)) >= 0;
            rowok |= (ok ? 1u : 0u) << (i * 4 + r);
        }

    const size_t tile = (size_t)tm + (size_t)tn * a.tiles_m;
    double2 sachs = make_double2(0.0, 0.0);
    for (int w0 = 0; w0 < a.nw; w0 += NWC) {
        double2 sach[NWC], sasx[NWC];
#pragma unroll
        for (int q = 0; q < NWC; ++q) { sach[q] = make_double2(0.0, 0.0); sasx[q] = make_double2(0.0, 0.0); }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int col = colbase + j * 16 + cl;
                E cwt[4], cie[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rowbase + i * 16 + acc_row<E>(lane, r);
                    const size_t e = (size_t)min(row, a.G - 1) * a.C + min(col, nmax);
                    cwt[r] = gld<TABLE>(a.wtilde, e);
                    cie[r] = gld<TABLE>(a.I_eps, e);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const bool ok = ((rowok >> (i * 4 + r)) & 1u) && col <= nmax;
                    const double2 B = acc_value(i, j, r);
                    const double2 Bo = widen(occ[((i * NB + j) * 4 + r) * 64 + lane]);
                    pole_element<NWC>(a.pp, w0, ok, widen(cwt[r]), widen(cie[r]), B, Bo, want_ssx, sach, sasx);
                    if constexpr (ACHS) {
                        if (w0 == 0 && ok) sachs = csub(sachs, cmul(B, widen(cie[r])));
                    }
                }
            }
        // wave reduction, then a fixed-order sum over the waves
#pragma unroll
        for (int q = 0; q < NWC; ++q) {
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
        if (tid < 4 * NWC) {
            const int q = tid >> 2, comp = tid & 3;
            const int iw = w0 + q;
            if (iw < a.nw) {
                double sum = red[0][tid];
#pragma unroll
                for (int w = 1; w < NWAVES; ++w) sum += red[w][tid];
                // comp 0,1: achtemp re,im; 2,3: asxtemp re,im
                gst<TABLE>(reinterpret_cast<double *>(a.partials + (tile * a.nw + iw) * 2), comp, sum);
            }
        }
        __syncthreads();
    }
    if constexpr (ACHS) {
        const double v0 = wave_sum(sachs.x), v1 = wave_sum(sachs.y);
        if (lane == 0) {
            red[wave][0] = v0;
            red[wave][1] = v1;
        }
        __syncthreads();
        if (tid < 2) {
            double sum = red[0][tid];
#pragma unroll
            for (int w = 1; w < NWAVES; ++w) sum += red[w][tid];
            gst<TABLE>(reinterpret_cast<double *>(a.achs_part + tile), tid, sum);
        }
        __syncthreads();
    }
}

// One problem, one tile per workgroup.
template <class E, int MB, int NB, int P, int WGM, int WGN, int MINW, int NWC, bool ACHS>
__global__ __launch_bounds__(64 * WGM * WGN, MINW) void gpp_fused_kernel(GppArgs<E> a)
{
    __shared__ TileLds<E, MB, NB, WGM * WGN, NWC> sh;
    // The second workgroup slot of each CU starts about half a tile late, so
    // that one slot's epilogue (VALU, memory) overlaps the other's MFMA loop;
    // later workgroups inherit the offset.  The first slot runs alone at the
    // full MFMA rate meanwhile.  The wait is bounded by the tick count.
    if ((int)blockIdx.x >= a.stagger_lo && (int)blockIdx.x < a.stagger_hi) {
        const uint64_t t0 = wall_clock64();
        while (wall_clock64() - t0 < (uint64_t)a.stagger_ticks) __builtin_amdgcn_s_sleep(32);
    }
    int tm, tn;
    tile_coords(blockIdx.x, a.tiles_m, a.tiles_n, a.group_m, tm, tn);
    gpp_tile<E, MB, NB, P, WGM, WGN, NWC, ACHS>(a, tm, tn, sh);
}

// Many problems in one launch: a fixed grid of workgroups takes tiles from a
// shared counter, in the order of the tile table (largest K first), until
// the table is exhausted.  tiles[t] = (problem, tm, tn).  (Measured on the
// 64-problem FP64 batch, 115.5 ms in this form: per-XCD segments of the
// table with a counter each, 119.5 ms; tiles in bands of 4 row tiles, 116.8
// ms; the second workgroup of each CU started half a tile late, 115.9 ms.)
struct BatchTile { int prob, tm, tn; };

template <class E, int MB, int NB, int P, int WGM, int WGN, int MINW, int NWC>
__global__ __launch_bounds__(64 * WGM * WGN, MINW) void gpp_batch_kernel(const GppArgs<E> *__restrict__ probs,
                                                                         const BatchTile *__restrict__ tiles,
                                                                         int ntiles, int *counter)
{
    __shared__ TileLds<E, MB, NB, WGM * WGN, NWC> sh;
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        // one value for the whole workgroup: read-only, uniform addresses let
        // the problem's arguments live in scalar registers, as a kernel
        // argument does in the single-problem kernel
        const int t = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (t >= ntiles) break;
        const BatchTile bt = tiles[t];
        const int prob = __builtin_amdgcn_readfirstlane(bt.prob);
        const int tm = __builtin_amdgcn_readfirstlane(bt.tm), tn = __builtin_amdgcn_readfirstlane(bt.tn);
        gpp_tile<E, MB, NB, P, WGM, WGN, NWC, false, true>(probs[prob], tm, tn, sh);
    }
}

__global__ void zero_outputs_kernel(double2 *a, double2 *b, int n, double2 *c)
{
    int i = threadIdx.x;
    if (i < n) { a[i] = make_double2(0.0, 0.0); b[i] = make_double2(0.0, 0.0); }
    if (c && i == 0) *c = make_double2(0.0, 0.0);
}

// out = sum_t part[t], one workgroup, fixed order
__global__ __launch_bounds__(256) void sum_partials_kernel(const double2 *part, int n, double2 *out)
{
    __shared__ double s_re[256], s_im[256];
    double re = 0.0, im = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) {
        re += part[t].x;
        im += part[t].y;
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
    if (threadIdx.x == 0) *out = make_double2(s_re[0], s_im[0]);
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

// Band-invariant pole factor of one frequency for the per-band output:
// X[my_igp][ig] = mask * delw * I_eps at wx (IEEE quotient, the reference's
// evaluation order), zero where the mask is off.  FP64 from either storage.
template <class E>
__global__ void pole_factor_kernel(const E *I_eps, const E *wtilde, int64_t n, double wx, double limittwo,
                                   double limitone, int mask_on, double2 *X)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double2 wt = widen(wtilde[i]);
        const double2 ie = widen(I_eps[i]);
        const double wdr = wx - wt.x;
        const double wdiffr = pole_wdiffr(wdr, -wt.y);
        double2 delw;
        double delwr;
        pole_delw(wt, wdr, 1.0 / wdiffr, delw, delwr);
        const bool m = (wdiffr > limittwo) && (delwr < limitone);
        X[i] = (m || !mask_on) ? cmul(delw, ie) : make_double2(0.0, 0.0);
    }
}

__global__ void widen_kernel(const float2 *src, int64_t n, double2 *dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = widen(src[i]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#ifndef FK_GPP_GROUP_M
#define FK_GPP_GROUP_M 4
#endif
// row tiles per band of the fused kernel's tile order (tile_coords).  At
// 512/2/32768/20 (26 x 512 tiles of 64 x 64) FETCH_SIZE per launch, x2 to
// bytes: whole columns 7.91 GB, bands of 4 or 8 row tiles 4.44 GB, of 13
// 5.3 GB (algorithmic 2.00 GB); the kernel time is the same for all.
constexpr int kGroupM = FK_GPP_GROUP_M;

struct TileCfg { int bm, bn, p, minw; };
// config ids (fk_options.config = id + 1): wave tile, ring depth, waves, WG/CU
//   0: 32x64 waves, ring 4, 1x4 waves (32x256 tile), 1 workgroup/CU
//   1: 32x32 waves, ring 3, 2x2 waves (64x64 tile),  2 workgroups/CU
//   2: 64x32 waves, ring 3, 2x2 waves (128x64 tile), 2 workgroups/CU
//   3: 32x64 waves, ring 3, 2x2 waves (64x128 tile), 2 workgroups/CU
constexpr TileCfg kCfgs[] = {{32, 256, 4, 1}, {64, 64, 3, 2}, {128, 64, 3, 2}, {64, 128, 3, 2}};
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <class E, int NWC, bool ACHS>
void launch_fused_nw(int cfg, int tiles, hipStream_t s, const GppArgs<E> &a)
{
    switch (cfg) {
    case 0: hipLaunchKernelGGL((gpp_fused_kernel<E, 2, 4, 4, 1, 4, 1, NWC, ACHS>), dim3(tiles), dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((gpp_fused_kernel<E, 2, 2, 3, 2, 2, 2, NWC, ACHS>), dim3(tiles), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((gpp_fused_kernel<E, 4, 2, 3, 2, 2, 2, NWC, ACHS>), dim3(tiles), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((gpp_fused_kernel<E, 2, 4, 3, 2, 2, 2, NWC, ACHS>), dim3(tiles), dim3(256), 0, s, a); break;
    }
}

// Frequencies per epilogue pass that the kernels hold in registers without
// scratch (gfx950 build, checked with -Rpass-analysis=kernel-resource-usage):
// FP64 configurations 0, 2 and 3 spill 800 B per lane at 4; the FP32 ones 2
// and 3 spill 270-360 B at 1..3 and less or none at 4.
template <class E> struct PassWidth;
template <> struct PassWidth<double2> {
    static constexpr int lo[kNumCfgs] = {1, 1, 1, 1}, hi[kNumCfgs] = {3, 4, 3, 3};
};
template <> struct PassWidth<float2> {
    static constexpr int lo[kNumCfgs] = {1, 1, 4, 4}, hi[kNumCfgs] = {4, 4, 4, 4};
};

template <class E, bool ACHS>
void launch_fused_achs(int cfg, int tiles, hipStream_t s, const GppArgs<E> &a)
{
    switch (std::max(PassWidth<E>::lo[cfg], std::min(a.nw, PassWidth<E>::hi[cfg]))) {
    case 1: launch_fused_nw<E, 1, ACHS>(cfg, tiles, s, a); break;
    case 2: launch_fused_nw<E, 2, ACHS>(cfg, tiles, s, a); break;
    case 3: launch_fused_nw<E, 3, ACHS>(cfg, tiles, s, a); break;
    default: launch_fused_nw<E, 4, ACHS>(cfg, tiles, s, a); break;
    }
}

// Frequencies per epilogue pass: all of them up to the configuration's
// width, else passes of that width; the static-CH remainder only in the
// kernels of calls that ask for it.
template <class E>
void launch_fused(int cfg, int tiles, hipStream_t s, const GppArgs<E> &a)
{
    if (a.achs_part) launch_fused_achs<E, true>(cfg, tiles, s, a);
    else launch_fused_achs<E, false>(cfg, tiles, s, a);
}

int choose_cfg(const fk_gpp_problem *p, const fk_options *o)
{
    (void)p;
    if (o->config > 0) return o->config - 1;
    // 32x32 wave tiles, two staggered workgroups per CU: the fastest measured
    // at 512/2/32768/20 and 128/2/8192/20
    return 1;
}

struct Layout {
    int cfg, Kpad, nvpad, ldc, tiles_m, tiles_n;
    size_t off_c, off_igp, off_part, off_achs;
    size_t off_stop;   // the gppsum = 1 form: the occupied loop's stop per column
};

Layout plan_layout(const fk_gpp_problem *p, const fk_options *o, Scratch &ws, size_t elem)
{
    Layout L;
    const bool sym = p->semantics == FK_GPP_GPPSUM;
    L.cfg = sym ? 1 : choose_cfg(p, o);   // the gppsum = 1 kernel has 64 x 64 tiles
    const TileCfg &t = kCfgs[L.cfg];
    L.nvpad = div_up(p->nvband, 4) * 4;
    const int rows = std::max(L.nvpad + p->number_bands - p->nvband, 1);
    L.Kpad = div_up(rows, 4 * t.p) * 4 * t.p;
    L.ldc = div_up(p->ngpown, t.bm) * t.bm;
    L.tiles_m = L.ldc / t.bm;
    L.tiles_n = div_up(p->ncouls, t.bn);
    L.off_c = ws.plan((size_t)L.Kpad * L.ldc * elem);
    L.off_igp = ws.plan(sizeof(int) * (size_t)p->ngpown);
    L.off_part = ws.plan(sizeof(double2) * 2 * (size_t)p->nfreq * L.tiles_m * L.tiles_n);
    L.off_achs = ws.plan(sizeof(double2) * (size_t)L.tiles_m * L.tiles_n);   // used when achstemp is asked for
    L.off_stop = sym ? ws.plan(sizeof(int) * (size_t)p->ngpown) : 0;
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
    if (p->semantics != FK_GPP_CANONICAL && p->semantics != FK_GPP_GPPSUM) {
        set_error("gpp_kernel: unknown semantics %d", p->semantics);
        return FK_ERR_ARG;
    }
    if (p->semantics == FK_GPP_GPPSUM && (o->precision != FK_PREC_FP64 || o->mask != 1 || o->config != 0)) {
        set_error("gpp_kernel: the gppsum = 1 form runs in FP64 with the default mask and configuration only");
        return FK_ERR_UNSUPPORTED;
    }
    return FK_OK;
}

// c[Kpad][ldc] and the device copy of the column map (gpp_coef_kernel, in
// chunks of kColChunk columns whose map travels in the kernel arguments).
template <class E>
int launch_coef(const fk_gpp_problem *p, const Layout &L, const std::vector<int32_t> &cols, const E *aqsm,
                const double *vcoul, E *c, int *igp, hipStream_t s)
{
    for (int m_begin = 0; m_begin < L.ldc; m_begin += kColChunk) {
        CoefArgs<E> ca;
        ca.aqsm = aqsm; ca.vcoul = vcoul; ca.c = c; ca.igp_out = igp;
        ca.nb = p->number_bands; ca.nv = p->nvband; ca.nvpad = L.nvpad;
        ca.C = p->ncouls; ca.G = p->ngpown; ca.Kpad = L.Kpad; ca.ldc = L.ldc;
        ca.m_begin = m_begin;
        ca.m_count = std::min(kColChunk, L.ldc - m_begin);
        for (int mm = 0; mm < kColChunk; ++mm) {
            int m = m_begin + mm;
            ca.igp[mm] = (mm < ca.m_count && m < p->ngpown) ? cols[m] : -1;
        }
        int64_t n = (int64_t)L.Kpad * ca.m_count;
        int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(gpp_coef_kernel<E>, dim3(blocks), dim3(256), 0, s, ca);
        FK_HIP(hipGetLastError());
    }
    return FK_OK;
}

// The gppsum = 1 form (fk_gpp_sym.hip) on FP64 inputs in device memory.
int launch_gpp_sym(const fk_gpp_problem *p, const fk_options *o, const Layout &L, Scratch &ws,
                   const std::vector<int32_t> &cols, const double2 *aqsm, const double2 *aqsn, const double2 *ieps,
                   const double2 *wt, const double *vcoul, double2 *d_ach, double2 *d_asx, hipStream_t s)
{
    if (p->number_bands == 0) {
        hipLaunchKernelGGL(zero_outputs_kernel, dim3(1), dim3(64), 0, s, d_ach, d_asx, p->nfreq, nullptr);
        FK_HIP(hipGetLastError());
        return FK_OK;
    }
    double2 *c = static_cast<double2 *>(ws.at(L.off_c));
    int *igp = static_cast<int *>(ws.at(L.off_igp));
    double2 *part = static_cast<double2 *>(ws.at(L.off_part));
    FK_TRY(launch_coef<double2>(p, L, cols, aqsm, vcoul, c, igp, s));
    FK_TRY(gpp_sym_launch(p, c, L.ldc, L.Kpad, L.nvpad, igp, static_cast<int *>(ws.at(L.off_stop)), aqsm, aqsn,
                          ieps, wt, part, L.tiles_m, L.tiles_n, o->kernel_ms, s));
    return gpp_reduce_launch(part, L.tiles_m * L.tiles_n, p->nfreq, 2.0 * p->occ, d_ach, d_asx, s);
}

template <class E>
int launch_gpp(const fk_gpp_problem *p, const fk_options *o, const Layout &L, Scratch &ws,
               const std::vector<int32_t> &cols, const E *aqsm, const E *aqsn, const E *ieps, const E *wt,
               const double *vcoul, double2 *d_ach, double2 *d_asx, double2 *d_achs, hipStream_t s)
{
    const int C = p->ncouls, G = p->ngpown;
    if (p->number_bands == 0) {   // no band: every sum is empty
        hipLaunchKernelGGL(zero_outputs_kernel, dim3(1), dim3(64), 0, s, d_ach, d_asx, p->nfreq, d_achs);
        FK_HIP(hipGetLastError());
        return FK_OK;
    }
    E *c = static_cast<E *>(ws.at(L.off_c));
    int *igp = static_cast<int *>(ws.at(L.off_igp));
    double2 *part = static_cast<double2 *>(ws.at(L.off_part));
    FK_TRY(launch_coef<E>(p, L, cols, aqsm, vcoul, c, igp, s));

    GppArgs<E> a;
    a.c = c; a.aqsn = aqsn; a.I_eps = ieps; a.wtilde = wt; a.igp = igp; a.partials = part;
    a.achs_part = d_achs ? static_cast<double2 *>(ws.at(L.off_achs)) : nullptr;
    a.nb = p->number_bands; a.nv = p->nvband; a.C = C; a.G = G; a.nw = p->nfreq; a.ldc = L.ldc; a.Kpad = L.Kpad;
    a.nvpad = L.nvpad; a.nks = L.Kpad / 4;
    a.tiles_m = L.tiles_m; a.tiles_n = L.tiles_n;
    a.group_m = std::min(L.tiles_m, kGroupM);
    const int tiles = L.tiles_m * L.tiles_n;
    a.stagger_lo = a.stagger_hi = 0;
    a.stagger_ticks = 0;
    if (kCfgs[L.cfg].minw == 2) {
        int dev = 0, ncu = 0;
        FK_HIP(hipGetDevice(&dev));
        FK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (tiles > 2 * ncu) {
            // half a tile of MFMA work at half a CU's rate (two workgroups
            // share it): FP64 MFMA 128 flop/cycle/CU, FP32 256; 2.4 GHz clock,
            // 100 MHz wall clock
            const double cu_rate = sizeof(E) == sizeof(double2) ? 128.0 : 256.0;
            const double tile_cycles = 8.0 * kCfgs[L.cfg].bm * kCfgs[L.cfg].bn * L.Kpad / (cu_rate / 2);
            a.stagger_lo = ncu;
            a.stagger_hi = 2 * ncu;
            a.stagger_ticks = (int)std::min(0.5 * tile_cycles / 24.0, 1.0e6);
        }
    }
    a.pp = make_pole_params(p->wx, p->nfreq, p->limittwo, p->limitone, p->to1, o->mask != 0);
    KernelTimer timer(o->kernel_ms, s);
    FK_TRY(timer.start());
    launch_fused<E>(L.cfg, tiles, s, a);
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());
    hipLaunchKernelGGL(gpp_reduce_kernel, dim3(2 * p->nfreq), dim3(256), 0, s, part, tiles, p->nfreq,
                       2.0 * p->occ, d_ach, d_asx);
    FK_HIP(hipGetLastError());
    if (d_achs) FK_TRY(gpp_sum_launch(a.achs_part, tiles, d_achs, s));
    return timer.finish();
}

// Scratch of the per-band output: X [G][C], widened aqsm and aqsn
// (complex64 inputs) and the row contraction's own scratch
struct N1Layout {
    size_t off_x = 0, off_m = 0, off_n = 0, off_rows = 0, rows_bytes = 0;
};
N1Layout plan_n1(const fk_gpp_problem *p, bool wide, Scratch &ws)
{
    const int nb = std::max(p->number_bands, 1), C = p->ncouls, G = p->ngpown;
    N1Layout N;
    N.off_x = ws.plan(sizeof(double2) * (size_t)G * C);
    if (!wide) {
        N.off_m = ws.plan(sizeof(double2) * (size_t)nb * C);
        N.off_n = ws.plan(sizeof(double2) * (size_t)nb * C);
    }
    N.rows_bytes = ff_rows_weighted_bytes(nb, C, G);
    N.off_rows = ws.plan(N.rows_bytes);
    return N;
}
size_t n1_bytes(const fk_gpp_problem *p, bool wide)
{
    Scratch ws;
    plan_n1(p, wide, ws);
    return ws.total();
}

// acht_n1_loc[n1] = sum_my_igp 0.5 vcoul[igp] conj(aqsm[n1][igp]) sum_ig aqsn[n1][ig] X[my_igp][ig]
// with X the pole factor at frequency index 2 (BGW-Kernels/gppKer.cpp:424):
// one more complex GEMM over all bands, run as the FF row contraction.
// Scratch: `base` (n1_bytes of caller workspace) or, with base = NULL, a
// stream-ordered allocation for the call.
template <class E>
int launch_acht_n1(const fk_gpp_problem *p, const fk_options *o, const std::vector<int32_t> &cols, const E *aqsm,
                   const E *aqsn, const E *ieps, const E *wt, const double *vcoul, double2 *d_n1, void *base,
                   hipStream_t s)
{
    const int nb = p->number_bands, C = p->ncouls, G = p->ngpown;
    if (nb == 0) return FK_OK;
    const bool wide = sizeof(E) == sizeof(double2);
    Scratch ws;
    const N1Layout N = plan_n1(p, wide, ws);
    const size_t off_x = N.off_x, off_m = N.off_m, off_n = N.off_n;
    FK_TRY(ws.acquire(base, base ? ws.total() : 0, s));
    double2 *X = static_cast<double2 *>(ws.at(off_x));
    const int64_t gc = (int64_t)G * C;
    const int blocks = (int)std::min<int64_t>((gc + 255) / 256, 8192);
    hipLaunchKernelGGL(pole_factor_kernel<E>, dim3(blocks), dim3(256), 0, s, ieps, wt, gc, p->wx[2], p->limittwo,
                       p->limitone, o->mask ? 1 : 0, X);
    FK_HIP(hipGetLastError());
    const double2 *am, *an;
    if constexpr (sizeof(E) == sizeof(double2)) {
        am = aqsm;
        an = aqsn;
    } else {
        double2 *wm = static_cast<double2 *>(ws.at(off_m)), *wn = static_cast<double2 *>(ws.at(off_n));
        const int64_t nbc = (int64_t)nb * C;
        const int b2 = (int)std::min<int64_t>((nbc + 255) / 256, 8192);
        hipLaunchKernelGGL(widen_kernel, dim3(b2), dim3(256), 0, s, aqsm, nbc, wm);
        hipLaunchKernelGGL(widen_kernel, dim3(b2), dim3(256), 0, s, aqsn, nbc, wn);
        FK_HIP(hipGetLastError());
        am = wm;
        an = wn;
    }
    return ff_rows_weighted(nb, C, G, am, an, X, vcoul, cols, 0.5, d_n1, ws.at(N.off_rows), N.rows_bytes, s);
}

// The part of the caller's workspace after its first `used` bytes, when it
// holds `need` bytes more; else NULL.
void *workspace_tail(const fk_options *o, size_t used, size_t need)
{
    if (!o->workspace || need == 0 || o->workspace_bytes < used + need) return nullptr;
    return static_cast<char *>(o->workspace) + used;
}

// Host-memory path: upload, run, download (synchronous).
template <class E, class Launch>
int run_host(const fk_gpp_problem *p, const fk_options *o, const std::vector<int32_t> &cols, double2 *ach,
             double2 *asx, double2 *n1, double2 *achs, void *n1_base, hipStream_t s, Launch &&launch)
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
    const size_t nout = 2 * (size_t)p->nfreq + 1 + (n1 ? (size_t)p->number_bands : 0);
    alloc((void **)&dout, sizeof(double2) * nout);
    copy(dm, p->aqsmtemp, nbC * sizeof(E));
    copy(dn, p->aqsntemp, nbC * sizeof(E));
    copy(die, p->I_eps_array, GC * sizeof(E));
    copy(dwt, p->wtilde_array, GC * sizeof(E));
    copy(dv, p->vcoul, sizeof(double) * p->ncouls);
    double2 *d_achs = achs ? dout + 2 * p->nfreq : nullptr;
    double2 *d_n1 = dout + 2 * p->nfreq + 1;
    if (rc == FK_OK) rc = launch(dm, dn, die, dwt, dv, dout, dout + p->nfreq, d_achs);
    if (rc == FK_OK && n1) rc = launch_acht_n1<E>(p, o, cols, dm, dn, die, dwt, dv, d_n1, n1_base, s);
    if (rc == FK_OK) {
        hipError_t e = hipMemcpyAsync(ach, dout, sizeof(double2) * p->nfreq, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(asx, dout + p->nfreq, sizeof(double2) * p->nfreq, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && n1 && p->number_bands > 0)
            e = hipMemcpyAsync(n1, d_n1, sizeof(double2) * p->number_bands, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && achs)
            e = hipMemcpyAsync(achs, d_achs, sizeof(double2), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error("download: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
    } else {
        (void)hipStreamSynchronize(s);
    }
    for (void *ptr : {(void *)dm, (void *)dn, (void *)die, (void *)dwt, (void *)dv, (void *)dout})
        if (ptr) (void)hipFree(ptr);
    return rc;
}

// ---------------------------------------------------------------------------
// Batched problems: one coefficient launch, one persistent tile launch and
// one reduction launch for the whole batch (captured once into a graph).
// ---------------------------------------------------------------------------
template <class E>
struct CoefBatch {
    const E *aqsm;
    const double *vcoul;
    const int *igp;          // resolved column map in device memory
    E *c;
    int nb, nv, nvpad, C, G, Kpad, ldc;
    int64_t begin;           // first element of this problem in the flat index
};

template <class E>
__global__ void gpp_coef_batch_kernel(const CoefBatch<E> *cb, int nprob, int64_t total)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        int lo = 0, hi = nprob - 1;            // last problem with begin <= idx
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cb[mid].begin <= idx) lo = mid;
            else hi = mid - 1;
        }
        const CoefBatch<E> &a = cb[lo];
        const int64_t local = idx - a.begin;
        const int kr = (int)(local / a.ldc), m = (int)(local % a.ldc);
        const int g = m < a.G ? a.igp[m] : -1;
        double2 v = make_double2(0.0, 0.0);
        const int band = band_of(kr, a.nb, a.nv, a.nvpad);
        if (band >= 0 && g >= 0) {
            const double2 am = widen(a.aqsm[(size_t)band * a.C + g]);
            const double sc = 0.5 * a.vcoul[g];
            v = make_double2(am.x * sc, -am.y * sc);
        }
        E *dst = a.c + (size_t)kr * a.ldc + m;
        if constexpr (sizeof(E) == sizeof(double2)) *dst = v;
        else *dst = make_float2((float)v.x, (float)v.y);
    }
}

struct ReduceBatch {
    const double2 *partials; // [tiles][nw][2]
    int tiles, nw;
    double asx_scale;
};

// blockIdx.y = problem, blockIdx.x = 2 iw + which; the same fixed-order sum
// as gpp_reduce_kernel; out[p][0][iw] = achtemp, out[p][1][iw] = asxtemp
__global__ __launch_bounds__(256) void gpp_reduce_batch_kernel(const ReduceBatch *rb, double2 *out)
{
    __shared__ double s_re[256], s_im[256];
    const ReduceBatch &a = rb[blockIdx.y];
    const int iw = blockIdx.x >> 1, which = blockIdx.x & 1;
    if (iw >= a.nw) return;
    double re = 0.0, im = 0.0;
    for (int t = threadIdx.x; t < a.tiles; t += 256) {
        const double2 v = a.partials[((size_t)t * a.nw + iw) * 2 + which];
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
        const double sc = which ? a.asx_scale : 1.0;
        out[((size_t)blockIdx.y * 2 + which) * kMaxFreq + iw] = make_double2(s_re[0] * sc, s_im[0] * sc);
    }
}

__global__ void set_int_kernel(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

// ach[p][iw] = out[p][0][iw], asx[p][iw] = out[p][1][iw] for iw < nw[p]
__global__ void batch_scatter_kernel(const double2 *out, const int *nw, double2 *const *ach, double2 *const *asx)
{
    const int p = blockIdx.x, iw = threadIdx.x;
    if (iw < nw[p]) {
        ach[p][iw] = out[((size_t)p * 2) * kMaxFreq + iw];
        asx[p][iw] = out[((size_t)p * 2 + 1) * kMaxFreq + iw];
    }
}

// frequencies per epilogue pass in batches (3..11 frequencies: 4 takes the
// mixed batch from 93 to 87 ms; the FP64 batch takes the same time with 2, 3
// or 4), 3 for the FP64 128 x 64 tiles, which spill 800 B per lane at 4
template <class E> constexpr int batch_nwc(int cfg) { return sizeof(E) == sizeof(double2) && cfg == 2 ? 3 : 4; }

template <class E>
void launch_batch(int cfg, int grid, hipStream_t s, const GppArgs<E> *probs, const BatchTile *tiles, int ntiles,
                  int *counter)
{
    switch (cfg) {
    case 0:
        hipLaunchKernelGGL((gpp_batch_kernel<E, 2, 4, 4, 1, 4, 1, batch_nwc<E>(0)>), dim3(grid), dim3(256), 0, s,
                           probs, tiles, ntiles, counter);
        break;
    case 1:
        hipLaunchKernelGGL((gpp_batch_kernel<E, 2, 2, 3, 2, 2, 2, batch_nwc<E>(1)>), dim3(grid), dim3(256), 0, s,
                           probs, tiles, ntiles, counter);
        break;
    case 2:
        hipLaunchKernelGGL((gpp_batch_kernel<E, 4, 2, 3, 2, 2, 2, batch_nwc<E>(2)>), dim3(grid), dim3(256), 0, s,
                           probs, tiles, ntiles, counter);
        break;
    default:
        hipLaunchKernelGGL((gpp_batch_kernel<E, 2, 4, 3, 2, 2, 2, batch_nwc<E>(3)>), dim3(grid), dim3(256), 0, s,
                           probs, tiles, ntiles, counter);
        break;
    }
}

// Host side of a batch: the layout of every problem, the device workspace,
// the tile table for the chosen configuration and the captured graph.
struct BatchProb {
    int nb, nv, C, G, nw, Kpad, nvpad, ldc, max_tiles;
    size_t off_c, off_igp, off_part;
};

struct GppBatch {
    int n = 0, mixed = 0, cfg = -1, device = 0, ncu = 0, ntiles = 0, max_tiles_total = 0;
    std::vector<fk_gpp_problem> src;
    std::vector<BatchProb> prob;
    void *mem = nullptr;
    size_t bytes = 0, off_args = 0, off_tiles = 0, off_counter = 0, off_out = 0, off_coef = 0, off_red = 0,
           off_nw = 0, off_ptrs = 0;
    int64_t coef_total = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    fk_options opt;

    char *at(size_t off) const { return static_cast<char *>(mem) + off; }
    size_t elem() const { return mixed ? sizeof(float2) : sizeof(double2); }

    ~GppBatch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cs) {
            (void)hipStreamSynchronize(cs);
            (void)hipStreamDestroy(cs);
        }
        if (mem) (void)hipFree(mem);
    }

    // GppArgs and tile table of configuration `c`, uploaded on the batch's stream
    template <class E>
    int prepare(int c)
    {
        const TileCfg &t = kCfgs[c];
        std::vector<GppArgs<E>> args(n);
        std::vector<BatchTile> tiles;
        std::vector<std::pair<std::pair<int, int>, int>> order;   // ((-nks, -nw), problem)
        for (int i = 0; i < n; ++i) {
            const BatchProb &b = prob[i];
            const fk_gpp_problem &p = src[i];
            GppArgs<E> &a = args[i];
            a.c = reinterpret_cast<const E *>(at(b.off_c));
            a.aqsn = static_cast<const E *>(p.aqsntemp);
            a.I_eps = static_cast<const E *>(p.I_eps_array);
            a.wtilde = static_cast<const E *>(p.wtilde_array);
            a.igp = reinterpret_cast<const int *>(at(b.off_igp));
            a.partials = reinterpret_cast<double2 *>(at(b.off_part));
            a.nb = b.nb; a.nv = b.nv; a.C = b.C; a.G = b.G; a.nw = b.nw;
            a.ldc = b.ldc; a.nvpad = b.nvpad;
            // The workspace is padded for every configuration (bands to 48,
            // rows to 128); the tiles walk only this configuration's own
            // padding, as a single call of it does (bands to 4 x ring depth,
            // rows to the tile height).  The coefficient rows past the bands
            // are zero and rows past G are masked, so reads stay inside the
            // larger padding.  (With the common padding the batch ran 3.5%
            // slower than the single calls.)
            const int kstep = 4 * t.p;
            a.Kpad = div_up(std::max(b.nvpad + b.nb - b.nv, 1), kstep) * kstep;
            a.nks = a.Kpad / 4;
            a.tiles_m = div_up(b.G, t.bm);
            a.tiles_n = div_up(b.C, t.bn);
            a.group_m = a.tiles_m;   // unused: the batch walks its own tile table
            a.stagger_lo = a.stagger_hi = a.stagger_ticks = 0;
            a.pp = make_pole_params(p.wx, p.nfreq, p.limittwo, p.limitone, p.to1, opt.mask != 0);
            if (b.nb > 0) order.push_back({{-a.nks, -b.nw}, i});
        }
        std::stable_sort(order.begin(), order.end());
        for (const auto &o : order) {
            const GppArgs<E> &a = args[o.second];
            for (int tn = 0; tn < a.tiles_n; ++tn)
                for (int tm = 0; tm < a.tiles_m; ++tm) tiles.push_back({o.second, tm, tn});
        }
        ntiles = (int)tiles.size();
        std::vector<ReduceBatch> red(n);
        for (int i = 0; i < n; ++i) {
            red[i].partials = args[i].partials;
            red[i].tiles = prob[i].nb > 0 ? args[i].tiles_m * args[i].tiles_n : 0;
            red[i].nw = prob[i].nw;
            red[i].asx_scale = 2.0 * src[i].occ;
        }
        FK_HIP(hipMemcpyAsync(at(off_args), args.data(), sizeof(GppArgs<E>) * n, hipMemcpyHostToDevice, cs));
        if (ntiles)
            FK_HIP(hipMemcpyAsync(at(off_tiles), tiles.data(), sizeof(BatchTile) * ntiles, hipMemcpyHostToDevice, cs));
        FK_HIP(hipMemcpyAsync(at(off_red), red.data(), sizeof(ReduceBatch) * n, hipMemcpyHostToDevice, cs));
        FK_HIP(hipStreamSynchronize(cs));
        cfg = c;
        return FK_OK;
    }

    // the launches of one run
    template <class E>
    int enqueue(hipStream_t s)
    {
        hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<int *>(at(off_counter)), 0);
        const int cblocks = (int)std::min<int64_t>((coef_total + 255) / 256, 1 << 16);
        hipLaunchKernelGGL(gpp_coef_batch_kernel<E>, dim3(std::max(cblocks, 1)), dim3(256), 0, s,
                           reinterpret_cast<const CoefBatch<E> *>(at(off_coef)), n, coef_total);
        launch_batch<E>(cfg, ncu * kCfgs[cfg].minw, s, reinterpret_cast<const GppArgs<E> *>(at(off_args)),
                        reinterpret_cast<const BatchTile *>(at(off_tiles)), ntiles,
                        reinterpret_cast<int *>(at(off_counter)));
        hipLaunchKernelGGL(gpp_reduce_batch_kernel, dim3(2 * kMaxFreq, n), dim3(256), 0, s,
                           reinterpret_cast<const ReduceBatch *>(at(off_red)), reinterpret_cast<double2 *>(at(off_out)));
        FK_HIP(hipGetLastError());
        return FK_OK;
    }

    template <class E>
    int build(int config)
    {
        // layout: Kpad a multiple of 48 and ldc of 128 serve every tile configuration
        Scratch plan;
        std::vector<CoefBatch<E>> coef(n);
        off_args = plan.plan(sizeof(GppArgs<E>) * n);
        off_red = plan.plan(sizeof(ReduceBatch) * n);
        off_coef = plan.plan(sizeof(CoefBatch<E>) * n);
        off_nw = plan.plan(sizeof(int) * n);
        off_ptrs = plan.plan(sizeof(void *) * 2 * n);
        off_out = plan.plan(sizeof(double2) * 2 * kMaxFreq * n);
        off_counter = plan.plan(sizeof(int));
        max_tiles_total = 0;
        coef_total = 0;
        for (int i = 0; i < n; ++i) {
            BatchProb &b = prob[i];
            const fk_gpp_problem &p = src[i];
            b.nb = p.number_bands; b.nv = p.nvband; b.C = p.ncouls; b.G = p.ngpown; b.nw = p.nfreq;
            b.nvpad = div_up(b.nv, 4) * 4;
            b.Kpad = div_up(std::max(b.nvpad + b.nb - b.nv, 1), 48) * 48;
            b.ldc = div_up(b.G, 128) * 128;
            b.max_tiles = 0;
            for (int c = 0; c < kNumCfgs; ++c)
                b.max_tiles = std::max(b.max_tiles, (b.ldc / kCfgs[c].bm) * div_up(b.C, kCfgs[c].bn));
            max_tiles_total += b.nb > 0 ? b.max_tiles : 0;
            b.off_c = plan.plan(elem() * (size_t)b.Kpad * b.ldc);
            b.off_igp = plan.plan(sizeof(int) * (size_t)b.G);
            b.off_part = plan.plan(sizeof(double2) * 2 * (size_t)b.nw * std::max(b.max_tiles, 1));
            coef[i].begin = coef_total;
            coef_total += (int64_t)b.Kpad * b.ldc;
        }
        off_tiles = plan.plan(sizeof(BatchTile) * (size_t)std::max(max_tiles_total, 1));
        bytes = plan.total();
        if (hipMalloc(&mem, bytes) != hipSuccess) {
            mem = nullptr;
            set_error("gpp_batch_create: hipMalloc(%zu) failed", bytes);
            return FK_ERR_NOMEM;
        }
        std::vector<int> nw(n);
        for (int i = 0; i < n; ++i) {
            const BatchProb &b = prob[i];
            const fk_gpp_problem &p = src[i];
            std::vector<int32_t> cols;
            FK_TRY(resolve_columns(p.ncouls, p.ngpown, p.inv_igp_index, p.indinv, cols));
            FK_HIP(hipMemcpyAsync(at(b.off_igp), cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice, cs));
            CoefBatch<E> &c = coef[i];
            c.aqsm = static_cast<const E *>(p.aqsmtemp);
            c.vcoul = p.vcoul;
            c.igp = reinterpret_cast<const int *>(at(b.off_igp));
            c.c = reinterpret_cast<E *>(at(b.off_c));
            c.nb = b.nb; c.nv = b.nv; c.nvpad = b.nvpad; c.C = b.C; c.G = b.G; c.Kpad = b.Kpad; c.ldc = b.ldc;
            nw[i] = b.nw;
            FK_HIP(hipStreamSynchronize(cs));   // cols leaves scope
        }
        FK_HIP(hipMemcpyAsync(at(off_coef), coef.data(), sizeof(CoefBatch<E>) * n, hipMemcpyHostToDevice, cs));
        FK_HIP(hipMemcpyAsync(at(off_nw), nw.data(), sizeof(int) * n, hipMemcpyHostToDevice, cs));
        FK_HIP(hipStreamSynchronize(cs));

        if (config > 0) {
            FK_TRY(prepare<E>(config - 1));
        } else {
            // autotuner: time one run of each configuration (after a warm-up
            // run) on this batch and keep the fastest
            hipEvent_t e0, e1;
            FK_HIP(hipEventCreate(&e0));
            FK_HIP(hipEventCreate(&e1));
            int best = 1;
            float best_ms = 1e30f;
            int rc = FK_OK;
            for (int c = 0; c < kNumCfgs && rc == FK_OK; ++c) {
                rc = prepare<E>(c);
                if (rc == FK_OK) rc = enqueue<E>(cs);
                if (rc == FK_OK && hipEventRecord(e0, cs) != hipSuccess) rc = FK_ERR_HIP;
                if (rc == FK_OK) rc = enqueue<E>(cs);
                if (rc == FK_OK && hipEventRecord(e1, cs) != hipSuccess) rc = FK_ERR_HIP;
                float ms = 0.f;
                if (rc == FK_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = FK_ERR_HIP;
                if (rc == FK_OK && ms < best_ms) { best_ms = ms; best = c; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc != FK_OK) { set_error("gpp_batch_create: autotuning failed"); return rc; }
            FK_TRY(prepare<E>(best));
        }
        FK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue<E>(cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        graph = g;
        if (rc != FK_OK) return rc;
        if (e != hipSuccess) { set_error("gpp_batch_create: graph capture: %s", hipGetErrorString(e)); return FK_ERR_HIP; }
        FK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return FK_OK;
    }

    int run(fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp, hipStream_t s, float *ms)
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ms) {
            FK_HIP(hipEventCreate(&e0));
            FK_HIP(hipEventCreate(&e1));
            FK_HIP(hipEventRecord(e0, s));
        }
        FK_HIP(hipGraphLaunch(exec, s));
        if (ms) FK_HIP(hipEventRecord(e1, s));
        std::vector<void *> ptrs(2 * n);
        for (int i = 0; i < n; ++i) {
            ptrs[i] = achtemp[i];
            ptrs[n + i] = asxtemp[i];
        }
        FK_HIP(hipMemcpyAsync(at(off_ptrs), ptrs.data(), sizeof(void *) * 2 * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(batch_scatter_kernel, dim3(n), dim3(kMaxFreq), 0, s, reinterpret_cast<const double2 *>(at(off_out)),
                           reinterpret_cast<const int *>(at(off_nw)), reinterpret_cast<double2 *const *>(at(off_ptrs)),
                           reinterpret_cast<double2 *const *>(at(off_ptrs)) + n);
        FK_HIP(hipGetLastError());
        if (ms) {
            FK_HIP(hipEventSynchronize(e1));
            FK_HIP(hipEventElapsedTime(ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return FK_OK;
    }
};

// Mixed-precision batches on the int8-split contraction: the kernels of the
// single-problem mixed path (operand slices, int8 GEMM, pole epilogue) as
// one launch each for all problems (fk_gpp_oz.hip, gpp_oz_batch_launch),
// then the batched reduction; the launches of one run are captured into a
// graph on one stream, as for the FP64 batch.  Every problem keeps its own
// workspace (its operand slices and FP64 contraction B, [ldm][ldn]) in one
// allocation.  (An earlier form ran every problem's launches on parallel
// branches of a multi-stream capture; it ended in a host crash inside
// gpp_batch_create on its first run, with the cause not found, and was
// removed.)
struct OzBatch {
    int n = 0;
    std::vector<fk_gpp_problem> src;
    void *mem = nullptr;
    size_t bytes = 0, off_rec = 0, off_first = 0, off_red = 0, off_out = 0, off_nw = 0, off_ptrs = 0;
    int total_prep = 0, total_gemm = 0, total_pole = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    fk_options opt;

    char *at(size_t off) const { return static_cast<char *>(mem) + off; }

    ~OzBatch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cs) {
            (void)hipStreamSynchronize(cs);
            (void)hipStreamDestroy(cs);
        }
        if (mem) (void)hipFree(mem);
    }

    int enqueue(hipStream_t s)
    {
        const int *first = reinterpret_cast<const int *>(at(off_first));
        FK_TRY(gpp_oz_batch_launch(at(off_rec), first, first + (n + 1), first + 2 * (n + 1), n, total_prep, total_gemm,
                                   total_pole, s));
        hipLaunchKernelGGL(gpp_reduce_batch_kernel, dim3(2 * kMaxFreq, n), dim3(256), 0, s,
                           reinterpret_cast<const ReduceBatch *>(at(off_red)), reinterpret_cast<double2 *>(at(off_out)));
        FK_HIP(hipGetLastError());
        return FK_OK;
    }

    int create(int n_, const fk_gpp_problem *const *problems, const fk_options *o)
    {
        n = n_;
        opt = *o;
        opt.kernel_ms = nullptr;
        opt.workspace = nullptr;
        opt.workspace_bytes = 0;
        for (int i = 0; i < n; ++i) src.push_back(*problems[i]);
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
            cs = nullptr;
            set_error("gpp_batch_create: stream creation failed");
            return FK_ERR_HIP;
        }
        const size_t rec_bytes = gpp_oz_batch_record_bytes();
        Scratch plan;
        off_rec = plan.plan(rec_bytes * n);
        off_first = plan.plan(sizeof(int) * 3 * (size_t)(n + 1));
        off_red = plan.plan(sizeof(ReduceBatch) * n);
        off_out = plan.plan(sizeof(double2) * 2 * kMaxFreq * n);
        off_nw = plan.plan(sizeof(int) * n);
        off_ptrs = plan.plan(sizeof(void *) * 2 * n);
        std::vector<size_t> off_igp(n), off_ws(n), ws_bytes(n);
        for (int i = 0; i < n; ++i) {
            off_igp[i] = plan.plan(sizeof(int) * (size_t)src[i].ngpown);
            Scratch ws;
            OzLayout L;
            gpp_oz_plan(&src[i], ws, L);
            ws_bytes[i] = ws.total();       // also without bands: every record points into the allocation
            off_ws[i] = plan.plan(ws_bytes[i]);
        }
        bytes = plan.total();
        if (hipMalloc(&mem, bytes) != hipSuccess) {
            mem = nullptr;
            set_error("gpp_batch_create: hipMalloc(%zu) failed", bytes);
            return FK_ERR_NOMEM;
        }
        std::vector<char> recs(rec_bytes * n);
        std::vector<int> first(3 * (size_t)(n + 1), 0), nw(n);
        std::vector<ReduceBatch> red(n);
        for (int i = 0; i < n; ++i) {
            const fk_gpp_problem &p = src[i];
            std::vector<int32_t> cols;
            FK_TRY(resolve_columns(p.ncouls, p.ngpown, p.inv_igp_index, p.indinv, cols));
            FK_HIP(hipMemcpyAsync(at(off_igp[i]), cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice, cs));
            FK_HIP(hipStreamSynchronize(cs));   // cols leaves scope
            Scratch ws;
            OzLayout L;
            gpp_oz_plan(&p, ws, L);
            FK_TRY(ws.acquire(at(off_ws[i]), ws_bytes[i], cs));
            const OzBlocks nbk = gpp_oz_batch_record(&p, &opt, L, ws, reinterpret_cast<const int *>(at(off_igp[i])),
                                                     recs.data() + rec_bytes * i);
            first[i + 1] = first[i] + nbk.prep;
            first[(n + 1) + i + 1] = first[(n + 1) + i] + nbk.gemm;
            first[2 * (n + 1) + i + 1] = first[2 * (n + 1) + i] + nbk.pole;
            red[i].partials = static_cast<const double2 *>(ws.at(L.off_part));
            red[i].tiles = nbk.pole;
            red[i].nw = p.nfreq;
            red[i].asx_scale = 2.0 * p.occ;
            nw[i] = p.nfreq;
        }
        total_prep = first[n];
        total_gemm = first[(n + 1) + n];
        total_pole = first[2 * (n + 1) + n];
        FK_HIP(hipMemcpyAsync(at(off_rec), recs.data(), recs.size(), hipMemcpyHostToDevice, cs));
        FK_HIP(hipMemcpyAsync(at(off_first), first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, cs));
        FK_HIP(hipMemcpyAsync(at(off_red), red.data(), sizeof(ReduceBatch) * n, hipMemcpyHostToDevice, cs));
        FK_HIP(hipMemcpyAsync(at(off_nw), nw.data(), sizeof(int) * n, hipMemcpyHostToDevice, cs));
        FK_HIP(hipStreamSynchronize(cs));
        FK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        graph = g;
        if (rc != FK_OK) return rc;
        if (e != hipSuccess) { set_error("gpp_batch_create: graph capture: %s", hipGetErrorString(e)); return FK_ERR_HIP; }
        FK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return FK_OK;
    }

    int run(fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp, hipStream_t s, float *ms)
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ms) {
            FK_HIP(hipEventCreate(&e0));
            FK_HIP(hipEventCreate(&e1));
            FK_HIP(hipEventRecord(e0, s));
        }
        FK_HIP(hipGraphLaunch(exec, s));
        if (ms) FK_HIP(hipEventRecord(e1, s));
        std::vector<void *> ptrs(2 * n);
        for (int i = 0; i < n; ++i) {
            ptrs[i] = achtemp[i];
            ptrs[n + i] = asxtemp[i];
        }
        FK_HIP(hipMemcpyAsync(at(off_ptrs), ptrs.data(), sizeof(void *) * 2 * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(batch_scatter_kernel, dim3(n), dim3(kMaxFreq), 0, s, reinterpret_cast<const double2 *>(at(off_out)),
                           reinterpret_cast<const int *>(at(off_nw)), reinterpret_cast<double2 *const *>(at(off_ptrs)),
                           reinterpret_cast<double2 *const *>(at(off_ptrs)) + n);
        FK_HIP(hipGetLastError());
        if (ms) {
            FK_HIP(hipEventSynchronize(e1));
            FK_HIP(hipEventElapsedTime(ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return FK_OK;
    }
};

}  // namespace

int gpp_sum_launch(const double2 *partials, int n, double2 *out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partials, n, out);
    FK_HIP(hipGetLastError());
    return FK_OK;
}

int gpp_reduce_launch(const double2 *partials, int tiles, int nw, double asx_scale, double2 *achtemp,
                      double2 *asxtemp, hipStream_t s)
{
    hipLaunchKernelGGL(gpp_reduce_kernel, dim3(2 * nw), dim3(256), 0, s, partials, tiles, nw, asx_scale, achtemp,
                       asxtemp);
    FK_HIP(hipGetLastError());
    return FK_OK;
}

}  // namespace fk

using namespace fk;

extern "C" {

size_t gpp_kernel_workspace_size(const fk_gpp_problem *p, const fk_options *o)
{
    fk_options d = fk_default_options();
    if (!o) o = &d;
    if (check_problem(p, o) != FK_OK) return 0;
    Scratch ws;
    if (o->precision == FK_PREC_MIXED && o->config == 0) {
        OzLayout Z;
        gpp_oz_plan(p, ws, Z);
        ws.plan(sizeof(int) * (size_t)p->ngpown);
    } else {
        plan_layout(p, o, ws, o->precision == FK_PREC_MIXED ? sizeof(float2) : sizeof(double2));
    }
    return ws.total();
}

size_t gpp_kernel_ex_workspace_size(const fk_gpp_problem *p, const fk_options *o, const fk_gpp_outputs *out)
{
    fk_options d = fk_default_options();
    if (!o) o = &d;
    const size_t base = gpp_kernel_workspace_size(p, o);
    if (base == 0 || !out || !out->acht_n1_loc || p->semantics != FK_GPP_CANONICAL) return base;
    return base + n1_bytes(p, o->precision != FK_PREC_MIXED);
}

int gpp_kernel(const fk_gpp_problem *p, const fk_options *o, fk_cdouble *achtemp, fk_cdouble *asxtemp,
               fk_cdouble *acht_n1_loc)
{
    fk_gpp_outputs out;
    out.achtemp = achtemp;
    out.asxtemp = asxtemp;
    out.acht_n1_loc = acht_n1_loc;
    out.achstemp = nullptr;
    return gpp_kernel_ex(p, o, &out);
}

int gpp_kernel_ex(const fk_gpp_problem *p, const fk_options *o, const fk_gpp_outputs *out)
{
    clear_error();
    FK_TRY(check_problem(p, o));
    if (!out || !out->achtemp || !out->asxtemp) { set_error("gpp_kernel: NULL output"); return FK_ERR_ARG; }
    fk_cdouble *achtemp = out->achtemp, *asxtemp = out->asxtemp, *acht_n1_loc = out->acht_n1_loc;
    if (acht_n1_loc && p->nfreq < 3) {   // the per-band sum is taken at frequency index 2 (:424)
        set_error("gpp_kernel: acht_n1_loc needs nfreq >= 3 (got %d)", p->nfreq);
        return FK_ERR_ARG;
    }
    const bool sym = p->semantics == FK_GPP_GPPSUM;
    if (sym && (acht_n1_loc || out->achstemp)) {
        set_error("gpp_kernel: the gppsum = 1 form gives achtemp and asxtemp only (no acht_n1_loc, no achstemp)");
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
    // FK_PREC_MIXED with config 0: the int8-split contraction; configs 1..4
    // select the FP32 MFMA contraction with that tile configuration
    const bool oz = mixed && o->config == 0;
    if (oz) FK_TRY(gpp_oz_check(p));
    double2 *ach = reinterpret_cast<double2 *>(achtemp), *asx = reinterpret_cast<double2 *>(asxtemp);
    double2 *n1 = reinterpret_cast<double2 *>(acht_n1_loc);
    double2 *achs = reinterpret_cast<double2 *>(out->achstemp);
    int rc;
    if (oz) {
        Scratch ws;
        OzLayout Z;
        gpp_oz_plan(p, ws, Z);
        const size_t off_igp = ws.plan(sizeof(int) * (size_t)p->ngpown);
        void *n1_base = workspace_tail(o, ws.total(), n1 ? n1_bytes(p, false) : 0);
        rc = ws.acquire(o->workspace, o->workspace_bytes, s);
        int *d_igp = static_cast<int *>(ws.at(off_igp));
        if (rc == FK_OK) {
            hipError_t e = hipMemcpyAsync(d_igp, igp.data(), sizeof(int) * igp.size(), hipMemcpyHostToDevice, s);
            if (e != hipSuccess) { set_error("column map upload: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
        }
        auto launch = [&](const float2 *m, const float2 *n, const float2 *ie, const float2 *w, const double *v,
                          double2 *da, double2 *dx, double2 *dc) -> int {
            if (p->number_bands == 0) {
                hipLaunchKernelGGL(zero_outputs_kernel, dim3(1), dim3(64), 0, s, da, dx, p->nfreq, dc);
                FK_HIP(hipGetLastError());
                return FK_OK;
            }
            return gpp_oz_launch(p, o, Z, ws, d_igp, m, n, ie, w, v, da, dx, dc, s);
        };
        if (rc == FK_OK) {
            if (o->memory == FK_MEM_DEVICE) {
                const float2 *m = (const float2 *)p->aqsmtemp, *n = (const float2 *)p->aqsntemp;
                const float2 *ie = (const float2 *)p->I_eps_array, *w = (const float2 *)p->wtilde_array;
                rc = launch(m, n, ie, w, p->vcoul, ach, asx, achs);
                if (rc == FK_OK && n1)
                    rc = launch_acht_n1<float2>(p, o, igp, m, n, ie, w, p->vcoul, n1, n1_base, s);
            } else {
                rc = run_host<float2>(p, o, igp, ach, asx, n1, achs, n1_base, s, launch);
            }
        }
    } else {
        Scratch ws;
        Layout L = plan_layout(p, o, ws, mixed ? sizeof(float2) : sizeof(double2));
        void *n1_base = workspace_tail(o, ws.total(), n1 ? n1_bytes(p, !mixed) : 0);
        rc = ws.acquire(o->workspace, o->workspace_bytes, s);
        if (rc == FK_OK) {
            if (o->memory == FK_MEM_DEVICE) {
                if (mixed) {
                    const float2 *m = (const float2 *)p->aqsmtemp, *n = (const float2 *)p->aqsntemp;
                    const float2 *ie = (const float2 *)p->I_eps_array, *w = (const float2 *)p->wtilde_array;
                    rc = launch_gpp<float2>(p, o, L, ws, igp, m, n, ie, w, p->vcoul, ach, asx, achs, s);
                    if (rc == FK_OK && n1)
                        rc = launch_acht_n1<float2>(p, o, igp, m, n, ie, w, p->vcoul, n1, n1_base, s);
                } else {
                    const double2 *m = (const double2 *)p->aqsmtemp, *n = (const double2 *)p->aqsntemp;
                    const double2 *ie = (const double2 *)p->I_eps_array, *w = (const double2 *)p->wtilde_array;
                    if (sym) {
                        rc = launch_gpp_sym(p, o, L, ws, igp, m, n, ie, w, p->vcoul, ach, asx, s);
                    } else {
                        rc = launch_gpp<double2>(p, o, L, ws, igp, m, n, ie, w, p->vcoul, ach, asx, achs, s);
                        if (rc == FK_OK && n1)
                            rc = launch_acht_n1<double2>(p, o, igp, m, n, ie, w, p->vcoul, n1, n1_base, s);
                    }
                }
            } else if (mixed) {
                rc = run_host<float2>(p, o, igp, ach, asx, n1, achs, n1_base, s,
                                      [&](const float2 *m, const float2 *n, const float2 *ie, const float2 *w,
                                          const double *v, double2 *da, double2 *dx, double2 *dc) {
                                          return launch_gpp<float2>(p, o, L, ws, igp, m, n, ie, w, v, da, dx, dc,
                                                                    s);
                                      });
            } else {
                rc = run_host<double2>(p, o, igp, ach, asx, n1, achs, n1_base, s,
                                       [&](const double2 *m, const double2 *n, const double2 *ie, const double2 *w,
                                           const double *v, double2 *da, double2 *dx, double2 *dc) {
                                           if (sym) return launch_gpp_sym(p, o, L, ws, igp, m, n, ie, w, v, da, dx, s);
                                           return launch_gpp<double2>(p, o, L, ws, igp, m, n, ie, w, v, da, dx, dc,
                                                                      s);
                                       });
            }
        }
    }
    if (prev_dev >= 0) (void)hipSetDevice(prev_dev);
    return rc;
}

}  // extern "C"

struct fk_gpp_batch {
    fk::GppBatch b;
    std::unique_ptr<fk::OzBatch> oz;   // mixed precision, config 0: the int8-split contraction
};

namespace {

int create_batch(int32_t n, const fk_gpp_problem *const *problems, const fk_options *options, fk_gpp_batch **batch)
{
    if (!batch) { set_error("gpp_batch_create: NULL result pointer"); return FK_ERR_ARG; }
    *batch = nullptr;
    if (n < 1 || !problems || !options) { set_error("gpp_batch_create: need n >= 1, problems and options"); return FK_ERR_ARG; }
    if (options->memory != FK_MEM_DEVICE) {
        set_error("gpp_batch_create: batches take device memory (FK_MEM_DEVICE); call gpp_kernel per problem for host memory");
        return FK_ERR_UNSUPPORTED;
    }
    if (options->precision != FK_PREC_FP64 && options->precision != FK_PREC_MIXED) {
        set_error("gpp_batch_create: bad precision");
        return FK_ERR_ARG;
    }
    if (options->config < 0 || options->config > kNumCfgs) {
        set_error("gpp_batch_create: config %d out of range", options->config);
        return FK_ERR_ARG;
    }
    for (int i = 0; i < n; ++i) {
        if (!problems[i]) { set_error("gpp_batch_create: problem %d is NULL", i); return FK_ERR_ARG; }
        if (problems[i]->semantics != FK_GPP_CANONICAL) {
            set_error("gpp_batch_create: problem %d: batches run the canonical semantics (call gpp_kernel for "
                      "the gppsum = 1 form)", i);
            return FK_ERR_UNSUPPORTED;
        }
        fk_options oi = *options;
        oi.config = 0;
        const int rc = check_problem(problems[i], &oi);
        if (rc != FK_OK) return rc;
        std::vector<int32_t> cols;
        FK_TRY(resolve_columns(problems[i]->ncouls, problems[i]->ngpown, problems[i]->inv_igp_index,
                               problems[i]->indinv, cols));
    }
    if (options->precision == FK_PREC_MIXED && options->config == 0) {
        for (int i = 0; i < n; ++i) FK_TRY(gpp_oz_check(problems[i]));
        fk_gpp_batch *h = new fk_gpp_batch;
        h->oz.reset(new OzBatch);
        const int rc = h->oz->create(n, problems, options);
        if (rc != FK_OK) {
            delete h;
            return rc;
        }
        *batch = h;
        return FK_OK;
    }
    fk_gpp_batch *h = new fk_gpp_batch;
    GppBatch &b = h->b;
    b.n = n;
    b.mixed = options->precision == FK_PREC_MIXED;
    b.opt = *options;
    for (int i = 0; i < n; ++i) b.src.push_back(*problems[i]);
    b.prob.resize(n);
    int rc = FK_OK;
    if (hipGetDevice(&b.device) != hipSuccess ||
        hipDeviceGetAttribute(&b.ncu, hipDeviceAttributeMultiprocessorCount, b.device) != hipSuccess ||
        hipStreamCreateWithFlags(&b.cs, hipStreamNonBlocking) != hipSuccess) {
        set_error("gpp_batch_create: device query or stream creation failed");
        rc = FK_ERR_HIP;
    }
    if (rc == FK_OK) rc = b.mixed ? b.build<float2>(options->config) : b.build<double2>(options->config);
    if (rc != FK_OK) {
        delete h;
        return rc;
    }
    *batch = h;
    return FK_OK;
}

}  // namespace

extern "C" {

int gpp_batch_create(int32_t n, const fk_gpp_problem *const *problems, const fk_options *options,
                     fk_gpp_batch **batch)
{
    clear_error();
    return create_batch(n, problems, options, batch);
}

int gpp_batch_run(fk_gpp_batch *batch, fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp, void *stream,
                  float *elapsed_ms)
{
    clear_error();
    if (!batch || !achtemp || !asxtemp) { set_error("gpp_batch_run: NULL argument"); return FK_ERR_ARG; }
    const int n = batch->oz ? batch->oz->n : batch->b.n;
    for (int i = 0; i < n; ++i)
        if (!achtemp[i] || !asxtemp[i]) { set_error("gpp_batch_run: NULL output of problem %d", i); return FK_ERR_ARG; }
    if (batch->oz) return batch->oz->run(achtemp, asxtemp, static_cast<hipStream_t>(stream), elapsed_ms);
    return batch->b.run(achtemp, asxtemp, static_cast<hipStream_t>(stream), elapsed_ms);
}

int gpp_batch_config(const fk_gpp_batch *batch)
{
    if (!batch) return 0;
    return batch->oz ? 0 : batch->b.cfg + 1;
}

size_t gpp_batch_workspace_bytes(const fk_gpp_batch *batch)
{
    if (!batch) return 0;
    return batch->oz ? batch->oz->bytes : batch->b.bytes;
}

void gpp_batch_destroy(fk_gpp_batch *batch)
{
    delete batch;
}

int gpp_kernel_batched(int32_t n, const fk_gpp_problem *const *problems, const fk_options *options,
                       fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp)
{
    clear_error();
    fk_options o = options ? *options : fk_default_options();
    // one call: no autotuning; the single-problem default tiles, or (mixed)
    // the int8-split contraction
    if (o.config == 0 && o.precision != FK_PREC_MIXED) o.config = 2;
    fk_gpp_batch *b = nullptr;
    FK_TRY(create_batch(n, problems, &o, &b));
    hipStream_t s = static_cast<hipStream_t>(o.stream);
    int rc = gpp_batch_run(b, achtemp, asxtemp, s, o.kernel_ms);
    if (rc == FK_OK && hipStreamSynchronize(s) != hipSuccess) {
        set_error("gpp_kernel_batched: stream synchronisation failed");
        rc = FK_ERR_HIP;
    }
    gpp_batch_destroy(b);
    return rc;
}

}  // extern "C"
