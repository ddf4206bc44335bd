// fk_gpp_oz.hip -- mixed-precision GPP band sum on MI355X (gfx950): the
// band contraction B = sum_n1 c[n1] aqsn[n1] runs on int8 MFMA with exact
// int32 accumulation (an Ozaki-style split, Ozaki, Ogita, Oishi & Rump,
// "Error-free transformations of matrix multiplication", Numer. Algorithms
// 59 (2012); on integer matrix units: Ootomo, Ozaki & Yokota, IJHPCA 2024);
// everything else is FP64.
//
// Inputs are complex64 (fk_cfloat).  Each row of the coefficient matrix
// c[n1][my_igp] = 0.5 vcoul[igp] conj(aqsm[n1][igp]) (FP64) and each column
// of aqsn is scaled so that its largest |re|, |im| (M) maps to 127, and
// split into four int8 slices:
//     x = M / 127 * (s0 + s1/254 + s2/254^2 + s3/254^3),  |s_k| <= 127,
// 31 bits of the row's or column's range.  A complex product is a real GEMM
// over an interleaved K' = 2 nb axis:
//     re B = sum (cr, -ci) . (ar, ai),   im B = sum (ci, cr) . (ar, ai),
// and the slice pairs (i, j) with i + j = L share the weight 254^-L, so five
// int32 accumulators per component (L = 0..4; the pairs with i + j > 4 are
// below 254^-5 of the leading term and dropped) hold the contraction
// exactly.  The error against the FP64 evaluation on the same complex64
// values is the slice truncation's, a few 1e-10 relative.
//
// Kernels (one call = four launches on the caller's stream):
//   oz_prep_kernel       row / column maxima of the coefficients and aqsn,
//                        their int8 slices, FP64 occupied rows
//   gpp_oz_gemm_kernel   int8 MFMA tile, operands staged through LDS by
//                        direct global->LDS loads in a five-deep ring;
//                        B rebuilt in FP64 and stored
//   gpp_oz_pole_kernel   per (my_igp, ig): occupied-band sum Bo in FP64, the
//                        pole and SX factors in FP64, per-block partials
//   gpp_reduce_kernel    fixed-order sum of the partials (shared with FP64)
//
// Slices are stored in MFMA-lane order: a 1 KB block per (slice, 32 rows,
// 32 K') in which lane L's 16 bytes (row L%32, K' 16 (L/32) .. +15) sit at
// byte 16 L, so one direct-to-LDS load moves a whole operand fragment of
// the 32x32x32 int8 MFMA and the MFMA reads it back with one conflict-free
// ds_read_b128 per lane.  (The 16x16x64 int8 MFMA runs at half the int8
// rate on this part; the 32x32x32 form reaches it.)
#include "fk_complex.h"
#include "fk_gpp_common.h"
#include "fk_internal.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace fk {
namespace {

using namespace gpp;
typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int kSlices = 4;
constexpr int kLevels = 5;      // slice pairs with i + j < kLevels
constexpr int kKStep = 64;      // K' per pipeline step = 32 bands = two MFMA k-halves
constexpr int kKHalf = 32;      // K' per 32x32x32 int8 MFMA
constexpr int kBandsPerStep = kKStep / 2;
constexpr int kPlanesA = kSlices, kPlanesB = kSlices;
constexpr int kFrag = 1024;     // bytes per lane-ordered fragment
constexpr int kMaxBandsOz = 16000;   // keeps the level-3 int32 sums (4 pairs) below 2^31

// Slices of x, |x| <= M (scale = 127 / M), in units of M/127 (first) and
// /254 per level.  The scaling rounds once at 2^-53 relative; every later
// step is exact.
__device__ __forceinline__ void split4(double x, double scale, int8_t (&s)[kSlices])
{
    double t = x * scale;
    double q = rint(t);
    s[0] = (int8_t)q;
#pragma unroll
    for (int k = 1; k < kSlices; ++k) {
        t = (t - q) * 254.0;
        q = rint(t);
        s[k] = (int8_t)q;
    }
}


__device__ __forceinline__ double2 coef(const float2 *aqsm, const double *vcoul, int C, int n1, int g)
{
    const float2 am = aqsm[(size_t)n1 * C + g];
    const double s = 0.5 * vcoul[g];
    return make_double2((double)am.x * s, -(double)am.y * s);
}

// ---------------------------------------------------------------------------
// Slices of both operands in one launch.  A workgroup takes 4 coefficient
// rows (the first ldm/4 workgroups) or 16 aqsn columns (the rest): it finds
// their max(|re|, |im|) over all bands, then splits them band block by band
// block (8 bands = the 16 bytes of one fragment lane); the second pass
// finds the workgroup's slab in L2.  Thread t of an R-wide slab: row or
// column R x + t%R, band blocks t/R, t/R + 256/R, ...  Coefficient slices
// hold (cr, -ci) (the GEMM derives (ci, cr)); aqsn slices hold (ar, ai).
// ---------------------------------------------------------------------------
struct PrepArgs {
    const float2 *aqsm, *aqsn;
    const double *vcoul;
    const int *igp;
    int8_t *as, *bs;         // [4][ldm/32][Kp/32][1024], [4][ldn/32][Kp/32][1024]
    double *ra, *rb;         // M / 127 of each row, column
    double2 *c_occ;          // [nv][ldm] coefficients of the occupied bands
    int nb, nv, C, G, ldm, ldn, nkh, blocks_a;
};

// R rows (coefficients, IS_A) or columns (aqsn) per workgroup, 256 / R band
// groups: the coefficients are gathered at scattered igp columns, so their
// slabs are kept thin to spread the gathers over more workgroups.
template <int R, bool IS_A>
__device__ __forceinline__ void prep_slab(const PrepArgs &a, int slab)
{
    constexpr int NG = 256 / R;
    __shared__ double sm[NG][R];
    const int tid = threadIdx.x, c = tid % R, grp = tid / R;
    const int idx = slab * R + c;
    const int g = IS_A ? (idx < a.G ? a.igp[idx] : -1) : 0;
    const bool in = IS_A ? g >= 0 : idx < a.C;
    // element n1 of this row / column as (first, second) of its stored pair
    auto value = [&](int n1) -> double2 {
        if (!in || n1 >= a.nb) return make_double2(0.0, 0.0);
        if constexpr (IS_A) {
            const double2 cf = coef(a.aqsm, a.vcoul, a.C, n1, g);
            return make_double2(cf.x, -cf.y);
        } else {
            const float2 v = a.aqsn[(size_t)n1 * a.C + idx];
            return make_double2((double)v.x, (double)v.y);
        }
    };
    const int nblk = (a.nb + 7) / 8;
    double mx = 0.0;
    for (int j = grp; j < nblk; j += NG)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const double2 v = value(8 * j + b);
            mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
        }
    sm[grp][c] = mx;
    __syncthreads();
    mx = 0.0;
    for (int k = 0; k < NG; ++k) mx = fmax(mx, sm[k][c]);
    const double scale = mx > 0.0 ? 127.0 / mx : 0.0;
    if (grp == 0) (IS_A ? a.ra : a.rb)[idx] = mx / 127.0;

    int8_t *base = IS_A ? a.as : a.bs;
    const size_t blk = idx >> 5, nblks = (size_t)(IS_A ? a.ldm : a.ldn) / 32;
    for (int j = grp; j < a.nkh * 2; j += NG) {   // band blocks of the padded K'
        int8_t p[kSlices][16];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int n1 = 8 * j + b;
            const double2 v = value(n1);
            if (IS_A && n1 < a.nv) a.c_occ[(size_t)n1 * a.ldm + idx] = make_double2(v.x, -v.y);
            int8_t s0[kSlices], s1[kSlices];
            split4(v.x, scale, s0);
            split4(v.y, scale, s1);
#pragma unroll
            for (int s = 0; s < kSlices; ++s) {
                p[s][2 * b] = s0[s];
                p[s][2 * b + 1] = s1[s];
            }
        }
        const int lane = (idx & 31) + 32 * (j & 1), kh = j >> 1;
#pragma unroll
        for (int pl = 0; pl < kSlices; ++pl) {
            int8_t *dst = base + ((pl * nblks + blk) * a.nkh + kh) * kFrag + 16 * lane;
            *reinterpret_cast<i4 *>(dst) = *reinterpret_cast<const i4 *>(p[pl]);
        }
    }
}

constexpr int kPrepRowsA = 4, kPrepColsB = 16;

__device__ __forceinline__ void oz_prep_block(const PrepArgs &a, int b)
{
    if (b < a.blocks_a) prep_slab<kPrepRowsA, true>(a, b);
    else prep_slab<kPrepColsB, false>(a, b - a.blocks_a);
}

__global__ __launch_bounds__(256) void oz_prep_kernel(PrepArgs a)
{
    oz_prep_block(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// int8 MFMA GEMM.  64x64 workgroup tile, 2x2 waves of one 32x32 block
// each.  32x32x32 int8 MFMA: lane l supplies 16 bytes of row l%32 of A and
// of column l%32 of B at the same K' positions; accumulator register j of
// lane l holds row 8 (j/4) + 4 (l/32) + j%4, column l%32.
//
// LDS ring of kStages k-steps; a stage holds the workgroup's 16 A fragments
// ([slice][row block][k-half]) and 16 B fragments ([slice][column
// block][k-half]), 32 KB.  Each wave moves 8 fragments per step with direct
// global->LDS loads, issued four steps before the step that uses them.
// A step is two halves (32 K' each).  A half runs its 26 MFMAs from one
// register set and, between them at fixed positions, reads the next
// half's 8 fragments from LDS into the other set and derives their (ci, cr)
// pairs from the stored (cr, -ci) ones, so LDS traffic and the byte
// shuffles overlap the MFMAs.  Between the halves of step s the wave waits
// for its loads of step s+1 (later steps may stay in flight) and passes a
// barrier that covers every wave's loads; the second half then issues the
// loads of step s+4, spread over its MFMAs, into the stage step s-1 used.  Scheduling barriers hold the order; two
// MFMAs into one accumulator are at least five apart.
// ---------------------------------------------------------------------------
constexpr int kBM = 64, kBN = 64, kGroupM = 4, kStages = 5, kAhead = kStages - 1;
constexpr int kFragsA = kPlanesA * (kBM / 32) * 2, kFragsB = kPlanesB * (kBN / 32) * 2;
constexpr int kFragsPerStage = kFragsA + kFragsB;          // 32
constexpr int kLoadsPerWave = kFragsPerStage / 4;          // 8

// MFMA order of one half: (slice of A, slice of B); both variants follow
// each entry.  Levels si + sj: 3 4 2 3 4 1 3 2 4 3 0 2 1.
struct PairStep { int si, sj; };
constexpr int kPairs = 13;
constexpr int kMfmasHalf = 2 * kPairs;                     // 26
constexpr PairStep kOrder[kPairs] = {{0, 3}, {1, 3}, {0, 2}, {1, 2}, {2, 2}, {0, 1}, {2, 1},
                                     {1, 1}, {3, 1}, {3, 0}, {0, 0}, {2, 0}, {1, 0}};
// positions in a half's MFMA stream: the 8 LDS reads follow MFMAs 0..7, the
// 16 dword shuffles follow MFMAs kRotFirst .. kRotFirst + 15
constexpr int kRotFirst = 9;

struct GemmArgs {
    const int8_t *as, *bs;
    double2 *bmat;           // [ldm][ldn] B / (ra[m] rb[ig]) in FP64
    int ldm, ldn, nkh, tiles_m, tiles_n, group_m;
};

typedef int i16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ i16v mfma_i8(i4 a, i4 b, i16v c)
{
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

// (cr, -ci) byte pairs -> (ci, cr) in one dword: negate the high byte of
// each 16-bit half (two's complement within the half; the carry out of bit
// 15 is dropped by the 16-bit lane), then swap the two bytes of each half.
__device__ __forceinline__ int rotate_pair_dword(int x)
{
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    const unsigned d = (unsigned)x ^ 0xFF00FF00u;
    const u16x2 h = __builtin_bit_cast(u16x2, d) + u16x2{0x0100, 0x0100};
    const unsigned n = __builtin_bit_cast(unsigned, h);
    return (int)__builtin_amdgcn_perm(n, n, 0x02030001u);
}

// Workgroup b -> tile.  Each XCD runs a contiguous range of the logical
// order; the logical order walks bands of group_m row tiles, column by
// column, so a band's coefficient slices stay in the XCD's L2 while the
// aqsn slices stream past.
__device__ __forceinline__ void oz_tile(int b, int tiles_m, int tiles_n, int group_m, int &tm, int &tn)
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

// wait until at most n of this wave's fragment loads are outstanding, then
// pass the workgroup barrier (n = 8 x the steps issued after the awaited one)
__device__ __forceinline__ void wait_loads_barrier(int n)
{
    switch (n) {
    case 24: asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory"); break;
    case 16: asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(8)\n\ts_barrier" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory"); break;
    }
}

struct HalfFrags {
    i4 a[kSlices][2];        // [slice][variant]
    i4 b[kSlices];           // [slice]
};

// one 64x64 tile (workgroup b of the problem's launch)
__device__ __forceinline__ void oz_gemm_tile(const GemmArgs &a, int b)
{
    __shared__ __attribute__((aligned(16))) int8_t lds[kStages][kFragsPerStage][kFrag];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    oz_tile(b, a.tiles_m, a.tiles_n, a.group_m, tm, tn);
    const int nks = a.nkh / 2;

    // global source of this wave's fragment loads: stage fragment f is A
    // (slice f/4, row block 2 tm + (f/2)%2, k-half f%2) for f < 16, else B
    // (slice (f-16)/4, column block 2 tn + ((f-16)/2)%2, k-half f%2);
    // consecutive k-steps are two fragments apart
    const size_t nrb = (size_t)a.ldm / 32, ncb = (size_t)a.ldn / 32;
    const int8_t *src[kLoadsPerWave];
#pragma unroll
    for (int k = 0; k < kLoadsPerWave; ++k) {
        const int f = wave * kLoadsPerWave + k;
        const int8_t *base;
        if (f < kFragsA) {
            const int pl = f >> 2, rb = 2 * tm + ((f >> 1) & 1), kh = f & 1;
            base = a.as + (((size_t)pl * nrb + rb) * a.nkh + kh) * kFrag;
        } else {
            const int g = f - kFragsA, pl = g >> 2, cb = 2 * tn + ((g >> 1) & 1), kh = g & 1;
            base = a.bs + (((size_t)pl * ncb + cb) * a.nkh + kh) * kFrag;
        }
        src[k] = base + 16 * lane;
    }
    auto issue_one = [&](int s, int k) {
        int8_t *dst = &lds[s % kStages][wave * kLoadsPerWave][0];
        __builtin_amdgcn_global_load_lds(src[k] + (size_t)s * 2 * kFrag, dst + k * kFrag, 16, 0, 0);
    };
    auto issue = [&](int s) {
#pragma unroll
        for (int k = 0; k < kLoadsPerWave; ++k) issue_one(s, k);
    };
    // LDS read #k (0..7) of a half's fragments (stage st, k-half kh): B slice
    // 3 - k/2 for even k, A slice k/2 for odd k (the order the MFMAs first
    // need them)
    auto read_frag = [&](const int8_t *st, int kh, HalfFrags &f, int k) {
        const int sl = k >> 1;
        if ((k & 1) == 0)
            f.b[3 - sl] = *reinterpret_cast<const i4 *>(st + (kFragsA + (3 - sl) * 4 + wn * 2 + kh) * kFrag + 16 * lane);
        else
            f.a[sl][0] = *reinterpret_cast<const i4 *>(st + (sl * 4 + wm * 2 + kh) * kFrag + 16 * lane);
    };
    // shuffle #k (0..15): dword k%4 of A slice k/4
    auto rotate = [&](HalfFrags &f, int k) {
        f.a[k >> 2][1][k & 3] = rotate_pair_dword(f.a[k >> 2][0][k & 3]);
    };

    i16v acc[kLevels][2];
#pragma unroll
    for (int L = 0; L < kLevels; ++L)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[L][v] = i16v{};

    // one half: MFMAs from `cur`; with NEXT, the following half's fragments
    // (stage st, k-half kh) into `nxt`; with LOADS, the global->LDS loads of
    // step `ls`, one after every other MFMA
    auto half = [&](auto next_t, auto loads_t, const HalfFrags &cur, HalfFrags &nxt, const int8_t *st, int kh,
                    int ls) {
        constexpr bool NEXT = decltype(next_t)::value;
        constexpr bool LOADS = decltype(loads_t)::value;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int o = 0; o < kMfmasHalf; ++o) {
            const PairStep ps = kOrder[o >> 1];
            const int v = o & 1;
            acc[ps.si + ps.sj][v] = mfma_i8(cur.a[ps.si][v], cur.b[ps.sj], acc[ps.si + ps.sj][v]);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (NEXT) {
                if (o < 8) read_frag(st, kh, nxt, o);
                if (o >= kRotFirst && o < kRotFirst + 16) rotate(nxt, o - kRotFirst);
            }
            if constexpr (LOADS) {
                if ((o & 1) && o < 2 * kLoadsPerWave) issue_one(ls, o >> 1);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    using yes = std::integral_constant<bool, true>;
    using no = std::integral_constant<bool, false>;

    HalfFrags h0, h1;
#pragma unroll
    for (int s = 0; s < kAhead; ++s)
        if (s < nks) issue(s);
    wait_loads_barrier(8 * min(kAhead - 1, nks - 1));
#pragma unroll
    for (int k = 0; k < 8; ++k) read_frag(&lds[0][0][0], 0, h0, k);
#pragma unroll
    for (int k = 0; k < 16; ++k) rotate(h0, k);
    auto stage = [&](int s) { return &lds[s % kStages][0][0]; };
    int s = 0;
    // steady state: step s+1 is awaited with steps s+2, s+3 in flight, and
    // step s+4 exists
    for (; s + kAhead < nks; ++s) {
        half(yes{}, no{}, h0, h1, stage(s), 1, 0);                    // (s, 0); reads (s, 1)
        asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory");
        half(yes{}, yes{}, h1, h0, stage(s + 1), 0, s + kAhead);      // (s, 1); reads (s+1, 0)
    }
    // the last steps: fewer loads in flight, none issued
    for (; s + 1 < nks; ++s) {
        half(yes{}, no{}, h0, h1, stage(s), 1, 0);
        wait_loads_barrier(8 * min(2, nks - 2 - s));
        half(yes{}, no{}, h1, h0, stage(s + 1), 0, 0);
    }
    half(yes{}, no{}, h0, h1, stage(s), 1, 0);
    half(no{}, no{}, h1, h0, stage(s), 0, 0);

    // B = sum_L 254^-L acc_L in FP64 (int32 -> FP64 is exact); the row and
    // column scales are applied where B is used
    const double W[kLevels] = {1.0, 1.0 / 254.0, 1.0 / 64516.0, 1.0 / 16387064.0, 1.0 / 4162314256.0};
    const int col = tn * kBN + wn * 32 + (lane & 31);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = tm * kBM + wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        double bre = 0.0, bim = 0.0;
#pragma unroll
        for (int L = kLevels - 1; L >= 0; --L) {
            bre += W[L] * (double)acc[L][0][j];
            bim += W[L] * (double)acc[L][1][j];
        }
        a.bmat[(size_t)row * a.ldn + col] = make_double2(bre, bim);
    }
}

__global__ __launch_bounds__(256, 1) void gpp_oz_gemm_kernel(GemmArgs a)
{
    oz_gemm_tile(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Pole/SX epilogue over every (my_igp, ig): B from bmat, Bo = sum over the
// occupied bands in FP64, then the FP64 pole and SX factors.  Workgroup b
// takes row my_igp = b / chunks and the (b % chunks)-th of its columns, in
// blocks of kPoleCols, kPoleK columns per thread (tid, tid + 256, ...).  For
// each block it loads the elements once, forms the per-element products
// (B I_eps, Bo Omega2, Bo I_eps) once, and keeps them in registers across
// the frequency passes;
// a pass takes up to 4 frequencies (nw split into ceil(nw / 4) passes of
// near-equal width, so no frequency slot is evaluated for nothing).  The
// occupied coefficients of the row are staged in LDS.  One partial per
// workgroup and frequency, summed in a fixed order (over the threads per
// block, then over the blocks).
// ---------------------------------------------------------------------------
#ifndef FK_POLE_K
#define FK_POLE_K 2
#endif
constexpr int kPoleK = FK_POLE_K;             // columns per thread
constexpr int kPoleCols = 256 * kPoleK;       // columns per workgroup
constexpr int kOccLds = 64;                   // occupied coefficients staged in LDS

struct PoleArgs {
    const double2 *bmat;     // [ldm][ldn] B / (ra[m] rb[ig])
    const double *ra, *rb;   // row and column scales of the slices
    const double2 *c_occ;    // [nv][ldm]
    const float2 *aqsn, *I_eps, *wtilde;
    const int *igp;
    double2 *partials;       // [G * chunks][nw][2]
    double2 *achs_part;      // [G * chunks] static-CH remainder, or NULL
    int nv, C, G, ldm, ldn, chunks;
    PoleParams pp;
};

// LDS of the STREAM form (occupied coefficients, per-wave sums) and of a
// kernel with the BLOCKS form as well (plus the per-block sums)
struct StreamLds {
    double2 occ[kOccLds];
    double red[4][16];
};
struct PoleLds {
    double2 occ[kOccLds];
    double red[4][16];
    double val[16][256];     // [value][thread] of a workgroup sum
    double part[16][16];
};

// Workgroup sum of V <= 16 values per thread, once per workgroup: each
// wave's sum (fixed butterfly), then the four waves in order.  Thread q < V
// returns the total of value q.
template <int V>
__device__ __forceinline__ double wave_block_sum(const double (&v)[V], double (&red)[4][16])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const double x = wave_sum(v[q]);
        if (lane == 0) red[wave][q] = x;
    }
    __syncthreads();
    double x = 0.0;
    if (tid < V) x = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    return x;
}

// Fixed-order workgroup sum of V values per thread: 16 partial sums per
// value over threads s, s + 16, ..., s + 240, then their sum.  Thread q < V
// returns the total of value q.  Two barriers; L.val and L.part may be
// reused after the call.
template <int V>
__device__ __forceinline__ double block_sum(const double (&v)[V], PoleLds &L)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < V; ++q) L.val[q][tid] = v[q];
    __syncthreads();
    if (tid < 16 * V) {
        const int q = tid >> 4, s = tid & 15;
        double x = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) x += L.val[q][s + 16 * j];
        L.part[q][s] = x;
    }
    __syncthreads();
    double x = 0.0;
    if (tid < V) {
#pragma unroll
        for (int s = 0; s < 16; ++s) x += L.part[tid][s];
    }
    return x;
}

// The W frequencies wq[0..W-1] of one element: the per-frequency steps of
// pole_element's hoisted form, every frequency live.
template <int W>
__device__ __forceinline__ void pole_freqs(const PoleParams &pp, const double (&wq)[W], const double (&iq)[W],
                                           double2 wt, double2 Bie, double2 BoOm, double2 Boie, bool want_ssx,
                                           double2 (&sach)[W], double2 (&sasx)[W])
{
    const double tol1 = 1e-13 * pp.limitone, tol2 = 1e-13 * pp.to1;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const double wdr = wq[q] - wt.x;
        const double wdiffr = pole_wdiffr(wdr, -wt.y);
        double2 delw;
        double delwr;
        pole_delw(wt, wdr, rcp_fast<1>(wdiffr), delw, delwr);
        // points within 1e-13 of a branch threshold take the IEEE quotient,
        // so every branch decision is the reference's
        const bool near = fabs(delwr - pp.limitone) <= tol1 || fabs(delwr - pp.to1) <= tol2;
        if (__ballot(near)) {
            if (near) pole_delw(wt, wdr, 1.0 / wdiffr, delw, delwr);
        }
        const bool m = (wdiffr > pp.limittwo) && (delwr < pp.limitone);
        const double2 t = (m || !pp.mask_on) ? cmul(Bie, delw) : make_double2(0.0, 0.0);
        sach[q] = cadd(sach[q], t);
        if (want_ssx) {
            double2 S = m ? make_double2(BoOm.x * iq[q], BoOm.y * iq[q]) : make_double2(0.0, 0.0);
            const bool need2 = !m && delwr > pp.to1;
            if (__ballot(need2)) {
                if (need2) {
                    const double2 z = make_double2(delw.x + 0.5, delw.y);
                    const double sc = -0.25 * rcp_fast<1>(z.x * z.x + z.y * z.y);
                    const double2 v = cmul(Boie, cmul(delw, cconj(z)));
                    S = make_double2(sc * v.x, sc * v.y);
                }
            }
            sasx[q] = cadd(sasx[q], S);
        }
    }
}

struct PoleElems {
    double2 wt[kPoleK], Bie[kPoleK], BoOm[kPoleK], Boie[kPoleK];
};

// The W frequencies w0 .. w0 + W - 1 of the current column block's
// elements, added to this thread's sums
template <int W>
__device__ __forceinline__ void oz_pole_freqs(const PoleArgs &a, int w0, bool want_ssx, const PoleElems &E,
                                              double2 (&sach)[W], double2 (&sasx)[W])
{
    double wq[W], iq[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        wq[q] = a.pp.pc.wx[w0 + q];
        iq[q] = a.pp.pc.iwx2[w0 + q];
    }
#pragma unroll
    for (int e = 0; e < kPoleK; ++e)
        pole_freqs<W>(a.pp, wq, iq, E.wt[e], E.Bie[e], E.BoOm[e], E.Boie[e], want_ssx, sach, sasx);
}

// workgroup total of the thread sums; thread q < 4 W returns value q
// (frequency q / 4; component q % 4: achtemp re, im, asxtemp re, im)
template <int W>
__device__ __forceinline__ double oz_pole_total(const double2 (&sach)[W], const double2 (&sasx)[W], PoleLds &L)
{
    double v[4 * W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        v[4 * q + 0] = sach[q].x;
        v[4 * q + 1] = sach[q].y;
        v[4 * q + 2] = sasx[q].x;
        v[4 * q + 3] = sasx[q].y;
    }
    return block_sum<4 * W>(v, L);
}

// One column block of row m: loads first (all independent), then Bo, then
// the products every pass uses.  Columns past c_hi and skipped rows are
// dead elements: B = Bo = 0 with wtilde = i, so every quotient is finite
// (|wdiff|^2 >= 1) and every term they add is zero.
template <bool ACHS>
__device__ __forceinline__ void oz_pole_elems(const PoleArgs &a, const PoleLds &L, int m, bool ok, bool want_ssx,
                                              int c0, int c_hi, PoleElems &E, double2 &sachs)
{
    const int tid = threadIdx.x;
    const double ram = a.ra[m];
    const double2 *brow = a.bmat + (size_t)m * a.ldn;
    const float2 *wrow = a.wtilde + (size_t)m * a.C, *erow = a.I_eps + (size_t)m * a.C;
    bool live[kPoleK];
    double2 Bu[kPoleK], wt[kPoleK], ie[kPoleK], Bo[kPoleK];
    double rbv[kPoleK];
#pragma unroll
    for (int e = 0; e < kPoleK; ++e) {
        const int ig = c0 + e * 256 + tid;
        live[e] = ok && ig < c_hi;
        Bu[e] = live[e] ? brow[ig] : make_double2(0.0, 0.0);
        rbv[e] = live[e] ? a.rb[ig] : 0.0;
        wt[e] = live[e] ? widen(wrow[ig]) : make_double2(0.0, 1.0);
        ie[e] = live[e] ? widen(erow[ig]) : make_double2(0.0, 0.0);
        Bo[e] = make_double2(0.0, 0.0);
    }
    if (ok && want_ssx) {
        for (int n1 = 0; n1 < a.nv; ++n1) {
            const double2 co = n1 < kOccLds ? L.occ[n1] : a.c_occ[(size_t)n1 * a.ldm + m];
            const float2 *arow = a.aqsn + (size_t)n1 * a.C;
#pragma unroll
            for (int e = 0; e < kPoleK; ++e) {
                const int ig = c0 + e * 256 + tid;
                const double2 x = live[e] ? widen(arow[ig]) : make_double2(0.0, 0.0);
                Bo[e] = cadd(Bo[e], cmul(co, x));
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kPoleK; ++e) {
        const double sc = ram * rbv[e];
        const double2 B = make_double2(Bu[e].x * sc, Bu[e].y * sc);
        const double2 Om = cmul(cmul(wt[e], wt[e]), ie[e]);
        E.wt[e] = wt[e];
        E.Bie[e] = cmul(B, ie[e]);
        E.BoOm[e] = cmul(Bo[e], Om);
        E.Boie[e] = cmul(Bo[e], ie[e]);
        // static-CH remainder -sum B I_eps (gppkernel.deslippe.09.30.2014.f90:192-230)
        if constexpr (ACHS) sachs = csub(sachs, E.Bie[e]);
    }
}

// nw <= 4 (one pass): the chunk's columns one element per thread and
// iteration, each element's products formed and used at once (few live
// registers, so many waves hide the loads); the thread sums are summed
// over the workgroup once
template <int W, bool ACHS, class LDS>
__device__ __forceinline__ void oz_pole_stream(const PoleArgs &a, LDS &L, int b, int m, bool ok,
                                               bool want_ssx, int c_lo, int c_hi, double2 &sachs)
{
    double wq[W], iq[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        wq[q] = a.pp.pc.wx[q];
        iq[q] = a.pp.pc.iwx2[q];
    }
    double2 sach[W], sasx[W];
#pragma unroll
    for (int q = 0; q < W; ++q) { sach[q] = make_double2(0.0, 0.0); sasx[q] = make_double2(0.0, 0.0); }
    const int tid = threadIdx.x;
    if (ok) {
        const double ram = a.ra[m];
        const double2 *brow = a.bmat + (size_t)m * a.ldn;
        const float2 *wrow = a.wtilde + (size_t)m * a.C, *erow = a.I_eps + (size_t)m * a.C;
        const int nvl = min(a.nv, kOccLds);
        for (int ig = c_lo + tid; ig < c_hi; ig += 256) {
            const double2 Bu = brow[ig];
            const double sc = ram * a.rb[ig];
            const double2 B = make_double2(Bu.x * sc, Bu.y * sc);
            const double2 wt = widen(wrow[ig]), ie = widen(erow[ig]);
            double2 Bo = make_double2(0.0, 0.0);
            for (int n1 = 0; n1 < nvl; ++n1) Bo = cadd(Bo, cmul(L.occ[n1], widen(a.aqsn[(size_t)n1 * a.C + ig])));
            for (int n1 = nvl; n1 < a.nv; ++n1)
                Bo = cadd(Bo, cmul(a.c_occ[(size_t)n1 * a.ldm + m], widen(a.aqsn[(size_t)n1 * a.C + ig])));
            const double2 Om = cmul(cmul(wt, wt), ie);
            const double2 Bie = cmul(B, ie);
            pole_freqs<W>(a.pp, wq, iq, wt, Bie, cmul(Bo, Om), cmul(Bo, ie), want_ssx, sach, sasx);
            // static-CH remainder -sum B I_eps (gppkernel.deslippe.09.30.2014.f90:192-230)
            if constexpr (ACHS) sachs = csub(sachs, Bie);
        }
    }
    double v[4 * W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        v[4 * q + 0] = sach[q].x;
        v[4 * q + 1] = sach[q].y;
        v[4 * q + 2] = sasx[q].x;
        v[4 * q + 3] = sasx[q].y;
    }
    const double x = wave_block_sum<4 * W>(v, L.red);
    if (tid < 4 * W) reinterpret_cast<double *>(a.partials + ((size_t)b * W + (tid >> 2)) * 2)[tid & 3] = x;
}

// nw > 4: per block, each pass's sums are summed over the workgroup and
// added to the thread that owns them (thread q < 4 W of pass p: own[p])
template <int W>
__device__ __forceinline__ void oz_pole_pass(const PoleArgs &a, int w0, bool want_ssx, const PoleElems &E,
                                             PoleLds &L, double &own)
{
    double2 sach[W], sasx[W];
#pragma unroll
    for (int q = 0; q < W; ++q) { sach[q] = make_double2(0.0, 0.0); sasx[q] = make_double2(0.0, 0.0); }
    oz_pole_freqs<W>(a, w0, want_ssx, E, sach, sasx);
    own += oz_pole_total<W>(sach, sasx, L);
}

// workgroup b of the problem's launch: row b / chunks, the (b % chunks)-th
// of the row's columns.  Forms: STREAM for nw <= 4, BLOCKS (blocks of
// kPoleCols, near-equal passes) for nw > 4; a kernel built with both picks
// by nw.
constexpr int kPoleStream = 1, kPoleBlocks = 2;
template <bool ACHS, int FORMS, class LDS>
__device__ __forceinline__ void oz_pole_body(const PoleArgs &a, int b, LDS &L)
{
    const int tid = threadIdx.x;
    const int m = __builtin_amdgcn_readfirstlane(b / a.chunks);
    const int ch = __builtin_amdgcn_readfirstlane(b - m * a.chunks);
    const int c_lo = (int)((int64_t)ch * a.C / a.chunks), c_hi = (int)((int64_t)(ch + 1) * a.C / a.chunks);
    const bool ok = a.igp[m] >= 0;
    const bool want_ssx = a.nv > 0;
    const int nvl = min(a.nv, kOccLds);
    for (int n1 = tid; n1 < nvl; n1 += 256) L.occ[n1] = a.c_occ[(size_t)n1 * a.ldm + m];
    __syncthreads();
    const int nw = a.pp.nw, passes = (nw + 3) / 4;
    double2 sachs = make_double2(0.0, 0.0);

    if ((FORMS & kPoleStream) && ((FORMS & kPoleBlocks) == 0 || nw <= 4)) {
        switch (nw) {
        case 1: oz_pole_stream<1, ACHS>(a, L, b, m, ok, want_ssx, c_lo, c_hi, sachs); break;
        case 2: oz_pole_stream<2, ACHS>(a, L, b, m, ok, want_ssx, c_lo, c_hi, sachs); break;
        case 3: oz_pole_stream<3, ACHS>(a, L, b, m, ok, want_ssx, c_lo, c_hi, sachs); break;
        default: oz_pole_stream<4, ACHS>(a, L, b, m, ok, want_ssx, c_lo, c_hi, sachs); break;
        }
    } else if constexpr ((FORMS & kPoleBlocks) != 0) {
        double own[4] = {0.0, 0.0, 0.0, 0.0};     // this thread's totals, one per pass
        for (int c0 = c_lo; c0 < c_hi; c0 += kPoleCols) {
            PoleElems E;
            oz_pole_elems<ACHS>(a, L, m, ok, want_ssx, c0, c_hi, E, sachs);
            for (int p = 0, w0 = 0; p < passes; ++p) {
                const int W = (nw - w0 + (passes - p) - 1) / (passes - p);
                switch (W) {
                case 1: oz_pole_pass<1>(a, w0, want_ssx, E, L, own[p]); break;
                case 2: oz_pole_pass<2>(a, w0, want_ssx, E, L, own[p]); break;
                case 3: oz_pole_pass<3>(a, w0, want_ssx, E, L, own[p]); break;
                default: oz_pole_pass<4>(a, w0, want_ssx, E, L, own[p]); break;
                }
                w0 += W;
            }
        }
        for (int p = 0, w0 = 0; p < passes; ++p) {
            const int W = (nw - w0 + (passes - p) - 1) / (passes - p);
            if (tid < 4 * W) {
                const int q = tid >> 2, comp = tid & 3;
                reinterpret_cast<double *>(a.partials + ((size_t)b * nw + w0 + q) * 2)[comp] = own[p];
            }
            w0 += W;
        }
    }
    if constexpr (ACHS) {
        __syncthreads();     // L.red is free again
        const double v[2] = {sachs.x, sachs.y};
        const double x = wave_block_sum<2>(v, L.red);
        if (tid < 2) reinterpret_cast<double *>(a.achs_part + b)[tid] = x;
    }
}

template <bool ACHS, int FORMS>
__device__ __forceinline__ void oz_pole_block(const PoleArgs &a, int b)
{
    if constexpr (FORMS == kPoleStream) {
        __shared__ StreamLds L;
        oz_pole_body<ACHS, FORMS>(a, b, L);
    } else {
        __shared__ PoleLds L;
        oz_pole_body<ACHS, FORMS>(a, b, L);
    }
}

#ifdef FK_POLE_WAVES
#define FK_POLE_ATTR __attribute__((amdgpu_waves_per_eu(FK_POLE_WAVES, FK_POLE_WAVES)))
#else
#define FK_POLE_ATTR
#endif

template <bool ACHS, int FORMS>
__global__ __launch_bounds__(256) FK_POLE_ATTR void gpp_oz_pole_kernel(PoleArgs a)
{
    oz_pole_block<ACHS, FORMS>(a, blockIdx.x);
}

// column chunks of a row in the pole kernel (one workgroup each)
#ifndef FK_POLE_CHUNKS
#define FK_POLE_CHUNKS 4
#endif
inline int pole_chunks(int C) { return std::min(FK_POLE_CHUNKS, div_up(C, kPoleCols)); }

// ---------------------------------------------------------------------------
// Batches: one launch of each kernel above for many problems.  Every
// problem has a record of its three kernels' arguments; block b of a batched
// launch belongs to the problem p with first[p] <= b < first[p + 1] and runs
// that problem's block b - first[p].  A problem's GEMM range starts at a
// multiple of the XCD count, so its tile order keeps its XCD locality, and
// the blocks padding a range return before any barrier.
// ---------------------------------------------------------------------------
struct OzRecord {
    PrepArgs pr;
    GemmArgs ga;
    PoleArgs pa;
    int gemm_tiles;          // tiles of the problem (its GEMM range may be longer)
};

// the problem whose range of a launch holds block b (first: n + 1 entries)
__device__ __forceinline__ int batch_problem(const int *first, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__global__ __launch_bounds__(256) void oz_prep_batch_kernel(const OzRecord *rec, const int *first, int n)
{
    const int b = (int)blockIdx.x, p = batch_problem(first, n, b);
    oz_prep_block(rec[p].pr, b - first[p]);
}

__global__ __launch_bounds__(256, 1) void gpp_oz_gemm_batch_kernel(const OzRecord *rec, const int *first, int n)
{
    const int b = (int)blockIdx.x, p = batch_problem(first, n, b);
    const int local = b - first[p];
    if (local >= rec[p].gemm_tiles) return;
    oz_gemm_tile(rec[p].ga, local);
}

__global__ __launch_bounds__(256) FK_POLE_ATTR void gpp_oz_pole_batch_kernel(const OzRecord *rec, const int *first, int n)
{
    const int b = (int)blockIdx.x, p = batch_problem(first, n, b);
    oz_pole_block<false, kPoleStream | kPoleBlocks>(rec[p].pa, b - first[p]);
}

}  // namespace

size_t gpp_oz_batch_record_bytes() { return sizeof(OzRecord); }

OzBlocks gpp_oz_batch_record(const fk_gpp_problem *p, const fk_options *o, const OzLayout &L, const Scratch &ws,
                             const int *d_igp, void *record)
{
    OzRecord r;
    PrepArgs &pr = r.pr;
    pr.aqsm = static_cast<const float2 *>(p->aqsmtemp); pr.aqsn = static_cast<const float2 *>(p->aqsntemp);
    pr.vcoul = p->vcoul; pr.igp = d_igp;
    pr.as = static_cast<int8_t *>(ws.at(L.off_as));
    pr.bs = static_cast<int8_t *>(ws.at(L.off_bs));
    pr.ra = static_cast<double *>(ws.at(L.off_ea));
    pr.rb = static_cast<double *>(ws.at(L.off_eb));
    pr.c_occ = static_cast<double2 *>(ws.at(L.off_cocc));
    pr.nb = p->number_bands; pr.nv = p->nvband; pr.C = p->ncouls; pr.G = p->ngpown;
    pr.ldm = L.ldm; pr.ldn = L.ldn; pr.nkh = L.Kp / kKHalf; pr.blocks_a = L.ldm / kPrepRowsA;
    GemmArgs &ga = r.ga;
    ga.as = pr.as; ga.bs = pr.bs;
    ga.bmat = static_cast<double2 *>(ws.at(L.off_bmat));
    ga.ldm = L.ldm; ga.ldn = L.ldn; ga.nkh = pr.nkh;
    ga.tiles_m = L.tiles_m; ga.tiles_n = L.tiles_n; ga.group_m = kGroupM;
    PoleArgs &pa = r.pa;
    pa.bmat = ga.bmat; pa.ra = pr.ra; pa.rb = pr.rb; pa.c_occ = pr.c_occ; pa.aqsn = pr.aqsn;
    pa.I_eps = static_cast<const float2 *>(p->I_eps_array); pa.wtilde = static_cast<const float2 *>(p->wtilde_array);
    pa.igp = d_igp;
    pa.partials = static_cast<double2 *>(ws.at(L.off_part));
    pa.achs_part = nullptr;
    pa.nv = p->nvband; pa.C = p->ncouls; pa.G = p->ngpown; pa.ldm = L.ldm; pa.ldn = L.ldn;
    pa.chunks = pole_chunks(p->ncouls);
    pa.pp = make_pole_params(p->wx, p->nfreq, p->limittwo, p->limitone, p->to1, o->mask != 0);
    r.gemm_tiles = L.tiles_m * L.tiles_n;
    std::memcpy(record, &r, sizeof(r));
    OzBlocks nb;
    const bool any = p->number_bands > 0;
    nb.prep = any ? L.ldm / kPrepRowsA + L.ldn / kPrepColsB : 0;
    nb.gemm = any ? div_up(r.gemm_tiles, kXcds) * kXcds : 0;
    nb.pole = any ? p->ngpown * pa.chunks : 0;
    return nb;
}

int gpp_oz_batch_launch(const void *records, const int *first_prep, const int *first_gemm, const int *first_pole,
                        int n, int total_prep, int total_gemm, int total_pole, hipStream_t s)
{
    const OzRecord *rec = static_cast<const OzRecord *>(records);
    if (total_prep > 0)
        hipLaunchKernelGGL(oz_prep_batch_kernel, dim3(total_prep), dim3(256), 0, s, rec, first_prep, n);
    if (total_gemm > 0)
        hipLaunchKernelGGL(gpp_oz_gemm_batch_kernel, dim3(total_gemm), dim3(256), 0, s, rec, first_gemm, n);
    if (total_pole > 0)
        hipLaunchKernelGGL(gpp_oz_pole_batch_kernel, dim3(total_pole), dim3(256), 0, s, rec, first_pole, n);
    FK_HIP(hipGetLastError());
    return FK_OK;
}

void gpp_oz_plan(const fk_gpp_problem *p, Scratch &ws, OzLayout &L)
{
    const int nb = std::max(p->number_bands, 1);
    L.ldm = div_up(p->ngpown, kBM) * kBM;
    L.ldn = div_up(p->ncouls, kBN) * kBN;
    L.Kp = div_up(2 * nb, kKStep) * kKStep;
    L.tiles_m = L.ldm / kBM;
    L.tiles_n = L.ldn / kBN;
    L.off_as = ws.plan((size_t)kPlanesA * L.ldm * L.Kp);
    L.off_bs = ws.plan((size_t)kPlanesB * L.ldn * L.Kp);
    L.off_ea = ws.plan(sizeof(double) * (size_t)L.ldm);
    L.off_eb = ws.plan(sizeof(double) * (size_t)L.ldn);
    L.off_cocc = ws.plan(sizeof(double2) * (size_t)std::max(p->nvband, 1) * L.ldm);
    L.off_bmat = ws.plan(sizeof(double2) * (size_t)L.ldm * L.ldn);
    L.off_part = ws.plan(sizeof(double2) * 2 * (size_t)p->nfreq * p->ngpown * pole_chunks(p->ncouls));
    L.off_achs = ws.plan(sizeof(double2) * (size_t)p->ngpown * pole_chunks(p->ncouls));
}

int gpp_oz_check(const fk_gpp_problem *p)
{
    if (p->number_bands > kMaxBandsOz) {
        set_error("gpp_kernel: the mixed-precision path takes at most %d bands (got %d)", kMaxBandsOz,
                  p->number_bands);
        return FK_ERR_UNSUPPORTED;
    }
    return FK_OK;
}

int gpp_oz_launch(const fk_gpp_problem *p, const fk_options *o, const OzLayout &L, Scratch &ws, const int *d_igp,
                  const float2 *aqsm, const float2 *aqsn, const float2 *ieps, const float2 *wt, const double *vcoul,
                  double2 *d_ach, double2 *d_asx, double2 *d_achs, hipStream_t s)
{
    const int nks = L.Kp / kKStep, nkh = L.Kp / kKHalf;
    PrepArgs pr;
    pr.aqsm = aqsm; pr.aqsn = aqsn; pr.vcoul = vcoul; pr.igp = d_igp;
    pr.as = static_cast<int8_t *>(ws.at(L.off_as));
    pr.bs = static_cast<int8_t *>(ws.at(L.off_bs));
    pr.ra = static_cast<double *>(ws.at(L.off_ea));
    pr.rb = static_cast<double *>(ws.at(L.off_eb));
    pr.c_occ = static_cast<double2 *>(ws.at(L.off_cocc));
    pr.nb = p->number_bands; pr.nv = p->nvband; pr.C = p->ncouls; pr.G = p->ngpown;
    pr.ldm = L.ldm; pr.ldn = L.ldn; pr.nkh = nkh; pr.blocks_a = L.ldm / kPrepRowsA;
    hipLaunchKernelGGL(oz_prep_kernel, dim3(L.ldm / kPrepRowsA + L.ldn / kPrepColsB), dim3(256), 0, s, pr);
    FK_HIP(hipGetLastError());

    GemmArgs ga;
    ga.as = pr.as; ga.bs = pr.bs;
    ga.bmat = static_cast<double2 *>(ws.at(L.off_bmat));
    ga.ldm = L.ldm; ga.ldn = L.ldn; ga.nkh = nkh;
    ga.tiles_m = L.tiles_m; ga.tiles_n = L.tiles_n; ga.group_m = kGroupM;
    const int tiles = L.tiles_m * L.tiles_n;
    KernelTimer timer(o->kernel_ms, s);
    FK_TRY(timer.start());
    hipLaunchKernelGGL(gpp_oz_gemm_kernel, dim3(tiles), dim3(256), 0, s, ga);
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());

    PoleArgs pa;
    pa.bmat = ga.bmat; pa.ra = pr.ra; pa.rb = pr.rb; pa.c_occ = pr.c_occ; pa.aqsn = aqsn; pa.I_eps = ieps; pa.wtilde = wt; pa.igp = d_igp;
    pa.partials = static_cast<double2 *>(ws.at(L.off_part));
    pa.achs_part = d_achs ? static_cast<double2 *>(ws.at(L.off_achs)) : nullptr;
    pa.nv = p->nvband; pa.C = p->ncouls; pa.G = p->ngpown; pa.ldm = L.ldm; pa.ldn = L.ldn;
    pa.chunks = pole_chunks(p->ncouls);
    pa.pp = make_pole_params(p->wx, p->nfreq, p->limittwo, p->limitone, p->to1, o->mask != 0);
    const int blocks = p->ngpown * pa.chunks;
    if (p->nfreq <= 4) {
        if (pa.achs_part) hipLaunchKernelGGL((gpp_oz_pole_kernel<true, kPoleStream>), dim3(blocks), dim3(256), 0, s, pa);
        else hipLaunchKernelGGL((gpp_oz_pole_kernel<false, kPoleStream>), dim3(blocks), dim3(256), 0, s, pa);
    } else {
        if (pa.achs_part) hipLaunchKernelGGL((gpp_oz_pole_kernel<true, kPoleBlocks>), dim3(blocks), dim3(256), 0, s, pa);
        else hipLaunchKernelGGL((gpp_oz_pole_kernel<false, kPoleBlocks>), dim3(blocks), dim3(256), 0, s, pa);
    }
    FK_HIP(hipGetLastError());
    FK_TRY(gpp_reduce_launch(pa.partials, blocks, p->nfreq, 2.0 * p->occ, d_ach, d_asx, s));
    if (d_achs) FK_TRY(gpp_sum_launch(pa.achs_part, blocks, d_achs, s));
    return timer.finish();
}

}  // namespace fk
