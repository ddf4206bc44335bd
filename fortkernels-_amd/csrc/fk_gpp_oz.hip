// fk_gpp_oz.hip -- mixed-precision GPP band sum on MI355X (gfx950): the
// band contraction B = sum_n1 c[n1] aqsn[n1] runs on int8 MFMA with exact
// int32 accumulation (an Ozaki-style split, Ozaki, Ogita, Oishi & Rump,
// "Error-free transformations of matrix multiplication", Numer. Algorithms
// 59 (2012); on integer matrix units: Ootomo, Ozaki & Yokota, IJHPCA 2024);
// everything else is FP64.
//
// Inputs are complex64 (fk_cfloat).  The complex product runs as three real
// products (Gauss):
//     P0 = cr ar,   P1 = ci ai,   P2 = (cr + ci)(ar + ai),
//     re B = P0 - P1,   im B = P2 - P0 - P1,
// over the matrices cr, ci, cr + ci of the coefficients c[n1][my_igp] =
// 0.5 vcoul[igp] conj(aqsm[n1][igp]) (FP64) and ar, ai, ar + ai of aqsn.
// Each row of each coefficient matrix and each column of each aqsn matrix
// is scaled so that its largest magnitude (M) maps to 127 and split into
// four int8 slices:
//     x = M / 127 * (s0 + s1/254 + s2/254^2 + s3/254^3),  |s_k| <= 127,
// 31 bits of the row's or column's range.  The slice pairs (i, j) with
// i + j = L share the weight 254^-L, so four int32 accumulators per product
// (L = 0..3; the pairs with i + j > 3 are below 254^-4 of the leading term
// and dropped) hold the contraction exactly: 3 x 10 slice pairs, 30 int8
// MFMAs per 32 bands.  The error against the FP64 evaluation on the same
// complex64 values is the slice truncation's, a few 1e-9 relative at most
// (2.4e-9 at 512/2/32768/1638).
//
// Kernels (one call = four launches on the caller's stream):
//   oz_prep_kernel       row / column maxima of the six matrices, their
//                        int8 slices, FP64 occupied rows
//   gpp_oz_gemm_kernel   int8 MFMA tile, operands staged through a
//                        three-stage LDS ring by direct global->LDS loads;
//                        B rebuilt in FP64 and stored
//   gpp_oz_pole_kernel   per (my_igp, ig): occupied-band sum Bo in FP64, the
//                        pole and SX factors in FP64, per-block partials
//   gpp_reduce_kernel    fixed-order sum of the partials (shared with FP64)
//
// Slices are stored in MFMA-lane order: a 1 KB block per (plane, 32 rows,
// 32 bands) in which lane L's 16 bytes (row L%32, bands 16 (L/32) .. +15)
// sit at byte 16 L, so one direct-to-LDS load moves a whole operand
// fragment of the 32x32x32 int8 MFMA and the MFMA reads it back with one
// conflict-free ds_read_b128 per lane.  (The 16x16x64 int8 MFMA runs at
// half the int8 rate on this part; the 32x32x32 form reaches it.)
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
constexpr int kProds = 3;       // P0 = cr ar, P1 = ci ai, P2 = (cr + ci)(ar + ai)
constexpr int kLevels = 4;      // slice pairs with i + j < kLevels
constexpr int kKStep = 32;      // bands per pipeline step = the K of one 32x32x32 int8 MFMA
constexpr int kPlanes = kProds * kSlices;   // slice planes per operand
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
// rows (the first ldm/4 workgroups) or 16 aqsn columns (the rest).  Each
// row (column) feeds three real matrices, cr, ci, cr + ci (ar, ai, ar + ai),
// each with its own scale: the workgroup finds the three maxima over all
// bands, then splits the three values of each band into four slices, band
// block by band block (16 bands = the 16 bytes of one fragment lane); the
// second pass finds the workgroup's slab in L2.  Thread t of an R-wide
// slab: row or column R x + t%R, band blocks t/R, t/R + 256/R, ...
// ---------------------------------------------------------------------------
struct PrepArgs {
    const float2 *aqsm, *aqsn;
    const double *vcoul;
    const int *igp;
    int8_t *as, *bs;         // [12][ldm/32][nkb][1024], [12][ldn/32][nkb][1024]: plane = matrix * 4 + slice
    double *ra, *rb;         // [3][ldm], [3][ldn]: M / 127 of each row, column of each matrix
    double2 *c_occ;          // [nv][ldm] coefficients of the occupied bands
    int nb, nv, C, G, ldm, ldn, nkb, blocks_a;
};

// R rows (coefficients, IS_A) or columns (aqsn) per workgroup, 256 / R band
// groups: the coefficients are gathered at scattered igp columns, so their
// slabs are kept thin to spread the gathers over more workgroups.
template <int R, bool IS_A>
__device__ __forceinline__ void prep_slab(const PrepArgs &a, int slab)
{
    constexpr int NG = 256 / R;
    __shared__ double sm[kProds][NG][R];
    const int tid = threadIdx.x, c = tid % R, grp = tid / R;
    const int idx = slab * R + c;
    const int g = IS_A ? (idx < a.G ? a.igp[idx] : -1) : 0;
    const bool in = IS_A ? g >= 0 : idx < a.C;
    // element n1 of this row (the coefficient c) or column (aqsn)
    auto value = [&](int n1) -> double2 {
        if (!in || n1 >= a.nb) return make_double2(0.0, 0.0);
        if constexpr (IS_A) {
            return coef(a.aqsm, a.vcoul, a.C, n1, g);
        } else {
            const float2 v = a.aqsn[(size_t)n1 * a.C + idx];
            return make_double2((double)v.x, (double)v.y);
        }
    };
    const int nblk = (a.nb + 15) / 16;
    double mx[kProds] = {0.0, 0.0, 0.0};
    for (int j = grp; j < nblk; j += NG)
#pragma unroll 8
        for (int b = 0; b < 16; ++b) {
            const double2 v = value(16 * j + b);
            mx[0] = fmax(mx[0], fabs(v.x));
            mx[1] = fmax(mx[1], fabs(v.y));
            mx[2] = fmax(mx[2], fabs(v.x + v.y));
        }
#pragma unroll
    for (int pr = 0; pr < kProds; ++pr) sm[pr][grp][c] = mx[pr];
    __syncthreads();
    const int ld = IS_A ? a.ldm : a.ldn;
    double scale[kProds];
#pragma unroll
    for (int pr = 0; pr < kProds; ++pr) {
        double m = 0.0;
        for (int k = 0; k < NG; ++k) m = fmax(m, sm[pr][k][c]);
        scale[pr] = m > 0.0 ? 127.0 / m : 0.0;
        if (grp == 0) (IS_A ? a.ra : a.rb)[(size_t)pr * ld + idx] = m / 127.0;
    }

    int8_t *base = IS_A ? a.as : a.bs;
    const size_t blk = idx >> 5, nblks = (size_t)ld / 32;
    // 16-band blocks of the padded K, split in two halves of 8 bands (a
    // scheduling barrier keeps each half's loads and splits together, so
    // only 8 values are live at once; 4 waves per SIMD), then one 16-byte
    // store per plane
    for (int j = grp; j < a.nkb * 2; j += NG) {
        const int lane = (idx & 31) + 32 * (j & 1), kb = j >> 1;
        uint32_t q[kProds][kSlices][4] = {};   // the 16 bytes of each plane, packed
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int n1 = 16 * j + 8 * h + b, byte = 8 * h + b;
                const double2 v = value(n1);
                if (IS_A && n1 < a.nv) a.c_occ[(size_t)n1 * a.ldm + idx] = v;
                const double x[kProds] = {v.x, v.y, v.x + v.y};
#pragma unroll
                for (int pr = 0; pr < kProds; ++pr) {
                    int8_t sl[kSlices];
                    split4(x[pr], scale[pr], sl);
#pragma unroll
                    for (int k = 0; k < kSlices; ++k)
                        q[pr][k][byte >> 2] |= (uint32_t)(uint8_t)sl[k] << (8 * (byte & 3));
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int pr = 0; pr < kProds; ++pr)
#pragma unroll
            for (int k = 0; k < kSlices; ++k) {
                int8_t *dst = base + (((size_t)(pr * kSlices + k) * nblks + blk) * a.nkb + kb) * kFrag + 16 * lane;
                *reinterpret_cast<uint4 *>(dst) = make_uint4(q[pr][k][0], q[pr][k][1], q[pr][k][2], q[pr][k][3]);
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
// int8 MFMA GEMM.  64x64 workgroup tile of 2x2 blocks of 32x32.
// 32x32x32 int8 MFMA: lane l supplies 16 bytes of row l%32 of A and
// of column l%32 of B at the same 16 bands; accumulator register j of lane
// l holds row 8 (j/4) + 4 (l/32) + j%4, column l%32.
//
// The complex product as three real products (Gauss):
//     P0 = cr ar,   P1 = ci ai,   P2 = (cr + ci)(ar + ai),
//     re B = P0 - P1,   im B = P2 - P0 - P1,
// each matrix with its own row (column) scales.  Each product runs the ten
// slice pairs of levels 0..3: 30 MFMAs per 32 bands.  (The complex product
// as two real products over an interleaved K' = 2 nb with one scale per row
// and column needs all 13 pairs of levels 0..4 for the same accuracy: 52
// MFMAs.)
//
// LDS ring of kStages steps of 32 bands; a stage holds the workgroup's 24 A
// fragments ([matrix][slice][row block]) and 24 B fragments
// ([matrix][slice][column block]), 48 KB.  Twelve waves, three per SIMD:
// wave w owns the 32x32 block w % 4 and the product w / 4.  Each wave moves
// 4 fragments per step with direct global->LDS loads, issued over the first
// MFMAs of step s for step s+2, into the stage step s-1 used; after its
// fifth MFMA it waits for its loads of step s+1 and passes a barrier that
// covers every wave's loads, and its next MFMAs carry the 8 LDS reads of
// its fragments of step s+1 into the other register set.  With three waves
// per SIMD the MFMA pipe is fed by whichever wave is ready (one wave per
// SIMD, each computing all three products of its block, left the pipes
// busy 47% of the time).
// ---------------------------------------------------------------------------
constexpr int kBM = 64, kBN = 64, kGroupM = 4, kStages = 3;
constexpr int kFragsA = kPlanes * (kBM / 32), kFragsB = kPlanes * (kBN / 32);   // 24, 24
constexpr int kFragsPerStage = kFragsA + kFragsB;          // 48

// slice pairs of a wave's step in MFMA order; levels si + sj: 3 2 1 3 2 1
// 3 2 0 3 (an accumulator is reused at most every third MFMA)
struct PairStep { int si, sj; };
constexpr int kPairs = 10;
constexpr PairStep kOrder[kPairs] = {{0, 3}, {0, 2}, {0, 1}, {1, 2}, {1, 1},
                                     {1, 0}, {2, 1}, {2, 0}, {0, 0}, {3, 0}};

struct GemmArgs {
    const int8_t *as, *bs;
    const double *ra, *rb;   // [3][ldm], [3][ldn]: M / 127 of each row, column of each matrix
    double2 *bmat;           // [ldm][ldn] B in FP64
    int ldm, ldn, nkb, tiles_m, tiles_n, group_m;
};

typedef int i16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ i16v mfma_i8(i4 a, i4 b, i16v c)
{
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
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


// One direct global->LDS load of 16 bytes per lane: lane l's bytes land at
// LDS byte address lds + 16 l.  Written as inline asm so that the compiler's
// wait insertion, which cannot tell which LDS these loads write, does not
// make every LDS read wait for every load in flight; the kernel orders its
// loads and reads itself with explicit vmcnt waits and barriers.  The asm
// has no register outputs.
__device__ __forceinline__ void lds_load16(const int8_t *g, const int8_t *lds)
{
    // the LDS offset (low half of the shared-aperture address), uniform
    const unsigned la = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)lds);
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(la) : "memory", "m0");
}

// The same from a wave-uniform base (kept in SGPRs) plus a per-lane byte
// offset: one VGPR serves every load of the kernel.
__device__ __forceinline__ void lds_load16(const int8_t *base, unsigned voff, const int8_t *lds)
{
    const unsigned la = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)lds);
    const uint64_t bu = (uint64_t)(uintptr_t)base;
    // readfirstlane yields an int: widen through unsigned (no sign extension)
    const uint64_t ub = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((unsigned)(bu >> 32)) << 32) |
                        (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((unsigned)bu);
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(ub), "s"(la)
                 : "memory", "m0");
}

// ---------------------------------------------------------------------------
// The same 64x64 tile with 12 waves, three per SIMD: wave w owns the 32x32
// block w % 4 and the product w / 4 (4 accumulators, 10 MFMAs per step).
// The LDS ring and its layout are those above; each wave moves 4 of a
// step's 48 fragments and reads the 8 of its block and product.  With three
// waves per SIMD the MFMA pipe is fed by whichever wave is ready.  The
// epilogue meets the three products of a block in the (then free) ring:
// the waves of products 1 and 2 leave P there, those of product 0 form B.
// ---------------------------------------------------------------------------
constexpr int kWaves12 = 12;
constexpr int kLoads12 = kFragsPerStage / kWaves12;   // 4
constexpr int kWait12 = 4;                            // MFMA after which a wave waits
// a wave's 8 fragments in the order of their first use by its step
struct Frag8 { bool a; int sl; };
constexpr Frag8 kRead12[2 * kSlices] = {{true, 0}, {false, 3}, {false, 2}, {false, 1},
                                        {true, 1}, {false, 0}, {true, 2}, {true, 3}};

__device__ __forceinline__ void wait_loads_barrier4(int n)
{
    if (n == 4) asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

struct Frags12 {
    i4 a[kSlices];   // this wave's product, row block
    i4 b[kSlices];   // this wave's product, column block
};

// Tiles first, first + stride, ... (count >= 1 of them) in launch order on
// one workgroup.  With PERSIST, the last two steps of a tile load the first
// two of the next (the last tile reloads its own: one code path), and the
// tile's scales travel as three more direct loads (waves 0..2, 1 KB each)
// into a double-buffered LDS area; without it (one tile per workgroup) the
// last two steps load nothing.  The epilogue forms this wave's scaled P,
// waits for and reads the next tile's step-0 fragments, then meets the
// three products of its block in the ring stage no load is writing (two
// rounds of 8 rows), and the product-0 waves store B.
template <bool PERSIST>
__device__ __forceinline__ void oz_gemm12_tiles(const GemmArgs &a, int first, int stride, int count)
{
    __shared__ __attribute__((aligned(16))) int8_t lds[kStages][kFragsPerStage][kFrag];
    __shared__ __attribute__((aligned(16))) double scl[2][2 * kProds][kBM];   // [tile parity][ra 0..2, rb 0..2][64]

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int blk = wave & 3, pr = wave >> 2;
    const int wm = blk & 1, wn = blk >> 1;
    const int nks = a.nkb;

    // this wave's loads: stage fragment f = 4 wave + k, A (plane f/2, row
    // block 2 tm + f%2) for f < 24, else B (plane (f-24)/2, column block
    // 2 tn + f%2); waves 0..5 move A, 6..11 B.  koff: from the tile's base.
    const bool side_a = kLoads12 * wave < kFragsA;
    const size_t nblk = (size_t)(side_a ? a.ldm : a.ldn) / 32;
    const int8_t *uside = side_a ? a.as : a.bs;
    auto koff = [&](int k) -> size_t {
        const int q = kLoads12 * wave + k - (side_a ? 0 : kFragsA);
        return ((size_t)(q >> 1) * nblk + (q & 1)) * a.nkb * kFrag;
    };
    auto tile_src = [&](int b) -> const int8_t * {
        int tm, tn;
        oz_tile(b, a.tiles_m, a.tiles_n, a.group_m, tm, tn);
        return uside + (size_t)(2 * (side_a ? tm : tn)) * a.nkb * kFrag;
    };
    const unsigned voff = 16 * lane;
    auto stage = [&](int g) -> int8_t * { return &lds[g % kStages][0][0]; };
    auto issue = [&](const int8_t *src, int g, int k) {
        lds_load16(src + koff(k), voff, stage(g) + (kLoads12 * wave + k) * kFrag);
    };
    // scale pieces of tile b (waves 0..2): lanes 0..31 piece 2w, lanes
    // 32..63 piece 2w+1 (0..2 ra rows, 3..5 rb columns), 2 doubles a lane
    auto issue_scales = [&](int b, int t) {
        if (wave < kProds) {
            int tm, tn;
            oz_tile(b, a.tiles_m, a.tiles_n, a.group_m, tm, tn);
            const int q = 2 * wave + (lane >> 5);
            const double *g = q < kProds ? a.ra + (size_t)q * a.ldm + tm * kBM
                                         : a.rb + (size_t)(q - kProds) * a.ldn + tn * kBN;
            lds_load16(reinterpret_cast<const int8_t *>(g + 2 * (lane & 31)),
                       reinterpret_cast<const int8_t *>(&scl[t & 1][2 * wave][0]));
        }
    };
    auto read_frag = [&](const int8_t *st, Frags12 &f, int r) {
        const Frag8 fr = kRead12[r];
        if (fr.a) f.a[fr.sl] = *reinterpret_cast<const i4 *>(st + (pr * 8 + fr.sl * 2 + wm) * kFrag + 16 * lane);
        else f.b[fr.sl] = *reinterpret_cast<const i4 *>(st + (kFragsA + pr * 8 + fr.sl * 2 + wn) * kFrag + 16 * lane);
    };

    i16v acc[kLevels];

    // the g-th step of this workgroup: MFMAs from `cur`; with LOADS, the 4
    // loads from `src` into the stage of step g+2 (and with SCALES tile
    // b's scales after them); with NEXT, the wait for step g+1 after MFMA
    // kWait12 and its fragments read into `nxt`
    auto step = [&](auto next_t, auto loads_t, auto scales_t, const Frags12 &cur, Frags12 &nxt, int g,
                    const int8_t *src, int b, int t) {
        constexpr bool NEXT = decltype(next_t)::value;
        constexpr bool LOADS = decltype(loads_t)::value;
        constexpr bool SCALES = decltype(scales_t)::value;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int o = 0; o < kPairs; ++o) {
            const PairStep ps = kOrder[o];
            acc[ps.si + ps.sj] = mfma_i8(cur.a[ps.si], cur.b[ps.sj], acc[ps.si + ps.sj]);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (LOADS) {
                if (o < kLoads12) issue(src, g + 2, o);
                if constexpr (SCALES) {
                    if (o == kLoads12 - 1) issue_scales(b, t);
                }
            }
            if constexpr (NEXT) {
                if (o == kWait12) wait_loads_barrier4(LOADS ? 4 : 0);
                if (o > kWait12 && o <= kWait12 + kSlices) {
                    read_frag(stage(g + 1), nxt, 2 * (o - kWait12 - 1));
                    read_frag(stage(g + 1), nxt, 2 * (o - kWait12 - 1) + 1);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    using yes = std::integral_constant<bool, true>;
    using no = std::integral_constant<bool, false>;

    // Waits: vmcnt counts this wave's direct loads (4 a step, plus one
    // scale piece for waves 0..2 issued after a step's loads) and B stores;
    // the loads complete in issue order, and a wait allows only the 4 loads
    // issued after the step waited for, so it is correct whatever the
    // stores do.
    Frags12 h0, h1;
    const int8_t *src = tile_src(first);
#pragma unroll
    for (int k = 0; k < kLoads12; ++k) issue(src, 0, k);
    issue_scales(first, 0);
#pragma unroll
    for (int k = 0; k < kLoads12; ++k) issue(src + kFrag, 1, k);
    wait_loads_barrier4(4);
#pragma unroll
    for (int r = 0; r < 2 * kSlices; ++r) read_frag(stage(0), h0, r);

    typedef double d2v __attribute__((ext_vector_type(2)));
    const double W[kLevels] = {1.0, 1.0 / 254.0, 1.0 / 64516.0, 1.0 / 16387064.0};
    int g = 0, b = first;
    for (int t = 0; t < count; ++t, b += stride, g += nks) {
#pragma unroll
        for (int L = 0; L < kLevels; ++L) acc[L] = i16v{};
        int s = 0;
        for (; s + 2 < nks; s += 2) {
            step(yes{}, yes{}, no{}, h0, h1, g + s, src + (size_t)(s + 2) * kFrag, b, t);
            step(yes{}, yes{}, no{}, h1, h0, g + s + 1, src + (size_t)(s + 3) * kFrag, b, t);
        }
        const int8_t *nsrc = src;
        if constexpr (PERSIST) {
            const bool more = t + 1 < count;
            const int bn = more ? b + stride : b;
            nsrc = tile_src(bn);
            step(yes{}, yes{}, yes{}, h0, h1, g + s, nsrc, bn, t + 1);
            step(no{}, yes{}, no{}, h1, h0, g + s + 1, nsrc + kFrag, b, t);
        } else {
            step(yes{}, no{}, no{}, h0, h1, g + s, src, b, t);
            step(no{}, no{}, no{}, h1, h0, g + s + 1, src, b, t);
        }
        src = nsrc;

        // P of this wave's product, scaled
        int tm, tn;
        oz_tile(b, a.tiles_m, a.tiles_n, a.group_m, tm, tn);
        const double *sc = &scl[t & 1][0][0];
        const double rbc = sc[(kProds + pr) * kBM + wn * 32 + (lane & 31)];
        double P[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int r = wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
            double v = 0.0;
#pragma unroll
            for (int L = kLevels - 1; L >= 0; --L) v += W[L] * (double)acc[L][j];
            P[j] = v * sc[pr * kBM + r] * rbc;
        }
        __builtin_amdgcn_sched_barrier(0);
        // the next tile's step-0 fragments (PERSIST), or just a barrier:
        // after it every wave is done with the ring of this tile
        if constexpr (PERSIST) {
            wait_loads_barrier4(4);
#pragma unroll
            for (int r = 0; r < 2 * kSlices; ++r) read_frag(stage(g + nks), h0, r);
        } else {
            wait_loads_barrier4(0);
        }
        // the stage of this tile's last step is free (the next tile's steps
        // 0 and 1 occupy the other two): the products meet there, 8 rows a
        // round, [2 products][4 blocks][8][64] doubles = 32 KB
        double *X = reinterpret_cast<double *>(stage(g + nks - 1));
        const int col = tn * kBN + wn * 32 + (lane & 31);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (pr > 0) {
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) X[(((pr - 1) * 4 + blk) * 8 + jj) * 64 + lane] = P[8 * h + jj];
            }
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            if (pr == 0) {
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int j = 8 * h + jj;
                    const int r = wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
                    const double P1 = X[((0 * 4 + blk) * 8 + jj) * 64 + lane];
                    const double P2 = X[((1 * 4 + blk) * 8 + jj) * 64 + lane];
                    const d2v bv = {P[j] - P1, P2 - P[j] - P1};
                    __builtin_nontemporal_store(bv,
                                                reinterpret_cast<d2v *>(&a.bmat[(size_t)(tm * kBM + r) * a.ldn + col]));
                }
            }
            // the round's values are read before the next round (or the
            // next tile's loads, and the scale buffer's refill) overwrite
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        }
    }
    // a last tile's loads for a next tile may still be in flight: they
    // write LDS, which must not be released before they land
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// one workgroup per CU, each walking its share of the tiles
__global__ __launch_bounds__(768, 1) void gpp_oz_gemm_kernel(GemmArgs a, int tiles)
{
    const int first = blockIdx.x, stride = gridDim.x;
    if (first >= tiles) return;
    oz_gemm12_tiles<true>(a, first, stride, (tiles - 1 - first) / stride + 1);
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
    const double2 *bmat;     // [ldm][ldn] B
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
//
// The fast path forms delw with shorter numerators and |delw|^2 as
// |wtilde|^2 / |wdiff|^2 (equal to the reference's delw.x^2 + delw.y^2 up to
// rounding, a few 1e-15 relative with the one-step reciprocal).  Every point
// whose |delw|^2 falls within 1e-13 of a branch threshold is evaluated again
// in the reference's order with the IEEE quotient, so every branch decision
// is the reference's; |wdiff|^2, which the mask compares with limittwo, is
// always the reference's (contraction off).  Both conditional blocks are
// plain divergent branches, skipped when no lane of the wave takes them (a
// ballot in front of them adds two vector instructions per frequency).
template <int W>
__device__ __forceinline__ void pole_freqs(const PoleParams &pp, const double (&wq)[W], const double (&iq)[W],
                                           double2 wt, double2 Bie, double2 BoOm, double2 Boie, bool want_ssx,
                                           double2 (&sach)[W], double2 (&sasx)[W])
{
    const double tol1 = 1e-13 * pp.limitone, tol2 = 1e-13 * pp.to1;
    double wy2;
    {
#pragma clang fp contract(off)
        wy2 = wt.y * wt.y;                   // wdi^2, as the reference rounds it
    }
    const double wt2 = __builtin_fma(wt.x, wt.x, wy2);
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const double wdr = wq[q] - wt.x;
        double wdiffr;
        {
#pragma clang fp contract(off)
            wdiffr = wdr * wdr + wy2;        // = pole_wdiffr(wdr, -wt.y)
        }
        const double rden = rcp_fast<1>(wdiffr);
        // numerators wt.x wdr - wdi^2 and wt.y (wt.x + wdr) = wt.y wx
        double2 delw = make_double2(__builtin_fma(wt.x, wdr, -wy2) * rden, wt.y * wq[q] * rden);
        double delwr = wt2 * rden;
        const bool near = (int)(fabs(delwr - pp.limitone) <= tol1) | (int)(fabs(delwr - pp.to1) <= tol2);
        if (near) pole_delw(wt, wdr, 1.0 / wdiffr, delw, delwr);
        const bool m = (wdiffr > pp.limittwo) && (delwr < pp.limitone);
        const double2 t = (m || !pp.mask_on) ? cmul(Bie, delw) : make_double2(0.0, 0.0);
        sach[q] = cadd(sach[q], t);
        if (want_ssx) {
            // mask branch: Bo Omega2 / wx^2, added through a selected factor
            // (Bo Omega2 is finite, so a zero factor adds an exact zero)
            const double f = m ? iq[q] : 0.0;
            sasx[q].x = __builtin_fma(BoOm.x, f, sasx[q].x);
            sasx[q].y = __builtin_fma(BoOm.y, f, sasx[q].y);
            if (!m && delwr > pp.to1) {
                const double2 z = make_double2(delw.x + 0.5, delw.y);
                const double sc = -0.25 * rcp_fast<1>(z.x * z.x + z.y * z.y);
                const double2 v = cmul(Boie, cmul(delw, cconj(z)));
                sasx[q].x = __builtin_fma(sc, v.x, sasx[q].x);
                sasx[q].y = __builtin_fma(sc, v.y, sasx[q].y);
            }
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
    const double2 *brow = a.bmat + (size_t)m * a.ldn;
    const float2 *wrow = a.wtilde + (size_t)m * a.C, *erow = a.I_eps + (size_t)m * a.C;
    bool live[kPoleK];
    double2 B[kPoleK], wt[kPoleK], ie[kPoleK], Bo[kPoleK];
#pragma unroll
    for (int e = 0; e < kPoleK; ++e) {
        const int ig = c0 + e * 256 + tid;
        live[e] = ok && ig < c_hi;
        B[e] = live[e] ? brow[ig] : make_double2(0.0, 0.0);
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
        const double2 Om = cmul(cmul(wt[e], wt[e]), ie[e]);
        E.wt[e] = wt[e];
        E.Bie[e] = cmul(B[e], ie[e]);
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
        const double2 *brow = a.bmat + (size_t)m * a.ldn;
        const float2 *wrow = a.wtilde + (size_t)m * a.C, *erow = a.I_eps + (size_t)m * a.C;
        const int nvl = min(a.nv, kOccLds);
        for (int ig = c_lo + tid; ig < c_hi; ig += 256) {
            const double2 B = brow[ig];
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

// Occupancy the pole kernels are held to: 5 waves per SIMD (<= 96 VGPRs)
// for the stream form, 4 (<= 128) for a kernel with the blocks form.  To fit,
// the compiler spills a few registers (1 to 12 a lane) in the set-up and at
// block boundaries, outside the frequency loop; left free, it takes a few
// registers more and loses a wave, which measured slower.
constexpr int pole_waves(int forms) { return forms == kPoleStream ? 5 : 4; }

template <bool ACHS, int FORMS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(pole_waves(FORMS))))
void gpp_oz_pole_kernel(PoleArgs a)
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
// prep or pole launch belongs to the problem p with first[p] <= b <
// first[p + 1] and runs that problem's block b - first[p].  The GEMM launch
// is persistent and walks the problems in turn; a problem's GEMM range
// starts at a multiple of the XCD count, so its tile order keeps its XCD
// locality.
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
    const PrepArgs pr = rec[p].pr;   // in registers: not reloaded after each barrier
    oz_prep_block(pr, b - first[p]);
}

// The GEMM of every problem, persistent: one workgroup per CU walks the
// problems in order and takes, of problem p, the tiles f, f + grid, ... with
// f = (w - first[p]) mod grid.  The rotation continues from problem to
// problem (first[p] is the problem's running tile offset, a multiple of the
// XCD count, so a tile runs on the XCD the single-problem kernel gives it),
// and within a problem the next tile's first steps load during the current
// tile's last two, as in gpp_oz_gemm_kernel.  Between problems a workgroup
// needs no barrier: each wave's LDS slots are its own and every wave drains
// its loads and passes the epilogue's last barrier before returning.
__global__ __launch_bounds__(768, 1) void gpp_oz_gemm_batch_kernel(const OzRecord *rec, const int *first, int n)
{
    const int w = (int)blockIdx.x, grid = (int)gridDim.x;
    for (int p = 0; p < n; ++p) {
        const int tiles = __builtin_amdgcn_readfirstlane(rec[p].gemm_tiles);
        const int f = (w - __builtin_amdgcn_readfirstlane(first[p]) % grid + grid) % grid;
        if (f >= tiles) continue;
        // a copy in registers: the tile's direct loads are asm statements that
        // clobber memory, after which fields read through rec would be reloaded
        const GemmArgs ga = rec[p].ga;
        oz_gemm12_tiles<true>(ga, f, grid, (tiles - 1 - f) / grid + 1);
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(pole_waves(kPoleStream | kPoleBlocks))))
void gpp_oz_pole_batch_kernel(const OzRecord *rec, const int *first, int n)
{
    const int b = (int)blockIdx.x, p = batch_problem(first, n, b);
    // by reference: a register copy of the arguments (with the pole
    // parameters of up to 16 frequencies) made this kernel 1.7x slower
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
    pr.ldm = L.ldm; pr.ldn = L.ldn; pr.nkb = L.Kb / kKStep; pr.blocks_a = L.ldm / kPrepRowsA;
    GemmArgs &ga = r.ga;
    ga.as = pr.as; ga.bs = pr.bs; ga.ra = pr.ra; ga.rb = pr.rb;
    ga.bmat = static_cast<double2 *>(ws.at(L.off_bmat));
    ga.ldm = L.ldm; ga.ldn = L.ldn; ga.nkb = pr.nkb;
    ga.tiles_m = L.tiles_m; ga.tiles_n = L.tiles_n; ga.group_m = kGroupM;
    PoleArgs &pa = r.pa;
    pa.bmat = ga.bmat; pa.c_occ = pr.c_occ; pa.aqsn = pr.aqsn;
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
    if (total_gemm > 0) {
        int dev = 0, ncu = 0;
        FK_HIP(hipGetDevice(&dev));
        FK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int grid = std::max(1, std::min(total_gemm, ncu));
        hipLaunchKernelGGL(gpp_oz_gemm_batch_kernel, dim3(grid), dim3(kWaves12 * 64), 0, s, rec, first_gemm, n);
    }
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
    L.Kb = div_up(nb, 2 * kKStep) * 2 * kKStep;   // an even number of GEMM steps
    L.tiles_m = L.ldm / kBM;
    L.tiles_n = L.ldn / kBN;
    L.off_as = ws.plan((size_t)kPlanes * L.ldm * L.Kb);
    L.off_bs = ws.plan((size_t)kPlanes * L.ldn * L.Kb);
    L.off_ea = ws.plan(sizeof(double) * kProds * (size_t)L.ldm);
    L.off_eb = ws.plan(sizeof(double) * kProds * (size_t)L.ldn);
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
    const int nkb = L.Kb / kKStep;
    PrepArgs pr;
    pr.aqsm = aqsm; pr.aqsn = aqsn; pr.vcoul = vcoul; pr.igp = d_igp;
    pr.as = static_cast<int8_t *>(ws.at(L.off_as));
    pr.bs = static_cast<int8_t *>(ws.at(L.off_bs));
    pr.ra = static_cast<double *>(ws.at(L.off_ea));
    pr.rb = static_cast<double *>(ws.at(L.off_eb));
    pr.c_occ = static_cast<double2 *>(ws.at(L.off_cocc));
    pr.nb = p->number_bands; pr.nv = p->nvband; pr.C = p->ncouls; pr.G = p->ngpown;
    pr.ldm = L.ldm; pr.ldn = L.ldn; pr.nkb = nkb; pr.blocks_a = L.ldm / kPrepRowsA;
    hipLaunchKernelGGL(oz_prep_kernel, dim3(L.ldm / kPrepRowsA + L.ldn / kPrepColsB), dim3(256), 0, s, pr);
    FK_HIP(hipGetLastError());

    GemmArgs ga;
    ga.as = pr.as; ga.bs = pr.bs; ga.ra = pr.ra; ga.rb = pr.rb;
    ga.bmat = static_cast<double2 *>(ws.at(L.off_bmat));
    ga.ldm = L.ldm; ga.ldn = L.ldn; ga.nkb = nkb;
    ga.tiles_m = L.tiles_m; ga.tiles_n = L.tiles_n; ga.group_m = kGroupM;
    const int tiles = L.tiles_m * L.tiles_n;
    KernelTimer timer(o->kernel_ms, s);
    FK_TRY(timer.start());
    int dev = 0, ncu = 0;
    FK_HIP(hipGetDevice(&dev));
    FK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(gpp_oz_gemm_kernel, dim3(std::min(tiles, std::max(ncu, 1))), dim3(kWaves12 * 64), 0, s, ga,
                       tiles);
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());

    PoleArgs pa;
    pa.bmat = ga.bmat; pa.c_occ = pr.c_occ; pa.aqsn = aqsn; pa.I_eps = ieps; pa.wtilde = wt; pa.igp = d_igp;
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
