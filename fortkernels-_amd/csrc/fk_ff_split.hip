// fk_ff_split.hip -- the FF contraction recomputed in split BF16 on MFMA (gfx950).
//
// The GEMM of the FF contraction (ffkernel.deslippe.best.f90.back1:402-418),
//     T[n1][m] = sum_ig aqsn[n1][ig] X[m][ig],        X = I_eps as [nfreq*ngpown][ncouls],
// is evaluated from the same FP64 arrays with each real operand rounded to
// FP32 and written as three BF16 slices, x ~ x0 + x1 + x2, where x0 is the
// BF16 rounding of x, x1 that of x - x0, x2 that of x - x0 - x1 (every
// difference is exact in FP32).  The six slice products with i + j <= 2 run
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation; the three dropped
// products are below 2^-24 of |x||y|.
//
// FP64 refinement: a workgroup accumulates one chunk of ig (1024 by default)
// in FP32 and contracts its tile in FP64,
//     schDt part[n1] = sum_m cf[n1][m] T[n1][m],   schDv part[n1] = sum_m vg[m] cf[n1][m] T[n1][m],
// and the chunks and column tiles are summed in FP64, in a fixed order, by
// the reduction of fk_ff.hip.  FP32 rounding therefore spans one chunk only.
//
// The complex product as two real GEMMs over K' = 2*ncouls with re and im of
// each ig interleaved:
//     A[n1][2ig] = Re a,    A[n1][2ig+1] = Im a                  (a = aqsn)
//     B1[m][2ig] = Re x,    B1[m][2ig+1] = -Im x     Re T = A B1^T
//     B2[m][2ig] = Im x,    B2[m][2ig+1] = Re x      Im T = A B2^T
// X is read in its own layout; each slice dword (Re x_s, Im x_s) gives B1 by
// flipping the sign bit of its high half and B2 by swapping its halves.
//
// Tile: 256 bands x 64 columns of one frequency x one ig chunk per
// workgroup of 4 waves; wave w owns 64 bands (2 x 2 blocks of 32x32, re and
// im: 128 FP32 accumulators).  X is read from HBM once: 32 ig per stage,
// split in registers and written as BF16 slices (B1 and B2) to a
// double-buffered LDS stage that the four waves share.  Aqsn is split once
// by ff_split_a_kernel into MFMA fragment order, and each wave loads its own
// fragments from L2 one stage ahead.  The MFMAs of a k-step run in six
// groups of eight; the splitting of X and the loads ride between them.
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>

namespace fk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXcds = 8;
constexpr int kSlices = 3;
constexpr int kStageIg = 32;                 // ig per stage (K' = 64)
constexpr int kSteps = 4;                    // MFMA k-steps (K' = 16) per stage
constexpr int kTileN = 64;                   // columns per workgroup
constexpr int kTileM = 256;                  // bands per workgroup
constexpr int kWaves = 4;
constexpr int kRowDw = 36;                   // LDS row of one slice: 32 dwords + 4 pad
constexpr int kSliceDw = kTileN * kRowDw;    // one slice of B1 or B2 for one stage
constexpr int kBufDw = 2 * kSlices * kSliceDw;   // one stage: [B1, B2][slice][column][ig]
constexpr int kEpiRow = 33;                  // epilogue row: 32 float2 + 1 pad
constexpr int kEpiWave = 64 * kEpiRow;       // float2 per wave
constexpr int kLdsBytes = (2 * kBufDw * 4 > kWaves * kEpiWave * 8) ? 2 * kBufDw * 4 : kWaves * kEpiWave * 8;

// (bf16(lo), bf16(hi)) in one dword, round to nearest even
__device__ __forceinline__ uint32_t cvt_pk_bf16(float lo, float hi)
{
    uint32_t d;
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(d) : "v"(lo), "v"(hi));
    return d;
}

// (x, y) -> three slice dwords (x_s, y_s), s = 0, 1, 2
__device__ __forceinline__ void split3(float x, float y, uint32_t (&d)[kSlices])
{
#pragma unroll
    for (int s = 0; s < kSlices; ++s) {
        d[s] = cvt_pk_bf16(x, y);
        if (s + 1 < kSlices) {
            x -= __uint_as_float(d[s] << 16);
            y -= __uint_as_float(d[s] & 0xffff0000u);
        }
    }
}

__device__ __forceinline__ f32x16 mfma(uint4 a, uint4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// As[((q*RF + r)*3 + s)*64 + L] (16 bytes): slice s of A for fragment row
// 32r + L%32 and K' = 16q + 8(L/32) .. +7.  Zero past nb and past ncouls.
__global__ void ff_split_a_kernel(const double2 *aqsn, int nb, int C, int nq, int RF, uint4 *As)
{
    const int64_t n = (int64_t)nq * RF * 64;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        const int L = (int)(idx % 64);
        const int r = (int)((idx / 64) % RF);
        const int q = (int)(idx / (64 * RF));
        const int row = 32 * r + (L & 31);
        const int ig0 = 8 * q + 4 * (L >> 5);
        uint32_t d[4][kSlices];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ig = ig0 + e;
            double2 v = make_double2(0.0, 0.0);
            if (row < nb && ig < C) v = aqsn[(size_t)row * C + ig];
            split3((float)v.x, (float)v.y, d[e]);
        }
#pragma unroll
        for (int s = 0; s < kSlices; ++s)
            As[(((size_t)q * RF + r) * kSlices + s) * 64 + L] = make_uint4(d[0][s], d[1][s], d[2][s], d[3][s]);
    }
}

struct SplitArgs {
    const uint4 *As;        // split aqsn (ff_split_a_kernel)
    const double2 *X;       // [nf*G][C]
    const double2 *cf;      // [nb][G]
    const double *vg;       // [G]
    double2 *partials;      // [nf][nkc*ngt][nb][2]
    int nb, C, G, nf;
    int RF;                 // fragment rows of As
    int nstages;            // stages over ncouls
    int ns;                 // stages per chunk
    int nkc, ngt, nmt, total;
};

__global__ __launch_bounds__(256, 1) void ff_split_gemm_kernel(SplitArgs a)
{
    __shared__ __align__(16) uint32_t lds[kLdsBytes / 4];

    // workgroup -> tile: each XCD takes a contiguous range of tiles; f runs
    // fastest, so the workgroups that run together on one XCD share one
    // cf block and one chunk of As in its L2
    int t;
    {
        const int b = blockIdx.x, q = a.total / kXcds, r = a.total % kXcds, x = b % kXcds, i = b / kXcds;
        t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    }
    const int f = t % a.nf;
    int rest = t / a.nf;
    const int gt = rest % a.ngt;
    rest /= a.ngt;
    const int mt = rest % a.nmt;
    const int kc = rest / a.nmt;

    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    const int s0 = kc * a.ns;
    const int nst = min(a.ns, a.nstages - s0);
    // fragment rows of this wave; a wave past the last band repeats the last
    // two fragments and its results are dropped
    const int r0 = min(8 * mt + 2 * w, a.RF - 2);

    // X elements this thread loads: columns 16w + 2j + (L >> 5) of the tile
    // (j = 0..7, clamped to the last column), ig = stage*32 + (L & 31)
    // clamped to the last ig (where the split aqsn is zero).  Addresses are
    // 32-bit offsets from a per-wave base.
    const int kk = L & 31;
    const int wcol = min(gt * kTileN + 16 * w, a.G - 1);
    const double2 *xbase = a.X + ((size_t)f * a.G + wcol) * a.C;
    uint32_t xoff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int col = min(gt * kTileN + 16 * w + 2 * j + (L >> 5), a.G - 1);
        xoff[j] = (uint32_t)(col - wcol) * (uint32_t)a.C;
    }
    auto load_x = [&](int s, int j) -> double2 {
        const uint32_t ig = (uint32_t)min((s0 + s) * kStageIg + kk, a.C - 1);
        return xbase[xoff[j] + ig];
    };
    // split one X element into the B1 and B2 slices of LDS stage `buf`
    auto store_x = [&](int buf, int j, double2 v) {
        uint32_t d[kSlices];
        split3((float)v.x, (float)v.y, d);
        uint32_t *p = lds + buf * kBufDw + (16 * w + 2 * j + (L >> 5)) * kRowDw + kk;
#pragma unroll
        for (int s = 0; s < kSlices; ++s) {
            p[s * kSliceDw] = d[s] ^ 0x80000000u;                                  // B1: (Re x, -Im x)
            p[(kSlices + s) * kSliceDw] = __builtin_rotateright32(d[s], 16);      // B2: (Im x, Re x)
        }
    };
    const uint4 *As = a.As + (size_t)r0 * kSlices * 64 + L;
    const size_t qstride = (size_t)a.RF * kSlices * 64;   // uint4 per k-step
    auto load_a = [&](int s, int q, int i, int sl) -> uint4 {
        return As[((size_t)(s0 + s) * kSteps + q) * qstride + (i * kSlices + sl) * 64];
    };

    f32x16 acc[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[i][j][c][e] = 0.f;

    // A fragments of one stage (one set per k-step) and X of the next stage
    uint4 av[kSteps][2][kSlices];
    double2 xv[8];
#pragma unroll
    for (int q = 0; q < kSteps; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int sl = 0; sl < kSlices; ++sl) av[q][i][sl] = load_a(0, q, i, sl);
#pragma unroll
    for (int j = 0; j < 8; ++j) store_x(0, j, load_x(0, j));
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = load_x(min(1, nst - 1), j);
    __syncthreads();

    for (int s = 0; s < nst; ++s) {
        const uint32_t *buf = lds + (s & 1) * kBufDw;
        const int sn = min(s + 1, nst - 1), sx = min(s + 2, nst - 1);
#pragma unroll
        for (int q = 0; q < kSteps; ++q) {
            uint4 b[2][2][kSlices];   // [B1, B2][column block][slice]
#pragma unroll
            for (int v = 0; v < 2; ++v)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int sl = 0; sl < kSlices; ++sl)
                        b[v][j][sl] = *reinterpret_cast<const uint4 *>(buf + (v * kSlices + sl) * kSliceDw +
                                                                    (32 * j + (L & 31)) * kRowDw + 8 * q +
                                                                    4 * (L >> 5));
            // slice pairs (A, B) in six groups: (0,0) (0,1) (0,2) (1,0) (1,1) (2,0)
            int g = 0;
#pragma unroll
            for (int sa = 0; sa < kSlices; ++sa) {
#pragma unroll
                for (int sb = 0; sb < kSlices - sa; ++sb, ++g) {
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            acc[i][j][0] = mfma(av[q][i][sa], b[0][j][sb], acc[i][j][0]);
                            acc[i][j][1] = mfma(av[q][i][sa], b[1][j][sb], acc[i][j][1]);
                        }
                    // work between the groups: X of stage s+1 split into the
                    // other LDS buffer (last read before the previous
                    // barrier), each element reloaded with stage s+2 at once
                    // (k-steps 1 and 2), and the A fragments of this k-step
                    // reloaded with stage s+1 after their last use
                    const int jx = (q == 1) ? g : (q == 2 && g < 2) ? 6 + g : -1;
                    if (jx >= 0) {
                        store_x((s + 1) & 1, jx, xv[jx]);
                        xv[jx] = load_x(sx, jx);
                    }
                    if (sb == kSlices - sa - 1) {
#pragma unroll
                        for (int i = 0; i < 2; ++i) av[q][i][sa] = load_a(sn, q, i, sa);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        __syncthreads();
    }

    // epilogue: T (FP32) -> LDS, one band per thread, FP64 contraction over
    // the tile's columns in ascending order
    float2 *E = reinterpret_cast<float2 *>(lds) + w * kEpiWave;
    const int n1 = mt * kTileM + 64 * w + L;
    const bool own = n1 < a.nb;
    double2 st = make_double2(0.0, 0.0), sv = st;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = 32 * i + 8 * (e >> 2) + 4 * (L >> 5) + (e & 3);
                E[row * kEpiRow + (L & 31)] = make_float2(acc[i][j][0][e], acc[i][j][1][e]);
            }
        __syncthreads();
        const int c0 = gt * kTileN + 32 * j;
        const int cn = min(32, a.G - c0);
        if (own) {
            const double2 *cfr = a.cf + (size_t)n1 * a.G + c0;
            for (int c = 0; c < cn; ++c) {
                const float2 T = E[L * kEpiRow + c];
                const double2 tt = cmul(cfr[c], make_double2((double)T.x, (double)T.y));
                const double v = a.vg[c0 + c];
                st = cadd(st, tt);
                sv = cadd(sv, make_double2(v * tt.x, v * tt.y));
            }
        }
        __syncthreads();
    }
    if (own) {
        const int mtiles = a.nkc * a.ngt;
        double2 *p = a.partials + (((size_t)f * mtiles + kc * a.ngt + gt) * a.nb + n1) * 2;
        p[0] = st;
        p[1] = sv;
    }
}

// stages per ig chunk for fk_options.config = 0 (default), 1, 2, 3: ig
// chunks of 1024, 256, 512 and 1024.  At 256/4096/4096/32 on MI355X the
// chunk of 1024 runs 5.4 ms against 5.7 ms for 512 (fewer workgroups, each
// paying one prologue and one epilogue), with error 6.6e-7 against 4.9e-7.
constexpr int kSplitNs[] = {32, 8, 16, 32};

}  // namespace

int ff_split_configs() { return (int)(sizeof(kSplitNs) / sizeof(kSplitNs[0])) - 1; }

void ff_split_plan(int nb, int C, int G, int nf, int config, Scratch &ws, FfSplitLayout &L)
{
    L.nstages = div_up(C, kStageIg);
    L.ns = std::min(kSplitNs[std::max(0, std::min(config, ff_split_configs()))], L.nstages);
    L.nkc = div_up(L.nstages, L.ns);
    L.RF = 2 * div_up(nb, 64);
    L.nmt = div_up(nb, kTileM);
    L.ngt = div_up(G, kTileN);
    L.total = L.nkc * L.nmt * L.ngt * nf;
    const size_t nq = (size_t)L.nstages * kSteps;
    L.off_as = ws.plan(nq * L.RF * kSlices * 64 * sizeof(uint4));
}

int ff_split_launch(const FfSplitLayout &L, Scratch &ws, const double2 *aqsn, const double2 *X, const double2 *cf,
                    const double *vg, double2 *partials, int nb, int C, int G, int nf, KernelTimer &timer,
                    hipStream_t s)
{
    uint4 *As = static_cast<uint4 *>(ws.at(L.off_as));
    const int nq = L.nstages * kSteps;
    {
        const int64_t n = (int64_t)nq * L.RF * 64;
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(ff_split_a_kernel, dim3(blocks), dim3(256), 0, s, aqsn, nb, C, nq, L.RF, As);
        FK_HIP(hipGetLastError());
    }
    SplitArgs a;
    a.As = As; a.X = X; a.cf = cf; a.vg = vg; a.partials = partials;
    a.nb = nb; a.C = C; a.G = G; a.nf = nf;
    a.RF = L.RF; a.nstages = L.nstages; a.ns = L.ns;
    a.nkc = L.nkc; a.ngt = L.ngt; a.nmt = L.nmt; a.total = L.total;
    FK_TRY(timer.start());
    hipLaunchKernelGGL(ff_split_gemm_kernel, dim3(L.total), dim3(64 * kWaves), 0, s, a);
    FK_HIP(hipGetLastError());
    return timer.stop();
}

}  // namespace fk
