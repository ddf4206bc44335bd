// fk_util.hip -- C-ABI utilities: errors, options, device scratch, the
// counter-based synthetic fill and the complex-helper entry point.
#include "fk_complex.h"
#include "fk_internal.h"

#include <cmath>
#include <cstring>

namespace fk {

static thread_local std::string g_error;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

void clear_error() { g_error.clear(); }

int resolve_columns(int ncouls, int ngpown, const int32_t *inv_igp_index, const int32_t *indinv,
                    std::vector<int32_t> &igp_out)
{
    igp_out.resize(ngpown);
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
        int indigp = inv_igp_index[my_igp];
        if (indigp < 0 || indigp > ncouls) {
            set_error("inv_igp_index[%d] = %d is outside [0, ncouls=%d]", my_igp, indigp, ncouls);
            return FK_ERR_INDEX;
        }
        int igp = indinv[indigp];
        if (indigp == ncouls) igp = ncouls - 1;
        if (igp > ncouls || igp < 0) { igp_out[my_igp] = -1; continue; }
        if (igp == ncouls) {
            set_error("column %d maps to igp = ncouls = %d (out of bounds)", my_igp, ncouls);
            return FK_ERR_INDEX;
        }
        igp_out[my_igp] = igp;
    }
    return FK_OK;
}

Scratch::~Scratch()
{
    if (owned_ && base_) (void)hipFreeAsync(base_, stream_);
}

int Scratch::acquire(void *user, size_t user_bytes, hipStream_t stream)
{
    stream_ = stream;
    if (total_ == 0) return FK_OK;
    if (user) {
        if (user_bytes < total_) {
            set_error("workspace of %zu bytes is smaller than the %zu bytes needed", user_bytes, total_);
            return FK_ERR_WORKSPACE;
        }
        base_ = user;
        owned_ = false;
        return FK_OK;
    }
    hipError_t e = hipMallocAsync(&base_, total_, stream);
    if (e != hipSuccess) {
        base_ = nullptr;
        set_error("hipMallocAsync(%zu) failed: %s", total_, hipGetErrorString(e));
        return FK_ERR_NOMEM;
    }
    owned_ = true;
    return FK_OK;
}

KernelTimer::KernelTimer(float *out, hipStream_t s) : out_(out), s_(s) {}
KernelTimer::~KernelTimer()
{
    if (a_) (void)hipEventDestroy(a_);
    if (b_) (void)hipEventDestroy(b_);
}
int KernelTimer::start()
{
    if (!out_) return FK_OK;
    FK_HIP(hipEventCreate(&a_));
    FK_HIP(hipEventCreate(&b_));
    FK_HIP(hipEventRecord(a_, s_));
    return FK_OK;
}
int KernelTimer::stop()
{
    if (!out_) return FK_OK;
    FK_HIP(hipEventRecord(b_, s_));
    return FK_OK;
}
int KernelTimer::finish()
{
    if (!out_) return FK_OK;
    FK_HIP(hipEventSynchronize(b_));
    float ms = 0.f;
    FK_HIP(hipEventElapsedTime(&ms, a_, b_));
    *out_ = ms;
    return FK_OK;
}

// ---------------------------------------------------------------------------
// Counter-based synthetic fill (SplitMix64 finaliser keyed by seed/stream).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform_at(uint64_t key, uint64_t index, double lo, double span)
{
    uint64_t h = mix64(key + (index + 1) * 0x9E3779B97F4A7C15ULL);
    double u = (double)(h >> 11) * 0x1.0p-53;
    return __builtin_fma(span, u, lo);
}

static uint64_t host_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint64_t fill_key(uint64_t seed, uint64_t stream)
{
    return host_mix64(seed * 0x9E3779B97F4A7C15ULL + host_mix64(stream + 0x632BE59BD9B4E019ULL));
}

template <bool AS_F32>
__global__ void fill_complex_kernel(void *dst, int64_t n, uint64_t key, double re_lo, double re_span,
                                    double im_lo, double im_span)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        double re = uniform_at(key, (uint64_t)(2 * j), re_lo, re_span);
        double im = uniform_at(key, (uint64_t)(2 * j + 1), im_lo, im_span);
        if (AS_F32) static_cast<float2 *>(dst)[j] = make_float2((float)re, (float)im);
        else static_cast<double2 *>(dst)[j] = make_double2(re, im);
    }
}

__global__ void fill_real_kernel(double *dst, int64_t n, uint64_t key, double lo, double span)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
        dst[j] = uniform_at(key, (uint64_t)j, lo, span);
}

static int fill_grid(int64_t n)
{
    int64_t blocks = (n + 255) / 256;
    return (int)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
}

// ---------------------------------------------------------------------------
// Complex helpers, host and device instantiations of the same functions.
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 apply_op(int op, double2 a, double2 b, double s, double2 acc)
{
    switch (op) {
    case FK_COP_MUL: return cmul(a, b);
    case FK_COP_CONJ: return cconj(a);
    case FK_COP_ABS2: return make_double2(cabs2(a), 0.0);
    case FK_COP_ABS: return make_double2(cabs(a), 0.0);
    case FK_COP_SCALE: return cscale(a, s);
    case FK_COP_FMA: return cfma(acc, a, b);
    case FK_COP_FMS: return cfms(acc, a, b);
    case FK_COP_REAL_MINUS: return real_minus(s, a);
    case FK_COP_REAL_PLUS: return real_plus(s, a);
    case FK_COP_DIV: return cdiv(a, b);
    case FK_COP_DIV_REAL: return cdiv(a, s);
    case FK_COP_SQUARE: return csquare(a);
    case FK_COP_ADD: return cadd(a, b);
    case FK_COP_SUB: return csub(a, b);
    default: return make_double2(0.0, 0.0);
    }
}

__global__ void complex_apply_kernel(int op, int64_t n, const double2 *a, const double2 *b, const double *s,
                                     double2 *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 z = make_double2(0.0, 0.0);
    out[i] = apply_op(op, a ? a[i] : z, b ? b[i] : z, s ? s[i] : 0.0, out[i]);
}

// One 16x16x4 MFMA on fixed operands: A[i][k] = 1 + i + 16 k,
// B[k][j] = 1 + 3 j - k; lane l holds A[l%16][l/16] and B[l/16][l%16], and
// out[l*4 + r] is accumulator register r of lane l.  kind 0: FP64 MFMA,
// kind 1: FP32 MFMA (values converted to double).
__global__ void mfma_probe_kernel(int kind, double *out)
{
    typedef double d4 __attribute__((ext_vector_type(4)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    int l = threadIdx.x;
    double a = 1.0 + (l % 16) + 16.0 * (l / 16);
    double b = 1.0 + 3.0 * (l % 16) - (l / 16);
    if (kind == 0) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[l * 4 + r] = acc[r];
    } else {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, acc, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[l * 4 + r] = (double)acc[r];
    }
}

}  // namespace fk

using namespace fk;

extern "C" {

fk_options fk_default_options(void)
{
    fk_options o;
    std::memset(&o, 0, sizeof o);
    o.memory = FK_MEM_HOST;
    o.precision = FK_PREC_FP64;
    o.mask = 1;
    o.device = -1;
    return o;
}

const char *fk_last_error(void) { return g_error.c_str(); }

int fk_abi_version(void) { return FK_ABI_VERSION; }

const char *fk_build_target(void) { return "gfx950"; }

int fk_fill_complex(void *dst, int64_t n, uint64_t seed, uint64_t stream_id, double re_lo, double re_hi,
                    double im_lo, double im_hi, int32_t as_complex64, void *stream)
{
    clear_error();
    if (n < 0 || (n > 0 && !dst)) { set_error("fk_fill_complex: bad arguments"); return FK_ERR_ARG; }
    if (n == 0) return FK_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint64_t key = fill_key(seed, stream_id);
    if (as_complex64)
        hipLaunchKernelGGL(fill_complex_kernel<true>, dim3(fill_grid(n)), dim3(256), 0, s, dst, n, key, re_lo,
                           re_hi - re_lo, im_lo, im_hi - im_lo);
    else
        hipLaunchKernelGGL(fill_complex_kernel<false>, dim3(fill_grid(n)), dim3(256), 0, s, dst, n, key, re_lo,
                           re_hi - re_lo, im_lo, im_hi - im_lo);
    FK_HIP(hipGetLastError());
    return FK_OK;
}

int fk_fill_real(double *dst, int64_t n, uint64_t seed, uint64_t stream_id, double lo, double hi, void *stream)
{
    clear_error();
    if (n < 0 || (n > 0 && !dst)) { set_error("fk_fill_real: bad arguments"); return FK_ERR_ARG; }
    if (n == 0) return FK_OK;
    hipLaunchKernelGGL(fill_real_kernel, dim3(fill_grid(n)), dim3(256), 0, static_cast<hipStream_t>(stream), dst,
                       n, fill_key(seed, stream_id), lo, hi - lo);
    FK_HIP(hipGetLastError());
    return FK_OK;
}

int fk_complex_apply(int32_t op, int64_t n, const fk_cdouble *a, const fk_cdouble *b, const double *s,
                     fk_cdouble *out, int32_t on_device)
{
    clear_error();
    if (op < FK_COP_MUL || op > FK_COP_SUB || n < 0 || (n > 0 && !out)) {
        set_error("fk_complex_apply: bad arguments");
        return FK_ERR_ARG;
    }
    if (n == 0) return FK_OK;
    const double2 *A = reinterpret_cast<const double2 *>(a), *B = reinterpret_cast<const double2 *>(b);
    double2 *O = reinterpret_cast<double2 *>(out);
    if (!on_device) {
        double2 z = make_double2(0.0, 0.0);
        for (int64_t i = 0; i < n; ++i) O[i] = apply_op(op, A ? A[i] : z, B ? B[i] : z, s ? s[i] : 0.0, O[i]);
        return FK_OK;
    }
    size_t zb = sizeof(double2) * (size_t)n;
    double2 *dA = nullptr, *dB = nullptr, *dO = nullptr;
    double *dS = nullptr;
    int rc = FK_OK;
    hipError_t e = hipSuccess;
    if (A && e == hipSuccess) e = hipMalloc(&dA, zb);
    if (B && e == hipSuccess) e = hipMalloc(&dB, zb);
    if (s && e == hipSuccess) e = hipMalloc(&dS, sizeof(double) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&dO, zb);
    if (A && e == hipSuccess) e = hipMemcpy(dA, A, zb, hipMemcpyHostToDevice);
    if (B && e == hipSuccess) e = hipMemcpy(dB, B, zb, hipMemcpyHostToDevice);
    if (s && e == hipSuccess) e = hipMemcpy(dS, s, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dO, O, zb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(complex_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, op, n,
                           dA, dB, dS, dO);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(O, dO, zb, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("fk_complex_apply: %s", hipGetErrorString(e));
        rc = FK_ERR_HIP;
    }
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dS) (void)hipFree(dS);
    if (dO) (void)hipFree(dO);
    return rc;
}

// Diagnostic: raw accumulator of one 16x16x4 MFMA (256 doubles, host
// memory); the tests compare it with the layout the GEMM kernels assume.
int fk_mfma_probe(int32_t kind, double *out)
{
    clear_error();
    if (!out || (kind != 0 && kind != 1)) { set_error("fk_mfma_probe: bad arguments"); return FK_ERR_ARG; }
    double *d = nullptr;
    FK_HIP(hipMalloc(&d, 256 * sizeof(double)));
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, nullptr, (int)kind, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) { set_error("fk_mfma_probe: %s", hipGetErrorString(e)); return FK_ERR_HIP; }
    return FK_OK;
}

}  // extern "C"
