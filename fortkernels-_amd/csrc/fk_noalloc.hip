// fk_noalloc.hip -- gpp_kernel without any allocation, for callers that
// cannot allocate device memory (code under graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The workspace of gpp_kernel (coefficient matrix, column map, per-tile
// partials; for the int8-split mixed form also the operand slices and the
// contraction) lives in one file-scope device array of kNoallocScratchBytes.
// Each GPU the library runs on holds its own copy of the array, at its own
// address, until the process ends.  The address is looked up once per
// device (hipGetSymbolAddress, with that device current) on the first call
// there; host code only ever hands that looked-up address to the kernels.
//
// Every call on one device shares that device's array, and the array keeps
// the last call's contents: the kernels write every workspace element they
// read before reading it, so nothing needs clearing, but calls that overlap
// on one device (different streams, or different threads) race on it.
#include "fk_internal.h"

#include <mutex>

namespace fk {
namespace {

constexpr size_t kNoallocScratchBytes = size_t(64) << 20;   // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_address(void **out)
{
    int dev = -1;
    FK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("gpp_kernel_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return FK_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("gpp_kernel_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                      hipGetErrorString(e));
            return FK_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return FK_OK;
}

}  // namespace
}  // namespace fk

using namespace fk;

extern "C" {

size_t gpp_kernel_noalloc_max_workspace(void) { return kNoallocScratchBytes; }

int gpp_kernel_noalloc(const fk_gpp_problem *p, const fk_options *o, fk_cdouble *achtemp, fk_cdouble *asxtemp,
                       fk_cdouble *acht_n1_loc)
{
    clear_error();
    if (!p || !o) { set_error("gpp_kernel_noalloc: NULL problem or options"); return FK_ERR_ARG; }
    if (o->memory != FK_MEM_DEVICE) {
        set_error("gpp_kernel_noalloc: inputs and outputs must be device memory (FK_MEM_DEVICE); host memory "
                  "needs staging buffers");
        return FK_ERR_UNSUPPORTED;
    }
    if (acht_n1_loc) {
        set_error("gpp_kernel_noalloc: acht_n1_loc is not available (its pole-factor matrix is ngpown x ncouls)");
        return FK_ERR_UNSUPPORTED;
    }
    if (o->workspace) {
        set_error("gpp_kernel_noalloc: takes no workspace (it uses the library's own); call gpp_kernel instead");
        return FK_ERR_ARG;
    }
    if (o->kernel_ms) {
        set_error("gpp_kernel_noalloc: kernel timing creates events; call gpp_kernel to time");
        return FK_ERR_UNSUPPORTED;
    }
    const size_t need = gpp_kernel_workspace_size(p, o);
    if (need == 0) return FK_ERR_ARG;   // bad sizes, pointers or options; the message is set
    if (need > kNoallocScratchBytes) {
        set_error("gpp_kernel_noalloc: the problem needs %zu bytes of workspace, more than the %zu bytes of the "
                  "library's device scratch", need, kNoallocScratchBytes);
        return FK_ERR_WORKSPACE;
    }
    void *ptr = nullptr;
    FK_TRY(scratch_address(&ptr));
    fk_options d = *o;
    d.workspace = ptr;
    d.workspace_bytes = kNoallocScratchBytes;
    return gpp_kernel(p, &d, achtemp, asxtemp, nullptr);
}

}  // extern "C"
