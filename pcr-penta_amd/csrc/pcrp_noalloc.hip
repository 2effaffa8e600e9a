// Plain FP64 solve that allocates nothing and takes no workspace argument:
// the split path's workspace (n above the LDS threshold) is a fixed array in
// device memory, reserved on every GPU the library runs on until the process
// ends. For callers that cannot allocate (code under graph capture, serving
// loops, harnesses that forbid hipMalloc). No counterpart in the reference,
// whose driver allocates everything itself (main.cu:25-30).
//
// Every call on one GPU uses that GPU's one array: calls that overlap in
// time on one GPU (several streams or threads) race on it and must be
// serialised by the caller. Nothing in the array has to start at zero (the
// split passes write every element they or the LDS stage later read), so it
// is never cleared; it keeps the last call's intermediates.
#include "pcr_penta.h"
#include "pcrp_internal.h"

#include <mutex>

namespace pcrp {

namespace {

constexpr size_t kScratchBytes = PCRP_NOALLOC_SCRATCH_BYTES;

// The scratch array. Host code never uses this name as an address (on the
// host it is a stand-in symbol): the device address comes from
// hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void* g_scratch[kMaxDevices];
hipError_t g_status[kMaxDevices];

// The current device's address of `scratch`, looked up once per device with
// that device current.
int scratch_address(void** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PCRP_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return PCRP_ERR_UNSUPPORTED;
    std::call_once(g_once[dev], [dev] {
        void* p = nullptr;
        g_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_status[dev] == hipSuccess ? p : nullptr;
    });
    if (g_status[dev] != hipSuccess || g_scratch[dev] == nullptr) return PCRP_ERR_HIP;
    *out = g_scratch[dev];
    return PCRP_OK;
}

}  // namespace

}  // namespace pcrp

extern "C" {

int pcrp_solve_f64_noalloc(const double* a, const double* b, const double* c, const double* d,
                           const double* e, double* y, int n, int m, int layout, void* stream)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (!a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    if (layout != PCRP_LAYOUT_CONTIGUOUS && layout != PCRP_LAYOUT_INTERLEAVED)
        return PCRP_ERR_LAYOUT;
    const size_t need = pcrp_workspace_size(n, m, PCRP_PREC_F64);
    if (need > pcrp::kScratchBytes) return PCRP_ERR_SCRATCH;
    void* ws = nullptr;
    if (need > 0) {
        const int rc = pcrp::scratch_address(&ws);
        if (rc != PCRP_OK) return rc;
    }
    return pcrp_solve_f64(a, b, c, d, e, y, n, m, layout, ws, need, stream);
}

}  // extern "C"
