// Allocation-free apply of a Circulant / Toeplitz plan.
//
// No counterpart in the reference.  For callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc): the same pipeline as fm_conv_apply, with the four-step
// intermediate in a fixed scratch array of this library instead of a
// caller-given workspace.
//
// The scratch is one file-scope __device__ array of kScratchBytes (64 MiB):
// every device the library runs on holds its own copy until the process
// ends.  Host code never passes the array's name to a kernel (that is a
// host-side stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, taken once per device with that device current and
// kept per device.  All calls on one device share that device's array, so
// calls that overlap on one device (different streams, or threads) race;
// calls on one stream are ordered.  Columns are processed in chunks whose
// intermediate fits the array; a transform whose single column does not fit
// (L * element size > 64 MiB, i.e. L > 2^23 complex64 / 2^22 complex128) is
// rejected with FM_ERR_VALUE.
#include <mutex>

#include "fm_fft_plan.hpp"

namespace fm {
namespace {
constexpr int64_t kScratchBytes = int64_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device
}  // namespace

// Device address of the current device's scratch array (looked up once per
// device, with that device current).
static fm_status scratch_address(void **out) {
    int dev = 0;
    FM_CHECK_HIP(hipGetDevice(&dev));
    FM_REQUIRE(dev >= 0 && dev < kMaxDevices, FM_ERR_VALUE, "device ordinal %d out of range",
               dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        FM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FM_REQUIRE(p != nullptr, FM_ERR_HIP, "scratch array has no device address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FM_SUCCESS;
}
}  // namespace fm

using namespace fm;

extern "C" fm_status fm_conv_apply_noalloc(const fm_conv_plan *p, int adjoint, const void *x,
                                           int64_t ldx, void *y, int64_t ldy, int64_t k,
                                           void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    if (k == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y, FM_ERR_VALUE, "NULL data pointer");
    const int64_t nin = adjoint ? p->rows : p->cols;
    const int64_t mout = adjoint ? p->cols : p->rows;
    FM_REQUIRE(ldx >= nin && ldy >= mout, FM_ERR_VALUE, "leading dimension too small");
    const int64_t esz = int64_t(esize(p->prec));
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = x; io.ldx = ldx; io.y = y; io.ldy = ldy; io.k = k;
    io.n_in = nin; io.m_out = mout;
    io.D = p->D; io.conj_d = adjoint ? 1 : 0;
    hipStream_t s = as_stream(stream);
    if (p->geo.passes == 1) return run_pipeline_any(p->geo, io, s);
    const int64_t per_col = p->L * esz;
    FM_REQUIRE(per_col <= kScratchBytes, FM_ERR_VALUE,
               "transform length %lld needs %lld bytes of scratch per column; the "
               "allocation-free apply holds %lld (use fm_conv_apply with a workspace)",
               (long long)p->L, (long long)per_col, (long long)kScratchBytes);
    void *work = nullptr;
    fm_status st = scratch_address(&work);
    if (st) return st;
    const int64_t cc = kScratchBytes / per_col;
    for (int64_t c0 = 0; c0 < k; c0 += cc) {
        PipeIO sub = io;
        sub.k = k - c0 < cc ? k - c0 : cc;
        sub.x = (const char *)x + c0 * ldx * esz;
        sub.y = (char *)y + c0 * ldy * esz;
        sub.work = work;
        st = run_pipeline_any(p->geo, sub, s);
        if (st) return st;
    }
    return FM_SUCCESS;
}
