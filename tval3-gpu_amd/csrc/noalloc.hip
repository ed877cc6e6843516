// noalloc.hip -- tv3_matvec_noalloc: the products of tv3_matvec with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Of the products tv3_matvec runs, only the dense A x uses scratch memory:
// the split-K partial sums of its column slabs (dense.hip), which tv3_matvec
// keeps in a context workspace grown by hipMalloc.  Here they go to a
// file-scope device array of fixed size instead.  Each GPU holds its own copy
// of the array until the process ends; its address on a device is looked up
// once, with that device current, when the first context on the device is
// created (tv3_ctx_create), and kept per device.  Every other product (CSR,
// dense A^T x, dynamic, hybrid) uses no scratch and is called as is.
#include <mutex>

#include "tv3_internal.hpp"

namespace tv3 {
namespace {
constexpr size_t kScratchBytes = size_t(16) << 20;   // per GPU
constexpr int kMaxDevices = 64;
}  // namespace
}  // namespace tv3

// The partial sums of the allocation-free dense A x.  In host code this name
// is not a device address: use the per-device pointer from noalloc_setup.
static __device__ __align__(256) unsigned char scratch[tv3::kScratchBytes];

namespace tv3 {
namespace {
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};
}  // namespace

void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    TV3_REQUIRE(device >= 0 && device < kMaxDevices, TV3_E_UNSUPPORTED,
                "device index beyond the allocation-free path's table");
    if (g_scratch[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    TV3_REQUIRE(e == hipSuccess && p != nullptr, TV3_E_HIP,
                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
}

static void *scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

}  // namespace tv3

using namespace tv3;

tv3_status tv3_matvec_noalloc(tv3_op *op, tv3_trans trans, const float *x, float *y,
                              void *stream) {
    if (!op) return TV3_E_ARG;
    tv3_ctx *ctx = op->ctx;
    (void)hipGetLastError();         // an earlier, already reported error is not this call's
    try {
        TV3_REQUIRE(x && y, TV3_E_ARG, "x or y is NULL");
        TV3_REQUIRE(trans == TV3_N || trans == TV3_T, TV3_E_ARG, "bad trans");
        TV3_HIP(hipSetDevice(ctx->device));
        hipStream_t s = (hipStream_t)stream;   // NULL: the null stream
        if (op->kind == OpKind::DENSE && trans == TV3_N) {
            void *ws = scratch_of(op->ctx->device);
            TV3_REQUIRE(ws != nullptr, TV3_E_INTERNAL,
                        "the allocation-free scratch of this device is not set up");
            dense_gemv_ws(op, x, y, s, ws, kScratchBytes);
        } else {
            op_apply(op, trans == TV3_T, x, y, s);
        }
        TV3_HIP(hipGetLastError());
        return TV3_OK;
    } catch (const tv3_error &e) {
        ctx->err = e.what();
        return e.status;
    } catch (const std::exception &e) {
        ctx->err = std::string("internal error: ") + e.what();
        return TV3_E_INTERNAL;
    }
}
