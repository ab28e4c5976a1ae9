// Device context, grow-only buffers, events and scans.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace rtt {

Context &ctx()
{
    static Context c;
    return c;
}

int require_ctx()
{
    if (ctx().device < 0 || !ctx().stream)
        return fail(RTT_ERR_INVALID, "rtt_init() has not been called");
    return RTT_OK;
}

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr)
        return RTT_OK;
    release();
    // Round up so that slowly growing requests do not reallocate every call.
    size_t cap = need < 256 ? 256 : need + need / 8;
    hipError_t e = hipMalloc(&ptr, cap);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // Retry at the exact size before giving up.
        e = hipMalloc(&ptr, need < 256 ? 256 : need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            bytes = 0;
            return fail(RTT_ERR_OOM, "hipMalloc(%zu) failed", need);
        }
        cap = need < 256 ? 256 : need;
    }
    bytes = cap;
    return RTT_OK;
}

void DevBuf::release()
{
    // Work queued earlier may still read this buffer: drain the stream first.
    if (ptr && ctx().stream)
        (void)hipStreamSynchronize(ctx().stream);
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int PinnedBuf::ensure(size_t need)
{
    if (need <= bytes && ptr)
        return RTT_OK;
    if (ptr)
        (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (hipHostMalloc(&ptr, need < 64 ? 64 : need, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        return fail(RTT_ERR_OOM, "hipHostMalloc(%zu) failed", need);
    }
    bytes = need < 64 ? 64 : need;
    return RTT_OK;
}

PinnedBuf::~PinnedBuf()
{
    if (ptr)
        (void)hipHostFree(ptr);
}

int Events::create()
{
    if (ready)
        return RTT_OK;
    for (auto &e : ev)
        RTT_HIP(hipEventCreate(&e));
    ready = true;
    return RTT_OK;
}

Events::~Events()
{
    if (ready)
        for (auto &e : ev)
            (void)hipEventDestroy(e);
}

int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, DevBuf *tmp,
                       hipStream_t s)
{
    // out has n + 1 entries; scanning n + 1 inputs needs a trailing zero,
    // which callers provide by sizing `in` with n + 1 entries.
    size_t bytes = 0;
    RTT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, s));
    RTT_TRY(tmp->ensure(bytes));
    RTT_HIP(hipcub::DeviceScan::ExclusiveSum(tmp->ptr, bytes, in, out, n + 1, s));
    return RTT_OK;
}

} // namespace rtt
