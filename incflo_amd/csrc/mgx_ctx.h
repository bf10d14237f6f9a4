// mgx_ctx.h -- context object behind the mgx_ctx handle, and the C-ABI
// error-translation helper.
#pragma once

#include <string>

#include "mgx_common.h"

struct mgx_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    hipEvent_t events[16] = {};
    bool events_made = false;
    void make_current() const { (void)hipSetDevice(device); }
};

namespace mgx {

// Run `f`, translating exceptions into a status code + message on the context.
template <class F>
int guarded(mgx_ctx_s* ctx, F&& f) {
    try {
        if (ctx) ctx->make_current();
        f();
        return MGX_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        return MGX_ERR_ARG;
    }
}

inline unsigned grid1d(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace mgx
