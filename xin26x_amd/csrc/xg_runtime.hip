// xg_runtime.hip -- error state, logging, per-thread call contexts.
#include "xg_runtime.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>

namespace xg {

static thread_local int32_t t_last_error = XG_OK;
static std::atomic<xg_log_fn> g_logger{nullptr};

void log_msg(int32_t level, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    xg_log_fn fn = g_logger.load();
    if (fn) fn(level, buf);
    else fprintf(stderr, "[xg266:%s] %s\n", level == 0 ? "error" : level == 1 ? "warn" : "info", buf);
}

void set_error(int32_t code, const char *fmt, ...)
{
    t_last_error = code;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    log_msg(0, "%s", buf);
}

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_error(XG_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return false;
}

void pack2d(void *dst, const void *src, intptr_t srcStride, int w, int h, size_t elem)
{
    char *d = (char *)dst;
    const char *s = (const char *)src;
    for (int y = 0; y < h; y++) memcpy(d + (size_t)y * w * elem, s + (intptr_t)y * srcStride * (intptr_t)elem, (size_t)w * elem);
}

void unpack2d(void *dst, intptr_t dstStride, const void *src, int w, int h, size_t elem)
{
    char *d = (char *)dst;
    const char *s = (const char *)src;
    for (int y = 0; y < h; y++) memcpy(d + (intptr_t)y * dstStride * (intptr_t)elem, s + (size_t)y * w * elem, (size_t)w * elem);
}

// Contexts are never destroyed: a thread's stream and buffers live until the
// process ends (destroying them from a thread_local destructor can run after
// the HIP runtime has shut down).
static thread_local CallCtx *t_ctx = nullptr;

CallCtx *call_ctx(size_t devBytes, size_t hostBytes)
{
    if (!device_init()) return nullptr;
    if (!t_ctx) {
        CallCtx *c = new CallCtx();
        if (!XG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking))) {
            delete c;
            return nullptr;
        }
        t_ctx = c;
    }
    CallCtx *c = t_ctx;
    if (devBytes > c->devCap) {
        if (c->dev) (void)hipFree(c->dev);
        c->dev = nullptr;
        c->devCap = 0;
        size_t cap = align_up(devBytes < (1u << 20) ? (1u << 20) : devBytes, 4096);
        if (!XG_HIP(hipMalloc((void **)&c->dev, cap))) {
            set_error(XG_ERR_NOMEM, "device staging of %zu bytes failed", cap);
            return nullptr;
        }
        c->devCap = cap;
    }
    if (hostBytes > c->hostCap) {
        if (c->host) (void)hipHostFree(c->host);
        c->host = nullptr;
        c->hostCap = 0;
        size_t cap = align_up(hostBytes < (1u << 20) ? (1u << 20) : hostBytes, 4096);
        if (!XG_HIP(hipHostMalloc((void **)&c->host, cap, hipHostMallocDefault))) {
            set_error(XG_ERR_NOMEM, "pinned staging of %zu bytes failed", cap);
            return nullptr;
        }
        c->hostCap = cap;
    }
    return c;
}

}  // namespace xg

extern "C" {

int32_t xg_last_error(void) { return xg::t_last_error; }
void xg_clear_error(void) { xg::t_last_error = XG_OK; }
void xg_set_logger(xg_log_fn fn) { xg::g_logger.store(fn); }
const char *xg_version(void) { return "xg266 1.0 gfx950"; }

}  // extern "C"
