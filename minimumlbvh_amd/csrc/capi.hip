// Context lifetime, error plumbing and host-side helpers of the C ABI.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "mlbvh_ctx.h"

namespace mlbvh {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_status(hipError_t e, const char* what)
{
    set_error("%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return MLBVH_E_HIP;
}

int set_device(mlbvh_ctx* ctx)
{
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_status(e, "hipSetDevice");
    return MLBVH_OK;
}

}  // namespace mlbvh

using namespace mlbvh;

extern "C" int mlbvh_abi_version(void) { return MLBVH_ABI_VERSION; }

extern "C" const char* mlbvh_last_error(void) { return g_err; }

extern "C" int mlbvh_ctx_create(int device, mlbvh_ctx** out)
{
    if (!out) {
        set_error("ctx_create: null out pointer");
        return MLBVH_E_INVALID_ARG;
    }
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= device || device < 0) {
        set_error("ctx_create: device %d not available (%d devices, %s)", device, count,
                  e == hipSuccess ? "ok" : hipGetErrorString(e));
        return MLBVH_E_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_status(e, "hipGetDeviceProperties");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("ctx_create: device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
        return MLBVH_E_NO_DEVICE;
    }
    mlbvh_ctx* ctx = new (std::nothrow) mlbvh_ctx();
    if (!ctx) {
        set_error("ctx_create: out of host memory");
        return MLBVH_E_INVALID_ARG;
    }
    ctx->device = device;
    ctx->cuCount = prop.multiProcessorCount;
    int rc = set_device(ctx);
    if (!rc) rc = noalloc_setup(device);  // the allocation-free build's scratch, once per device
    size_t one = 0, cap = 0;
    if (!rc) rc = ensure(ctx, &ctx->hist, &one, (size_t)8 * 256, "hipMalloc(hist)");
    cap = 0;
    if (!rc) rc = ensure(ctx, &ctx->ctrl, &cap, (size_t)kCtrlWords, "hipMalloc(ctrl)");
    if (!rc) {
        hipError_t e = hipMemset(ctx->ctrl, 0, kCtrlWords * sizeof(uint32_t));
        if (e != hipSuccess) rc = hip_status(e, "hipMemset(ctrl)");
    }
    cap = 0;
    if (!rc) rc = ensure(ctx, &ctx->sceneBox, &cap, (size_t)kSceneBoxWords, "hipMalloc(sceneBox)");
    if (!rc) {  // zeroed slots carry tag 0, which no build uses
        hipError_t e = hipMemset(ctx->sceneBox, 0, kSceneBoxWords * sizeof(uint64_t));
        if (e != hipSuccess) rc = hip_status(e, "hipMemset(sceneBox)");
    }
    if (rc) {
        mlbvh_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return MLBVH_OK;
}

extern "C" int mlbvh_ctx_destroy(mlbvh_ctx* ctx)
{
    if (!ctx) return MLBVH_OK;
    hipSetDevice(ctx->device);
    void* bufs[] = {ctx->status, ctx->hist,     ctx->ctrl,     ctx->recA,    ctx->recB,
                    ctx->deltas, ctx->flags,    ctx->sceneBox, ctx->meshBox, ctx->partials, ctx->frontier,
                    ctx->idxA,   ctx->idxB,     ctx->superList, ctx->blockCounts, ctx->l1Status, ctx->topList};
    int rc = MLBVH_OK;
    for (void* p : bufs) {
        if (!p) continue;
        hipError_t e = hipFree(p);
        if (e != hipSuccess && rc == MLBVH_OK) rc = hip_status(e, "hipFree");
    }
    delete ctx;
    return rc;
}

// RayGenerator::lookat, camera.h:10-22, evaluated on the host in IEEE single
// precision in the reference's operation order.
extern "C" int mlbvh_raygen_lookat(mlbvh_raygen* out, const float eye[3], const float center[3], const float up[3],
                                   float fovy, int width, int height)
{
    if (!out || !eye || !center || !up || width <= 0 || height <= 0) {
        set_error("raygen_lookat: invalid argument");
        return MLBVH_E_INVALID_ARG;
    }
    struct V {
        float x, y, z;
    };
    auto sub = [](V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto muls = [](V a, float s) { return V{a.x * s, a.y * s, a.z * s}; };
    auto dotp = [](V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    auto crossp = [](V a, V b) { return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
    auto norm = [&](V v) { return muls(v, 1.0f / sqrtf(dotp(v, v))); };
    V e{eye[0], eye[1], eye[2]}, c{center[0], center[1], center[2]}, u0{up[0], up[1], up[2]};
    V f = norm(sub(c, e));
    V s = norm(crossp(f, u0));
    V u = crossp(s, f);
    float tanThetaY = tanf(fovy * 0.5f);
    float tanThetaX = tanThetaY / (float)height * (float)width;
    V r = muls(s, tanThetaX), uu = muls(u, tanThetaY);
    out->origin[0] = e.x;
    out->origin[1] = e.y;
    out->origin[2] = e.z;
    out->right[0] = r.x;
    out->right[1] = r.y;
    out->right[2] = r.z;
    out->up[0] = uu.x;
    out->up[1] = uu.y;
    out->up[2] = uu.z;
    return MLBVH_OK;
}
