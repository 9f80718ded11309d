// post.hip -- the renderer's post stack on the device: guide decode, edge-aware a-trous
// denoise, FXAA, TXAA resolve and tone mapping to RGBA8, on the frame the integrator left in
// HBM, so a render returns 4 bytes per pixel instead of the 28 of linear rgba + compact guides.
//
// Restates (reference = windows_portable/): decodeOctNormal and the compact guide decode
// (gpu_opencl_windows.go:552-562, 686-716), luminance (renderer.go:2142-2144), atrousDenoise
// (:2228-2287), edgeAwareSharpen (:2146-2183), applyFXAA / applyTXAAResolve (:2557-2595),
// toneMapHDR / toRGBA (:2033-2064), in the order of RenderState.Start with HBAO, temporal reuse,
// bloom, debug views and upscaling off (:2842-2895).  Arithmetic is fp32 (the Go code is fp64).
// Every kernel is one thread per pixel over a 2-D launch; the stencils read neighbours that
// the workgroup's own pixels share, so they are served by L1/L2, and each pass is HBM-bound at
// 16 B in + 16 B out per pixel (plus 16 B of decoded guide for the denoiser).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/beamcast_hip.h"
#include "api_internal.hpp"

namespace bcd {

constexpr int kPostX = 64, kPostY = 4;

__device__ __forceinline__ float plum(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float pclamp(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }

// Compact guides -> (normal.xyz, depth); depth < 0 marks a pixel without a primary hit.
__global__ __launch_bounds__(kPostX* kPostY) void k_post_guides(const uint32_t* __restrict__ g, int w, int h,
                                                                float4* __restrict__ out) {
    const int x = blockIdx.x * kPostX + threadIdx.x, y = blockIdx.y * kPostY + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    const uint32_t packed = g[3 * i + 1];
    if (!(packed & 0x80000000u)) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        return;
    }
    float nx = (float)(packed & 1023u) / 1023.0f * 2.0f - 1.0f;
    float ny = (float)((packed >> 10) & 1023u) / 1023.0f * 2.0f - 1.0f;
    float nz = 1.0f - fabsf(nx) - fabsf(ny);
    if (nz < 0.0f) {
        const float ox = nx, oy = ny;
        nx = (1.0f - fabsf(oy)) * copysignf(1.0f, ox);
        ny = (1.0f - fabsf(ox)) * copysignf(1.0f, oy);
    }
    const float l = sqrtf(nx * nx + ny * ny + nz * nz);
    const float il = l < 1e-12f ? 0.0f : 1.0f / l;
    out[i] = make_float4(nx * il, ny * il, nz * il, __uint_as_float(g[3 * i]));
}

// One a-trous pass with tap spacing `step`.
__global__ __launch_bounds__(kPostX* kPostY) void k_atrous(const float4* __restrict__ cur, float4* __restrict__ next,
                                                           const float4* __restrict__ G, int w, int h, int step) {
    const int x = blockIdx.x * kPostX + threadIdx.x, y = blockIdx.y * kPostY + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    const float4 center = cur[i], g0 = G[i];
    const bool v0 = g0.w >= 0.0f;
    const float l0 = plum(center);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, wsum = 0.0f;
    for (int oy = -1; oy <= 1; oy++) {
        const int yy = y + oy * step;
        if (yy < 0 || yy >= h) continue;
        for (int ox = -1; ox <= 1; ox++) {
            const int xx = x + ox * step;
            if (xx < 0 || xx >= w) continue;
            const size_t j = (size_t)yy * (size_t)w + (size_t)xx;
            const float4 g1 = G[j];
            if (v0 != (g1.w >= 0.0f)) continue;
            const float4 q = cur[j];
            float wgt = 1.0f;
            if (v0) {
                const float nd = pclamp(g0.x * g1.x + g0.y * g1.y + g0.z * g1.z, -1.0f, 1.0f);
                wgt *= expf((nd - 1.0f) * 48.0f);
                const float ds = fmaxf(0.15f, 0.04f * fmaxf(g0.w, 1.0f));
                wgt *= expf(-fabsf(g0.w - g1.w) / ds);
            }
            wgt *= expf(-fabsf(l0 - plum(q)) * 3.5f);
            if (ox == 0 && oy == 0) wgt *= 1.5f;
            sx += q.x * wgt;
            sy += q.y * wgt;
            sz += q.z * wgt;
            wsum += wgt;
        }
    }
    if (wsum > 1e-12f) {
        const float iw = 1.0f / wsum;
        next[i] = make_float4(sx * iw, sy * iw, sz * iw, center.w);
    } else {
        next[i] = center;
    }
}

__global__ __launch_bounds__(kPostX* kPostY) void k_fxaa(const float4* __restrict__ img, float4* __restrict__ out, int w,
                                                         int h, float threshold, float blend) {
    const int x = blockIdx.x * kPostX + threadIdx.x, y = blockIdx.y * kPostY + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    const float4 c = img[i];
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) {
        out[i] = c;
        return;
    }
    const float4 pn = img[i - w], ps = img[i + w], pw = img[i - 1], pe = img[i + 1];
    const float ln = plum(pn), ls = plum(ps), lw = plum(pw), le = plum(pe), lc = plum(c);
    const float lo = fminf(lc, fminf(fminf(ln, ls), fminf(lw, le)));
    const float hi = fmaxf(lc, fmaxf(fmaxf(ln, ls), fmaxf(lw, le)));
    if (hi - lo < threshold * fmaxf(1.0f, hi)) {
        out[i] = c;
        return;
    }
    const float ax = (((pn.x + ps.x) + pw.x) + pe.x) * 0.25f;
    const float ay = (((pn.y + ps.y) + pw.y) + pe.y) * 0.25f;
    const float az = (((pn.z + ps.z) + pw.z) + pe.z) * 0.25f;
    const float k = 1.0f - blend;
    out[i] = make_float4(c.x * k + ax * blend, c.y * k + ay * blend, c.z * k + az * blend, c.w);
}

__global__ __launch_bounds__(kPostX* kPostY) void k_sharpen(const float4* __restrict__ img, float4* __restrict__ out, int w,
                                                            int h, float amount) {
    const int x = blockIdx.x * kPostX + threadIdx.x, y = blockIdx.y * kPostY + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    const float4 center = img[i];
    const float cl = plum(center);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, count = 0.0f, ew = 0.0f;
    for (int oy = -1; oy <= 1; oy++) {
        const int yy = y + oy;
        if (yy < 0 || yy >= h) continue;
        for (int ox = -1; ox <= 1; ox++) {
            const int xx = x + ox;
            if (xx < 0 || xx >= w) continue;
            const float4 nb = img[(size_t)yy * (size_t)w + (size_t)xx];
            sx += nb.x;
            sy += nb.y;
            sz += nb.z;
            count += 1.0f;
            ew += fabsf(cl - plum(nb));
        }
    }
    const float local = amount * pclamp(ew / 2.0f, 0.35f, 1.0f);
    const float ic = 1.0f / fmaxf(count, 1.0f);
    out[i] = make_float4(fmaxf(0.0f, center.x + (center.x - sx * ic) * local),
                         fmaxf(0.0f, center.y + (center.y - sy * ic) * local),
                         fmaxf(0.0f, center.z + (center.z - sz * ic) * local), center.w);
}

__device__ __forceinline__ float tone(float x, int hdr) {
    if (hdr == 0) return pclamp(x, 0.0f, 1.0f);
    if (hdr == 1) return x / (1.0f + x);
    if (hdr == 3) return pclamp((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
    x = fmaxf(0.0f, x - 0.004f);
    return pclamp((x * (6.2f * x + 0.5f)) / (x * (6.2f * x + 1.7f) + 0.06f), 0.0f, 1.0f);
}

__global__ __launch_bounds__(kPostX* kPostY) void k_to_rgba(const float4* __restrict__ img, uint32_t* __restrict__ out,
                                                            int w, int h, float escale, float inv_gamma, int hdr) {
    const int x = blockIdx.x * kPostX + threadIdx.x, y = blockIdx.y * kPostY + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * (size_t)w + (size_t)x;
    const float4 c = img[i];
    const float r = powf(fmaxf(0.0f, tone(c.x * escale, hdr)), inv_gamma);
    const float g = powf(fmaxf(0.0f, tone(c.y * escale, hdr)), inv_gamma);
    const float b = powf(fmaxf(0.0f, tone(c.z * escale, hdr)), inv_gamma);
    const uint32_t R = (uint32_t)(pclamp(r, 0.0f, 0.999f) * 256.0f);
    const uint32_t Gc = (uint32_t)(pclamp(g, 0.0f, 0.999f) * 256.0f);
    const uint32_t Bc = (uint32_t)(pclamp(b, 0.0f, 0.999f) * 256.0f);
    out[i] = 0xff000000u | (Bc << 16) | (Gc << 8) | R;
}

}  // namespace bcd

// ------------------------------------------------------------------------------ host side
struct PostState {
    void* buf[4] = {};  // ping, pong, decoded guides, pixels (bc_render_post)
    size_t cap[4] = {};
};

void bc_post_release(bc_ctx* c) {
    if (!c || !c->ps) return;
    for (void* p : c->ps->buf)
        if (p) (void)hipFree(p);
    delete c->ps;
    c->ps = nullptr;
}

extern "C" int bc_postprocess_device(bc_ctx* c, const bc_post_desc* pd, int w, int h, const float* d_rgba,
                                     const uint32_t* d_guides, uint32_t* d_pixels, float* d_linear) {
    if (!c || !pd || !d_rgba || w <= 0 || h <= 0 || (int64_t)w * h > ((int64_t)1 << 28)) return BC_ERR_ARG;
    if (pd->denoise_passes < 0 || pd->denoise_passes > 16 || pd->fxaa < 0 || pd->txaa < 0 ||
        (pd->denoise_passes > 0 && !d_guides) || (!d_pixels && !d_linear)) {
        c->err = "invalid post-process description";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    if (!c->ps) c->ps = new PostState();
    PostState& S = *c->ps;
    const size_t np = (size_t)w * (size_t)h;
    for (int k = 0; k < 3; k++)
        if (!bc_ensure(c, S.buf[k], S.cap[k], np * 16, "post-process buffers")) return BC_ERR_NOMEM;
    float4* ping = (float4*)S.buf[0];
    float4* pong = (float4*)S.buf[1];
    float4* G = (float4*)S.buf[2];
    hipStream_t st = c->stream;
    const dim3 blk(bcd::kPostX, bcd::kPostY), grd((unsigned)((w + bcd::kPostX - 1) / bcd::kPostX),
                                                   (unsigned)((h + bcd::kPostY - 1) / bcd::kPostY));
    const float4* cur = (const float4*)d_rgba;
    auto out_buf = [&](const float4* in) { return in == ping ? pong : ping; };
    if (pd->denoise_passes > 0) {
        hipLaunchKernelGGL(bcd::k_post_guides, grd, blk, 0, st, d_guides, w, h, G);
        for (int it = 0; it < pd->denoise_passes; it++) {
            float4* nx = out_buf(cur);
            hipLaunchKernelGGL(bcd::k_atrous, grd, blk, 0, st, cur, nx, (const float4*)G, w, h, 1 << it);
            cur = nx;
        }
    }
    static const float thr[4] = {0.0f, 0.16f, 0.10f, 0.065f}, bl[4] = {0.0f, 0.35f, 0.50f, 0.65f};
    if (pd->fxaa > 0 && w >= 3 && h >= 3) {
        const int lv = std::min(pd->fxaa, 3);
        float4* nx = out_buf(cur);
        hipLaunchKernelGGL(bcd::k_fxaa, grd, blk, 0, st, cur, nx, w, h, thr[lv], bl[lv]);
        cur = nx;
    }
    if (pd->txaa > 0) {  // applyTXAAResolve: FXAA one level up, then a light sharpen
        const int lv = std::min(pd->txaa + 1, 3);
        if (w >= 3 && h >= 3) {
            float4* nx = out_buf(cur);
            hipLaunchKernelGGL(bcd::k_fxaa, grd, blk, 0, st, cur, nx, w, h, thr[lv], bl[lv]);
            cur = nx;
        }
        float4* nx = out_buf(cur);
        hipLaunchKernelGGL(bcd::k_sharpen, grd, blk, 0, st, cur, nx, w, h, 0.05f + 0.04f * (float)pd->txaa);
        cur = nx;
    }
    if (d_pixels)
        hipLaunchKernelGGL(bcd::k_to_rgba, grd, blk, 0, st, cur, d_pixels, w, h, exp2f(pd->exposure),
                           1.0f / std::max(0.1f, pd->gamma), pd->hdr);
    if (d_linear && d_linear != (const float*)cur &&
        !bc_check(c, hipMemcpyAsync(d_linear, cur, np * 16, hipMemcpyDeviceToDevice, st), "post output"))
        return BC_ERR_DEVICE;
    if (!bc_check(c, hipGetLastError(), "post-process launch") || !bc_check(c, hipStreamSynchronize(st), "post-process"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

extern "C" int bc_render_post(bc_ctx* c, const bc_render_desc* d, const bc_post_desc* pd, uint32_t* pixels,
                              float* linear, uint64_t* rays_out) {
    if (!c || !d || !pd || !pixels) return BC_ERR_ARG;
    const size_t np = (size_t)d->width * (size_t)std::max(d->height, 0);
    if (d->width <= 0 || d->height <= 0) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    if (!bc_ensure(c, c->out_buf, c->out_cap, np * 28 + 64, "render output")) return BC_ERR_NOMEM;
    float* d_rgba = (float*)c->out_buf;
    uint32_t* d_g = (uint32_t*)((char*)c->out_buf + np * 16);
    int rc = bc_render_device(c, d, 0, d->height, d_rgba, pd->denoise_passes > 0 ? d_g : nullptr, rays_out, nullptr);
    if (rc != BC_OK) return rc;
    if (!c->ps) c->ps = new PostState();
    if (!bc_ensure(c, c->ps->buf[3], c->ps->cap[3], np * 4 + 64, "pixel buffer")) return BC_ERR_NOMEM;
    uint32_t* d_px = (uint32_t*)c->ps->buf[3];
    rc = bc_postprocess_device(c, pd, d->width, d->height, d_rgba, d_g, d_px, linear ? d_rgba : nullptr);
    if (rc != BC_OK) return rc;
    if (!bc_d2h_pair(c, pixels, d_px, np * 4, linear, d_rgba, linear ? np * 16 : 0, c->stream, "copy pixels"))
        return BC_ERR_DEVICE;
    return BC_OK;
}
