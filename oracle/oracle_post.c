/*
 * oracle_post.c -- CPU restatement of the renderer's post stack on the GPU path's outputs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the device post stack.
 *
 * Restates, in fp64 as the Go code computes them (reference = windows_portable/):
 *   - decodeOctNormal                                     gpu_opencl_windows.go:552-562
 *   - compact guide -> Guide{N, Depth, Valid}            gpu_opencl_windows.go:686-716
 *   - luminance                                           renderer.go:2142-2144
 *   - atrousDenoise                                       renderer.go:2228-2287
 *   - edgeAwareSharpen                                    renderer.go:2146-2183
 *   - applyFXAA / applyTXAAResolve                        renderer.go:2557-2595
 *   - toneMapHDR / toRGBA                                 renderer.go:2033-2064
 * in the order RenderState.Start applies them when HBAO, temporal reuse, bloom, debug views
 * and upscaling are off (renderer.go:2842-2895): denoise, FXAA, TXAA, toRGBA.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }
static double lum3(const double* c) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }

void orc_decode_oct_normal(uint32_t qx, uint32_t qy, double* n) {
    double x = (double)qx / 1023.0 * 2 - 1;
    double y = (double)qy / 1023.0 * 2 - 1;
    double z = 1 - fabs(x) - fabs(y);
    if (z < 0) {
        double ox = x, oy = y;
        x = (1 - fabs(oy)) * copysign(1, ox);
        y = (1 - fabs(ox)) * copysign(1, oy);
    }
    double l = sqrt(x * x + y * y + z * z);
    if (l < 1e-12) { n[0] = n[1] = n[2] = 0; return; }
    n[0] = x * (1 / l); n[1] = y * (1 / l); n[2] = z * (1 / l);
}

typedef struct { double n[3], depth; int valid; } guide_t;

static void decode_guides(const uint32_t* g, int64_t np, guide_t* out) {
    for (int64_t i = 0; i < np; i++) {
        uint32_t packed = g[3 * i + 1];
        memset(&out[i], 0, sizeof out[i]);
        if (!(packed & 0x80000000u)) continue;
        union { uint32_t u; float f; } cv;
        cv.u = g[3 * i];
        out[i].depth = (double)cv.f;
        orc_decode_oct_normal(packed & 1023u, (packed >> 10) & 1023u, out[i].n);
        out[i].valid = 1;
    }
}

static void atrous(const double* cur, double* next, const guide_t* gd, int w, int h, int step) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t i = (int64_t)y * w + x;
            const double* center = cur + 3 * i;
            const guide_t* g0 = &gd[i];
            double l0 = lum3(center), sum[3] = {0, 0, 0}, wsum = 0;
            for (int oy = -1; oy <= 1; oy++) {
                int yy = y + oy * step;
                if (yy < 0 || yy >= h) continue;
                for (int ox = -1; ox <= 1; ox++) {
                    int xx = x + ox * step;
                    if (xx < 0 || xx >= w) continue;
                    int64_t j = (int64_t)yy * w + xx;
                    const guide_t* g1 = &gd[j];
                    if (g0->valid != g1->valid) continue;
                    const double* q = cur + 3 * j;
                    double wgt = 1.0;
                    if (g0->valid) {
                        double nd = clampd(g0->n[0] * g1->n[0] + g0->n[1] * g1->n[1] + g0->n[2] * g1->n[2], -1, 1);
                        wgt *= exp((nd - 1) * 48.0);
                        double ds = fmax(0.15, 0.04 * fmax(g0->depth, 1.0));
                        wgt *= exp(-fabs(g0->depth - g1->depth) / ds);
                    }
                    wgt *= exp(-fabs(l0 - lum3(q)) * 3.5);
                    if (ox == 0 && oy == 0) wgt *= 1.5;
                    for (int k = 0; k < 3; k++) sum[k] += q[k] * wgt;
                    wsum += wgt;
                }
            }
            for (int k = 0; k < 3; k++) next[3 * i + k] = wsum > 1e-12 ? sum[k] * (1 / wsum) : center[k];
        }
}

static void sharpen(const double* img, double* out, int w, int h, double amount) {
    if (amount <= 0) { memcpy(out, img, sizeof(double) * 3 * (size_t)w * (size_t)h); return; }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const double* center = img + 3 * ((int64_t)y * w + x);
            double cl = lum3(center), sum[3] = {0, 0, 0}, count = 0, ew = 0;
            for (int oy = -1; oy <= 1; oy++) {
                int yy = y + oy;
                if (yy < 0 || yy >= h) continue;
                for (int ox = -1; ox <= 1; ox++) {
                    int xx = x + ox;
                    if (xx < 0 || xx >= w) continue;
                    const double* nb = img + 3 * ((int64_t)yy * w + xx);
                    for (int k = 0; k < 3; k++) sum[k] += nb[k];
                    count++;
                    ew += fabs(cl - lum3(nb));
                }
            }
            double local = amount * clampd(ew / 2.0, 0.35, 1.0);
            double* o = out + 3 * ((int64_t)y * w + x);
            for (int k = 0; k < 3; k++) {
                double blur = sum[k] * (1 / fmax(count, 1));
                o[k] = fmax(0, center[k] + (center[k] - blur) * local);
            }
        }
}

static void fxaa(const double* img, double* out, int w, int h, int level) {
    memcpy(out, img, sizeof(double) * 3 * (size_t)w * (size_t)h);
    if (level <= 0 || w < 3 || h < 3) return;
    static const double thr[4] = {0, 0.16, 0.10, 0.065}, bl[4] = {0, 0.35, 0.50, 0.65};
    int lv = level < 3 ? level : 3;
    double threshold = thr[lv], blend = bl[lv];
    for (int y = 1; y < h - 1; y++)
        for (int x = 1; x < w - 1; x++) {
            int64_t i = (int64_t)y * w + x;
            const double *c = img + 3 * i, *pn = img + 3 * (i - w), *ps = img + 3 * (i + w), *pw = img + 3 * (i - 1),
                         *pe = img + 3 * (i + 1);
            double ln = lum3(pn), ls = lum3(ps), lw = lum3(pw), le = lum3(pe), lc = lum3(c);
            double lo = fmin(lc, fmin(fmin(ln, ls), fmin(lw, le)));
            double hi = fmax(lc, fmax(fmax(ln, ls), fmax(lw, le)));
            if (hi - lo < threshold * fmax(1.0, hi)) continue;
            for (int k = 0; k < 3; k++) {
                double avg = (((pn[k] + ps[k]) + pw[k]) + pe[k]) * 0.25;
                out[3 * i + k] = c[k] * (1 - blend) + avg * blend;
            }
        }
}

static double tone(double x, int hdr) {
    switch (hdr) {
    case 0: return clampd(x, 0, 1);
    case 1: return x / (1 + x);
    case 3: {
        double a = 2.51, b = 0.03, cc = 2.43, d = 0.59, e = 0.14;
        return clampd((x * (a * x + b)) / (x * (cc * x + d) + e), 0, 1);
    }
    default:
        x = fmax(0, x - 0.004);
        return clampd((x * (6.2 * x + 0.5)) / (x * (6.2 * x + 1.7) + 0.06), 0, 1);
    }
}

uint32_t orc_to_rgba(const double* c, double exposure, double gamma, int hdr) {
    double es = pow(2, exposure), ig = 1.0 / fmax(0.1, gamma), v[3];
    for (int k = 0; k < 3; k++) v[k] = pow(fmax(0, tone(c[k] * es, hdr)), ig);
    uint32_t r = (uint32_t)(clampd(v[0], 0, .999) * 256);
    uint32_t g = (uint32_t)(clampd(v[1], 0, .999) * 256);
    uint32_t b = (uint32_t)(clampd(v[2], 0, .999) * 256);
    return 0xff000000u | (b << 16) | (g << 8) | r;
}

int orc_post(int w, int h, const float* rgba, const uint32_t* guides, int denoise_passes, int fxaa_level,
             int txaa_level, double exposure, double gamma, int hdr, double* linear_out, uint32_t* pixels_out) {
    int64_t np = (int64_t)w * h;
    double* a = (double*)malloc(sizeof(double) * 3 * (size_t)np);
    double* b = (double*)malloc(sizeof(double) * 3 * (size_t)np);
    guide_t* gd = (guide_t*)calloc((size_t)np, sizeof(guide_t));
    if (!a || !b || !gd) { free(a); free(b); free(gd); return -1; }
    for (int64_t i = 0; i < np; i++)
        for (int k = 0; k < 3; k++) a[3 * i + k] = (double)rgba[4 * i + k];
    if (denoise_passes > 0 && guides) {
        decode_guides(guides, np, gd);
        for (int it = 0; it < denoise_passes; it++) {
            atrous(a, b, gd, w, h, 1 << it);
            double* t = a; a = b; b = t;
        }
    }
    if (fxaa_level > 0) {
        fxaa(a, b, w, h, fxaa_level);
        double* t = a; a = b; b = t;
    }
    if (txaa_level > 0) {  /* applyTXAAResolve: FXAA one level up, then a light sharpen */
        fxaa(a, b, w, h, txaa_level + 1 < 3 ? txaa_level + 1 : 3);
        sharpen(b, a, w, h, 0.05 + 0.04 * (double)txaa_level);
    }
    for (int64_t i = 0; i < np; i++) {
        if (linear_out) for (int k = 0; k < 3; k++) linear_out[3 * i + k] = a[3 * i + k];
        if (pixels_out) pixels_out[i] = orc_to_rgba(a + 3 * i, exposure, gamma, hdr);
    }
    free(a); free(b); free(gd);
    return 0;
}
