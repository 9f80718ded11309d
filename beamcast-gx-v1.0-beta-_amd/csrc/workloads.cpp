// workloads.cpp -- seeded synthetic scenes and ray sets, plus the C entry points of the
// host scene mirror (beamcast_host.h).  Every random value is a pure function of
// (seed, item index, draw index), so any sub-range of a workload can be regenerated.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/beamcast_host.h"
#include "bvh8_build.hpp"
#include "host_scene.hpp"

namespace {

using bcx::V;
using bcx::Vec3;

// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014)
inline uint64_t sm64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// uniform double in [0,1) for (seed, item, draw)
inline double u01(uint64_t seed, uint64_t item, uint64_t draw) {
    uint64_t h = sm64(seed ^ sm64(item * 0x9e3779b97f4a7c15ull + sm64(draw + 0x632be59bd9b4e019ull)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
inline Vec3 sphere_dir(double u, double v) {
    double z = 1 - 2 * u, phi = 2 * M_PI * v, r = std::sqrt(std::max(0.0, 1 - z * z));
    return V(r * std::cos(phi), r * std::sin(phi), z);
}

void add_quad(bcx::Scene& s, Vec3 a, Vec3 b, Vec3 c, Vec3 d, int mat) {
    s.AddTriangle(a, b, c, mat);
    s.AddTriangle(a, c, d, mat);
}

// axis-aligned box of size (sx,sy,sz) with base centre `base`, rotated by deg about +y
void add_box(bcx::Scene& s, Vec3 base, double sx, double sy, double sz, double deg, int mat) {
    double r = deg * M_PI / 180, c = std::cos(r), sn = std::sin(r);
    auto P = [&](double x, double y, double z) {
        double rx = c * x + sn * z, rz = -sn * x + c * z;
        return V(base.x + rx, base.y + y, base.z + rz);
    };
    double hx = sx / 2, hz = sz / 2;
    Vec3 v[8] = {P(-hx, 0, -hz), P(hx, 0, -hz), P(hx, sy, -hz), P(-hx, sy, -hz),
                 P(-hx, 0, hz),  P(hx, 0, hz),  P(hx, sy, hz),  P(-hx, sy, hz)};
    const int f[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};
    for (auto& q : f) add_quad(s, v[q[0]], v[q[1]], v[q[2]], v[q[3]], mat);
}

int mat_diffuse(bcx::Scene& s, Vec3 c) {
    bcx::Material m;
    m.Kind = bcx::MatDiffuse;
    m.Color = c;
    return s.AddMaterial(m);
}

void set_cam(double* out, Vec3 p, Vec3 t, Vec3 u, double fov) {
    if (!out) return;
    double v[10] = {p.x, p.y, p.z, t.x, t.y, t.z, u.x, u.y, u.z, fov};
    std::memcpy(out, v, sizeof v);
}

// closed box: walls at x=+-1, y=0, y=2, z=-2 (open towards the camera at z=0)
void closed_box(bcx::Scene& s, int subdiv, int nmat_extra, uint64_t seed) {
    int grey = mat_diffuse(s, V(.75, .75, .75));
    int red = mat_diffuse(s, V(.63, .065, .05));
    int green = mat_diffuse(s, V(.14, .45, .09));
    bcx::Material lm;
    lm.Kind = bcx::MatEmissive;
    lm.Emission = V(17, 12, 4);
    int light = s.AddMaterial(lm);
    std::vector<int> extra;
    for (int i = 0; i < nmat_extra; i++)
        extra.push_back(mat_diffuse(s, V(.2 + .6 * u01(seed, (uint64_t)i, 0), .2 + .6 * u01(seed, (uint64_t)i, 1),
                                         .2 + .6 * u01(seed, (uint64_t)i, 2))));
    int quad_id = 0;
    auto wall = [&](Vec3 a, Vec3 b, Vec3 c, Vec3 d, int mat) {
        // subdivide the quad a-b-c-d into subdiv x subdiv cells
        for (int i = 0; i < subdiv; i++)
            for (int j = 0; j < subdiv; j++) {
                double u0 = (double)i / subdiv, u1 = (double)(i + 1) / subdiv;
                double v0 = (double)j / subdiv, v1 = (double)(j + 1) / subdiv;
                auto L = [&](double u, double v) {
                    Vec3 ab = a + (b - a) * u, dc = d + (c - d) * u;
                    return ab + (dc - ab) * v;
                };
                int m = mat;
                if (!extra.empty() && mat != light) m = extra[(size_t)(sm64(seed ^ (uint64_t)quad_id) % extra.size())];
                quad_id++;
                add_quad(s, L(u0, v0), L(u1, v0), L(u1, v1), L(u0, v1), m);
            }
    };
    wall(V(-1, 0, 0), V(1, 0, 0), V(1, 0, -2), V(-1, 0, -2), grey);      // floor
    wall(V(-1, 2, -2), V(1, 2, -2), V(1, 2, 0), V(-1, 2, 0), grey);      // ceiling
    wall(V(-1, 0, -2), V(1, 0, -2), V(1, 2, -2), V(-1, 2, -2), grey);    // back
    wall(V(-1, 0, 0), V(-1, 0, -2), V(-1, 2, -2), V(-1, 2, 0), red);     // left
    wall(V(1, 0, -2), V(1, 0, 0), V(1, 2, 0), V(1, 2, -2), green);       // right
    add_box(s, V(-.35, 0, -1.3), .6, 1.2, .6, 17, grey);
    add_box(s, V(.35, 0, -.7), .6, .6, .6, -17, grey);
    add_quad(s, V(-.25, 1.98, -1.25), V(.25, 1.98, -1.25), V(.25, 1.98, -.75), V(-.25, 1.98, -.75), light);
}

// value noise on an integer lattice, smooth-step interpolated
double vnoise(uint64_t seed, double x, double y) {
    double fx = std::floor(x), fy = std::floor(y);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    double tx = x - fx, ty = y - fy;
    auto L = [&](int64_t a, int64_t b) { return u01(seed, (uint64_t)(a * 73856093) ^ (uint64_t)(b * 19349663), 7); };
    double sx = tx * tx * (3 - 2 * tx), sy = ty * ty * (3 - 2 * ty);
    double a = L(ix, iy), b = L(ix + 1, iy), c = L(ix, iy + 1), d = L(ix + 1, iy + 1);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
}

}  // namespace

extern "C" {

bch_scene* bch_scene_new(void) { return new bch_scene(); }
void bch_scene_free(bch_scene* s) {
    if (s) bcx::renderHIPForget(&s->s);  // the persistent runtime must not mistake a later scene for it
    delete s;
}
int bch_add_material(bch_scene* s, int kind, const double* d) {
    bcx::Material m;
    m.Kind = kind;
    m.Color = V(d[0], d[1], d[2]);
    m.Roughness = d[3];
    m.IOR = d[4];
    m.Emission = V(d[5], d[6], d[7]);
    return s->s.AddMaterial(m);
}
void bch_add_sphere(bch_scene* s, const double* c, double r, int mat) { s->s.AddSphere(V(c[0], c[1], c[2]), r, mat); }
void bch_add_triangle(bch_scene* s, const double* a, const double* b, const double* c, int mat) {
    s->s.AddTriangle(V(a[0], a[1], a[2]), V(b[0], b[1], b[2]), V(c[0], c[1], c[2]), mat);
}
void bch_build(bch_scene* s, int bins, int leaf) {
    s->s.BVHBins = bins;
    s->s.BVHLeafSize = leaf;
    s->s.Build();
}
int bch_flatten_check(bch_scene* s, int64_t* out4) {
    const bc_scene_view& v = s->s.View(false);
    bcx::Bvh8 b;
    std::string err;
    if (!bcx::build_bvh8(v, b, err)) return -1;
    std::vector<int> seen((size_t)v.nprim, 0);
    int64_t slots = 0;
    bool ok = true;
    for (size_t i = 0; i < b.nodes.size(); i++) {
        const uint32_t lmask = bc_node_lmask(b.nodes[i]);
        if (lmask & bc_node_imask(b.nodes[i])) ok = false;  // a slot is either internal or a leaf
        for (int k = 0; k < 8; k++) {
            const uint32_t id = b.ids[BC_REC_PER_NODE * i + (size_t)k];
            if (!((lmask >> k) & 1u)) {
                if (id != BC_REC_HOLE) ok = false;
                continue;
            }
            const uint32_t q = id & ~BC_SPHERE_BIT;
            if (id != BC_REC_HOLE && q < (uint32_t)v.nprim) seen[q]++; else ok = false;
            slots++;
        }
    }
    for (int x : seen) ok = ok && x == 1;
    // conservative boxes: every record's exact box (fp64 of the fp32 operands) lies inside the
    // quantised slot box of its leaf and of every internal slot on its path from the root, and
    // every grid origin is k 2^e with |k| < 2^20
    const size_t nn = b.nodes.size();
    std::vector<uint32_t> parent(nn, 0xFFFFFFFFu), pslot(nn, 0);
    for (size_t i = 0; i < nn && ok; i++) {
        const bc_node8& n = b.nodes[i];
        const uint32_t im = bc_node_imask(n);
        for (int a = 0; a < 3; a++) ok = ok && std::abs(bc_node_k(n, a)) < (1 << 20);
        for (int k = 0, r = 0; k < 8; k++) {
            if (!((im >> k) & 1u)) continue;
            const size_t c = (size_t)bc_node_child_base(n) + (size_t)r++;
            if (c >= nn || parent[c] != 0xFFFFFFFFu) { ok = false; break; }
            parent[c] = (uint32_t)i;
            pslot[c] = (uint32_t)k;
        }
    }
    auto inside = [&](size_t node, int slot, const double* lo, const double* hi) {
        const bc_node8& n = b.nodes[node];
        const uint8_t* ql[3] = {n.qlox, n.qloy, n.qloz};
        const uint8_t* qh[3] = {n.qhix, n.qhiy, n.qhiz};
        for (int a = 0; a < 3; a++) {
            const double step = std::ldexp(1.0, bc_node_e(n, a)), o = (double)bc_node_k(n, a) * step;
            if (o + ql[a][slot] * step > lo[a] || o + qh[a][slot] * step < hi[a]) return false;
        }
        return true;
    };
    for (size_t i = 0; i < nn && ok; i++) {
        const uint32_t lm = bc_node_lmask(b.nodes[i]);
        for (int k = 0; k < 8 && ok; k++) {
            if (!((lm >> k) & 1u)) continue;
            const size_t r = BC_REC_PER_NODE * i + (size_t)k;
            const float* g = b.prims[r].v;
            double lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                if (b.ids[r] & BC_SPHERE_BIT) {
                    lo[a] = (double)g[a] - (double)g[3];
                    hi[a] = (double)g[a] + (double)g[3];
                } else {
                    const double A = g[a], B = A + (double)g[3 + a], C = A + (double)g[6 + a];
                    lo[a] = std::min(A, std::min(B, C));
                    hi[a] = std::max(A, std::max(B, C));
                }
            }
            ok = inside(i, k, lo, hi);
            for (size_t c = i; ok && parent[c] != 0xFFFFFFFFu; c = parent[c]) ok = inside(parent[c], (int)pslot[c], lo, hi);
        }
    }
    out4[0] = (int64_t)b.nodes.size();
    out4[1] = b.depth;
    out4[2] = slots;
    out4[3] = ok ? 1 : 0;
    return 0;
}

void bch_build_median(bch_scene* s, int leaf) {
    s->s.BVHMode = "Median";
    s->s.BVHLeafSize = leaf;
    s->s.Build();
    s->s.BVHMode = "SAH";
}
int bch_build_device(bch_scene* s, bc_ctx* ctx, int bins, int leaf) {
    s->s.BVHBins = bins;
    s->s.BVHLeafSize = leaf;
    std::string err;
    return s->s.BuildDevice(ctx, err);
}
void bch_export_bounds(const bch_scene* s, double* box6, double* centroid3) {
    for (size_t i = 0; i < s->s.Primitives.size(); i++) {
        const bcx::Primitive& p = s->s.Primitives[i];
        const double b[6] = {p.Box.Min.x, p.Box.Min.y, p.Box.Min.z, p.Box.Max.x, p.Box.Max.y, p.Box.Max.z};
        std::memcpy(box6 + 6 * i, b, sizeof b);
        const double c[3] = {p.Centroid.x, p.Centroid.y, p.Centroid.z};
        std::memcpy(centroid3 + 3 * i, c, sizeof c);
    }
}
void bch_refit(bch_scene* s) { s->s.Refit(); }
void bch_counts(const bch_scene* s, int32_t* o) {
    o[0] = (int32_t)s->s.Primitives.size();
    o[1] = (int32_t)s->s.Nodes.size();
    o[2] = (int32_t)s->s.Materials.size();
}
void bch_revisions(const bch_scene* s, uint64_t* o) {
    o[0] = s->s.GPUTopologyRevision;
    o[1] = s->s.GPUGeometryRevision;
}
const bc_scene_view* bch_view(bch_scene* s, int nee) { return &s->s.View(nee != 0); }
void bch_export_prims(const bch_scene* s, int32_t* kind, int32_t* mat, double* g) {
    for (size_t i = 0; i < s->s.Primitives.size(); i++) {
        const bcx::Primitive& p = s->s.Primitives[i];
        kind[i] = p.Kind;
        mat[i] = p.Material;
        double* o = g + 9 * i;
        if (p.Kind == bcx::PrimSphere) {
            double v[9] = {p.Center.x, p.Center.y, p.Center.z, p.Radius, 0, 0, 0, 0, 0};
            std::memcpy(o, v, sizeof v);
        } else {
            double v[9] = {p.A.x, p.A.y, p.A.z, p.B.x, p.B.y, p.B.z, p.C.x, p.C.y, p.C.z};
            std::memcpy(o, v, sizeof v);
        }
    }
}
void bch_export_materials(const bch_scene* s, int32_t* kind, double* d) {
    for (size_t i = 0; i < s->s.Materials.size(); i++) {
        const bcx::Material& m = s->s.Materials[i];
        kind[i] = m.Kind;
        double v[8] = {m.Color.x, m.Color.y, m.Color.z, m.Roughness, m.IOR, m.Emission.x, m.Emission.y, m.Emission.z};
        std::memcpy(d + 8 * i, v, sizeof v);
    }
}
void bch_translate(bch_scene* s, int start, int count, const double* d) {  // translatePrimitive, renderer.go:282-303
    Vec3 dv = V(d[0], d[1], d[2]);
    int end = std::min((int)s->s.Primitives.size(), start + count);
    for (int i = std::max(0, start); i < end; i++) {
        bcx::Primitive& p = s->s.Primitives[(size_t)i];
        if (p.Kind == bcx::PrimSphere) {
            p.Center = p.Center + dv;
            p.Centroid = p.Center;
            Vec3 rv = V(p.Radius, p.Radius, p.Radius);
            p.Box = bcx::AABB{p.Center - rv, p.Center + rv};
            continue;
        }
        p.A = p.A + dv; p.B = p.B + dv; p.C = p.C + dv;
        p.E1 = p.B - p.A; p.E2 = p.C - p.A;
        p.Normal = bcx::Unit(bcx::Cross(p.E1, p.E2));
        p.Centroid = (p.A + p.B + p.C) * (1.0 / 3.0);
        Vec3 mn = V(std::min(p.A.x, std::min(p.B.x, p.C.x)), std::min(p.A.y, std::min(p.B.y, p.C.y)),
                    std::min(p.A.z, std::min(p.B.z, p.C.z)));
        Vec3 mx = V(std::max(p.A.x, std::max(p.B.x, p.C.x)), std::max(p.A.y, std::max(p.B.y, p.C.y)),
                    std::max(p.A.z, std::max(p.B.z, p.C.z)));
        p.Box = bcx::AABB{mn - V(1e-6, 1e-6, 1e-6), mx + V(1e-6, 1e-6, 1e-6)};
    }
}

int bch_quality(const char* name, const char* backend, int32_t* o, double* firefly, char* label, int cap) {
    bcx::RenderSettings q;
    if (backend) {
        const std::string b = backend;
        if (b != "Auto" && b != "CPU" && b != "OpenCL GPU") return -1;
        q = bcx::UltraRealismSettings(b);
    } else {
        q = bcx::Quality(name ? name : "");
    }
    const int sampler = q.Sampler == "Halton" ? 1 : (q.Sampler == "R2" ? 2 : 0);
    const int be = q.Backend == "CPU" ? 1 : (q.Backend == "OpenCL GPU" ? 2 : 0);
    const int32_t v[20] = {q.Width, q.Height, q.SPP, q.Bounces, q.Denoise, q.MIS, q.PowerLightSampling, q.AdaptiveSampling,
                           q.ProgressiveUpdates, sampler, q.MSAA, q.TSAA, q.TAA, q.TXAA, q.FXAA, q.HBAO, q.HBAOPlus,
                           q.RRDepth, be, q.BVHBins};
    std::memcpy(o, v, sizeof v);
    if (firefly) *firefly = q.FireflyClamp;
    if (label && cap > 0) std::snprintf(label, (size_t)cap, "%s", q.Quality.c_str());
    return 0;
}

int bch_effective_spp(const char* name, int spp, int msaa, int tsaa) {
    bcx::RenderSettings q = bcx::Quality(name ? name : "");
    q.SPP = spp;
    q.MSAA = msaa;
    q.TSAA = tsaa;
    return bcx::effectiveSPP(q);
}

void bch_camera_basis(const double* pos, const double* target, const double* up, double fov, double aspect, double* o) {
    bcx::Camera c;
    c.Position = V(pos[0], pos[1], pos[2]);
    c.Target = V(target[0], target[1], target[2]);
    c.Up = V(up[0], up[1], up[2]);
    c.FOV = fov;
    Vec3 a, b, cc, d;
    c.basis(aspect, a, b, cc, d);
    double v[12] = {a.x, a.y, a.z, b.x, b.y, b.z, cc.x, cc.y, cc.z, d.x, d.y, d.z};
    std::memcpy(o, v, sizeof v);
}

bch_scene* bch_workload_scene(int kind, uint64_t seed, int64_t param, int bins, int leaf, double* cam) {
    bch_scene* w = new bch_scene();
    bcx::Scene& s = w->s;
    switch (kind) {
    case 1:
        s.Name = "closed box";
        closed_box(s, 1, 0, seed);
        set_cam(cam, V(0, 1, 3.4), V(0, 1, -1), V(0, 1, 0), 40);
        break;
    case 2: {
        s.Name = "triangle soup";
        int m = mat_diffuse(s, V(.75, .75, .75));
        for (int64_t i = 0; i < param; i++) {
            Vec3 c = V(u01(seed, (uint64_t)i, 0), u01(seed, (uint64_t)i, 1), u01(seed, (uint64_t)i, 2));
            Vec3 v[3];
            for (int k = 0; k < 3; k++) {
                double L = .002 + .018 * u01(seed, (uint64_t)i, 3 + 3 * (uint64_t)k);
                v[k] = c + sphere_dir(u01(seed, (uint64_t)i, 4 + 3 * (uint64_t)k), u01(seed, (uint64_t)i, 5 + 3 * (uint64_t)k)) * L;
            }
            s.AddTriangle(v[0], v[1], v[2], m);
        }
        set_cam(cam, V(.5, .5, -1.5), V(.5, .5, .5), V(0, 1, 0), 45);
        break;
    }
    case 3: {  // NewShowcaseScene, renderer.go:2944-2959
        s.Name = "Material Showcase";
        int ground = mat_diffuse(s, V(.55, .57, .6));
        int red = mat_diffuse(s, V(.72, .12, .10));
        bcx::Material g; g.Kind = bcx::MatMetal; g.Color = V(.93, .72, .28); g.Roughness = .08;
        int gold = s.AddMaterial(g);
        bcx::Material gl; gl.Kind = bcx::MatGlass; gl.Color = V(.98, .99, 1); gl.IOR = 1.5;
        int glass = s.AddMaterial(gl);
        bcx::Material lm; lm.Kind = bcx::MatEmissive; lm.Emission = V(18, 15, 11);
        int light = s.AddMaterial(lm);
        s.AddSphere(V(0, -1001, -4), 1000, ground);
        s.AddSphere(V(-1.8, 0, -4.2), 1, red);
        s.AddSphere(V(.25, 0, -3.2), 1, glass);
        s.AddSphere(V(2.2, 0, -4.4), 1, gold);
        s.AddSphere(V(0, 5, -4), .8, light);
        set_cam(cam, V(6, 3.2, 5.5), V(0, 0, -4), V(0, 1, 0), 40);
        break;
    }
    case 4: {  // NewCornellScene, renderer.go:2960-2979
        s.Name = "Cornell-like Room";
        int white = mat_diffuse(s, V(.72, .72, .72));
        int red = mat_diffuse(s, V(.70, .08, .06));
        int green = mat_diffuse(s, V(.08, .55, .12));
        bcx::Material gl; gl.Kind = bcx::MatGlass; gl.Color = V(.98, .99, 1); gl.IOR = 1.5;
        int glass = s.AddMaterial(gl);
        bcx::Material lm; lm.Kind = bcx::MatEmissive; lm.Emission = V(22, 20, 16);
        int light = s.AddMaterial(lm);
        s.AddTriangle(V(-3, -2, -8), V(3, -2, -8), V(3, -2, -2), white);
        s.AddTriangle(V(-3, -2, -8), V(3, -2, -2), V(-3, -2, -2), white);
        s.AddTriangle(V(-3, -2, -8), V(-3, -2, -2), V(-3, 4, -2), red);
        s.AddTriangle(V(-3, -2, -8), V(-3, 4, -2), V(-3, 4, -8), red);
        s.AddTriangle(V(3, -2, -2), V(3, -2, -8), V(3, 4, -8), green);
        s.AddTriangle(V(3, -2, -2), V(3, 4, -8), V(3, 4, -2), green);
        s.AddSphere(V(-.9, -1, -5.2), 1, glass);
        s.AddSphere(V(1.1, -1.2, -4.2), .8, white);
        s.AddSphere(V(0, 3.2, -5), .65, light);
        set_cam(cam, V(0, .5, 5), V(0, 0, -5), V(0, 1, 0), 38);
        break;
    }
    case 5: {  // NewStressScene, renderer.go:2980-2995
        s.Name = "Stress Grid";
        int ground = mat_diffuse(s, V(.45, .48, .52));
        bcx::Material b; b.Kind = bcx::MatMetal; b.Color = V(.25, .52, .92); b.Roughness = .12;
        int blue = s.AddMaterial(b);
        bcx::Material lm; lm.Kind = bcx::MatEmissive; lm.Emission = V(25, 22, 18);
        int light = s.AddMaterial(lm);
        s.AddSphere(V(0, -1001, -10), 1000, ground);
        for (int z = 0; z < 20; z++)
            for (int x = -20; x <= 20; x++) s.AddSphere(V((double)x * .58, -.65, -3 - (double)z * .58), .22, blue);
        s.AddSphere(V(0, 8, -8), 1.2, light);
        set_cam(cam, V(9, 5, 8), V(0, -.5, -8), V(0, 1, 0), 42);
        break;
    }
    case 6: {  // heightfield terrain: param x param cells over [0,1]^2, 8-octave fBm
        s.Name = "terrain";
        int n = (int)std::max<int64_t>(1, param);
        std::vector<int> mats;
        for (int i = 0; i < 64; i++) {
            double a = .2 + .6 * u01(seed, (uint64_t)i, 100);
            mats.push_back(mat_diffuse(s, V(a, a, a)));
        }
        std::vector<double> hgt((size_t)(n + 1) * (n + 1));
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                double x = (double)i / n, y = (double)j / n, amp = .5, fr = 4, hsum = 0;
                for (int o = 0; o < 8; o++) {
                    hsum += amp * vnoise(seed + (uint64_t)o, x * fr, y * fr);
                    amp *= .5;
                    fr *= 2;
                }
                hgt[(size_t)j * (n + 1) + i] = .25 * hsum;
            }
        s.Primitives.reserve((size_t)n * n * 2);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                auto P = [&](int a, int b) { return V((double)a / n, hgt[(size_t)b * (n + 1) + a], (double)b / n); };
                int m = mats[(size_t)(sm64(seed ^ ((uint64_t)j * (uint64_t)n + (uint64_t)i)) % 64)];
                s.AddTriangle(P(i, j), P(i, j + 1), P(i + 1, j + 1), m);
                s.AddTriangle(P(i, j), P(i + 1, j + 1), P(i + 1, j), m);
            }
        set_cam(cam, V(.5, .6, -.35), V(.5, .1, .5), V(0, 1, 0), 50);
        break;
    }
    case 8: {  // scene `param` of the multi-scene batch: boxes, spheres and 2-16 quad lights
        s.Name = "batch scene";
        const uint64_t i = (uint64_t)std::max<int64_t>(0, param);
        // primitive budget log-uniform in [1e4, 2e6]
        const double tris = std::exp(std::log(1e4) + u01(seed, i, 0) * std::log(200.0));
        std::vector<int> pal;
        for (int k = 0; k < 6; k++)
            pal.push_back(mat_diffuse(s, V(.15 + .7 * u01(seed, i * 64 + (uint64_t)k, 1), .15 + .7 * u01(seed, i * 64 + (uint64_t)k, 2),
                                           .15 + .7 * u01(seed, i * 64 + (uint64_t)k, 3))));
        bcx::Material metal;
        metal.Kind = bcx::MatMetal;
        metal.Color = V(.85, .8, .7);
        metal.Roughness = .15;
        pal.push_back(s.AddMaterial(metal));
        bcx::Material glass;
        glass.Kind = bcx::MatGlass;
        glass.Color = V(.95, .97, 1);
        glass.IOR = 1.5;
        pal.push_back(s.AddMaterial(glass));
        const int floor_m = mat_diffuse(s, V(.6, .6, .6));
        add_quad(s, V(-30, 0, -30), V(-30, 0, 30), V(30, 0, 30), V(30, 0, -30), floor_m);
        // analytic spheres (0.5% of the budget) and 2-16 emissive quads overhead; instanced boxes
        // of 12 triangles fill the rest of the budget, so the scene holds between the budget and
        // 11 primitives more
        const int64_t nsph = std::max<int64_t>(4, (int64_t)(tris / 200));
        const int nl = 2 + (int)(u01(seed, i, 14) * 15);
        const int64_t nbox = std::max<int64_t>(1, (int64_t)std::ceil((tris - 2.0 - (double)nsph - 2.0 * nl) / 12.0));
        const double lot = 10;
        for (int64_t b = 0; b < nbox; b++) {
            const uint64_t it = i * 0x100000000ull + (uint64_t)b;
            const double x = -lot + 2 * lot * u01(seed, it, 4), z = -lot + 2 * lot * u01(seed, it, 5);
            const double w = .05 + .5 * u01(seed, it, 6), d = .05 + .5 * u01(seed, it, 7), h = .05 + 1.5 * u01(seed, it, 8);
            const int m = pal[(size_t)(sm64(seed ^ it) % pal.size())];
            add_box(s, V(x, 0, z), w, h, d, 360 * u01(seed, it, 9), m);
        }
        for (int64_t b = 0; b < nsph; b++) {
            const uint64_t it = i * 0x100000000ull + 0x80000000ull + (uint64_t)b;
            const double r = .1 + .4 * u01(seed, it, 10);
            const double x = -lot + 2 * lot * u01(seed, it, 11), z = -lot + 2 * lot * u01(seed, it, 12);
            s.AddSphere(V(x, r + 2 * u01(seed, it, 13), z), r, pal[(size_t)(sm64(seed ^ it ^ 0x5555) % pal.size())]);
        }
        for (int l = 0; l < nl; l++) {
            const uint64_t it = i * 64 + (uint64_t)l;
            bcx::Material lm;
            lm.Kind = bcx::MatEmissive;
            const double warm = u01(seed, it, 15);
            const double e = 6 + 10 * u01(seed, it, 16);
            lm.Emission = V(e, e * (.75 + .2 * warm), e * (.55 + .45 * (1 - warm)));
            const int lmat = s.AddMaterial(lm);
            const double x = -8 + 16 * u01(seed, it, 17), z = -8 + 16 * u01(seed, it, 18), y = 6 + 3 * u01(seed, it, 19);
            const double hs = .3 + .6 * u01(seed, it, 20);
            add_quad(s, V(x - hs, y, z - hs), V(x + hs, y, z - hs), V(x + hs, y, z + hs), V(x - hs, y, z + hs), lmat);
        }
        const double az = 2 * M_PI * u01(seed, i, 21), dist = 14 + 6 * u01(seed, i, 22);
        set_cam(cam, V(dist * std::cos(az), 4 + 4 * u01(seed, i, 23), dist * std::sin(az)), V(0, .5, 0), V(0, 1, 0), 45);
        break;
    }
    case 7: {  // closed box retessellated to about `param` triangles with 64 materials
        s.Name = "closed box (retessellated)";
        int sub = (int)std::max<double>(1.0, std::floor(std::sqrt((double)param / 10.0)));
        closed_box(s, sub, 64, seed);
        set_cam(cam, V(0, 1, 3.4), V(0, 1, -1), V(0, 1, 0), 40);
        break;
    }
    default:
        delete w;
        return nullptr;
    }
    s.BVHBins = bins;
    s.BVHLeafSize = leaf;
    s.Build();
    return w;
}

int bch_workload_rays(int set, uint64_t seed, int64_t width, int any_hit, int64_t first, int64_t count, float* out) {
    if (width <= 0 || first < 0 || count < 0) return -1;
    Vec3 origin, lower, horiz, vert;
    if (set == 0) {
        bcx::Camera c;
        c.Position = V(.5, .5, -1.5);
        c.Target = V(.5, .5, .5);
        c.Up = V(0, 1, 0);
        c.FOV = 45;
        c.basis(1.0, origin, lower, horiz, vert);
    }
    const uint64_t rseed = seed ^ (set == 0 ? 0x0c0e0e0aull : 0x1c0e0e0bull);
    for (int64_t k = 0; k < count; k++) {
        const uint64_t i = (uint64_t)(first + k);
        Vec3 o, d;
        double tmax = 1e30;
        if (set == 0) {
            int64_t x = (int64_t)(i % (uint64_t)width), y = (int64_t)(i / (uint64_t)width);
            double u = ((double)x + .5) / (double)width, v = ((double)(width - 1 - y) + .5) / (double)width;
            o = origin;
            d = bcx::Unit(lower + horiz * u + vert * v - origin);
            if (any_hit) {  // tmax = entry distance into the unit cube + U(.05, 1)
                double tn = 0, tf = 1e30;
                const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
                for (int a = 0; a < 3; a++) {
                    if (std::fabs(dd[a]) < 1e-15) continue;
                    double t0 = (0 - oo[a]) / dd[a], t1 = (1 - oo[a]) / dd[a];
                    if (t0 > t1) std::swap(t0, t1);
                    tn = std::max(tn, t0);
                    tf = std::min(tf, t1);
                }
                tmax = (tn <= tf ? tn : 0) + .05 + .95 * u01(rseed, i, 9);
            }
        } else {
            o = V(u01(rseed, i, 0), u01(rseed, i, 1), u01(rseed, i, 2));
            d = sphere_dir(u01(rseed, i, 3), u01(rseed, i, 4));
            if (any_hit) tmax = .05 + .95 * u01(rseed, i, 9);
        }
        float* r = out + 8 * k;
        r[0] = (float)o.x; r[1] = (float)o.y; r[2] = (float)o.z; r[3] = 1e-4f;
        r[4] = (float)d.x; r[5] = (float)d.y; r[6] = (float)d.z; r[7] = (float)tmax;
    }
    return 0;
}

}  // extern "C"
