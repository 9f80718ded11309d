// host_scene.cpp -- C++ mirror of the Go scene build and GPU render entry.
// Each function names the renderer.go / gpu_opencl_windows.go code whose behaviour it keeps.
#include "host_scene.hpp"
#include "../../include/beamcast_host.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <mutex>

namespace bcx {

Vec3 Unit(Vec3 a) {  // renderer.go:32-38
    double l = std::sqrt(Dot(a, a));
    if (l < 1e-12) return V(0, 0, 0);
    return a * (1 / l);
}

namespace {
AABB emptyAABB() {
    const double inf = INFINITY;
    return AABB{V(inf, inf, inf), V(-inf, -inf, -inf)};
}
AABB expandPoint(AABB b, Vec3 p) {  // renderer.go:117-125
    b.Min = V(GoMin(b.Min.x, p.x), GoMin(b.Min.y, p.y), GoMin(b.Min.z, p.z));
    b.Max = V(GoMax(b.Max.x, p.x), GoMax(b.Max.y, p.y), GoMax(b.Max.z, p.z));
    return b;
}
AABB expandBox(AABB b, const AABB& c) { return expandPoint(expandPoint(b, c.Min), c.Max); }
double surfaceArea(const AABB& b) {  // renderer.go:129-135
    Vec3 e = b.Max - b.Min;
    if (e.x < 0 || e.y < 0 || e.z < 0) return 0;
    return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
}
double axisOf(const Vec3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
int largestAxis(const Vec3& e) {  // renderer.go:486-494
    if (e.y > e.x && e.y >= e.z) return 1;
    if (e.z > e.x && e.z > e.y) return 2;
    return 0;
}
}  // namespace

int Scene::AddMaterial(const Material& m) {
    Materials.push_back(m);
    return (int)Materials.size() - 1;
}

void Scene::AddSphere(Vec3 c, double r, int mat) {  // renderer.go:246-256
    if (r <= 0) return;
    Primitive p;
    p.Kind = PrimSphere;
    p.Material = mat;
    p.Center = c;
    p.Radius = r;
    p.Centroid = c;
    Vec3 rv = V(r, r, r);
    p.Box = AABB{c - rv, c + rv};
    Primitives.push_back(p);
}

void Scene::AddTriangle(Vec3 a, Vec3 b, Vec3 c, int mat) {  // renderer.go:257-270
    Vec3 e1 = b - a, e2 = c - a, n = Cross(e1, e2);
    if (Dot(n, n) < 1e-18) return;
    AABB box = expandPoint(expandPoint(expandPoint(emptyAABB(), a), b), c);
    const double eps = 1e-6;
    box.Min = box.Min - V(eps, eps, eps);
    box.Max = box.Max + V(eps, eps, eps);
    Primitive p;
    p.Kind = PrimTriangle;
    p.Material = mat;
    p.A = a; p.B = b; p.C = c; p.E1 = e1; p.E2 = e2;
    p.Normal = Unit(n);
    p.Box = box;
    p.Centroid = (a + b + c) * (1.0 / 3.0);
    Primitives.push_back(p);
}

void Scene::Build() {  // renderer.go:409-449
    GPUTopologyRevision++;
    GPUGeometryRevision++;
    if (BVHMode.empty()) BVHMode = "SAH";
    if (BVHBins < 4) BVHBins = 16;
    if (BVHBins > 32) BVHBins = 32;
    if (BVHLeafSize < 1) BVHLeafSize = 4;
    if (BVHLeafSize > 16) BVHLeafSize = 16;
    Nodes.clear();
    Nodes.reserve(Primitives.size() * 2);
    Order.resize(Primitives.size());
    for (size_t i = 0; i < Order.size(); i++) Order[i] = (int)i;
    if (Order.empty()) return;
    std::string mode = BVHMode;
    for (char& ch : mode) ch = (char)std::tolower((unsigned char)ch);
    if (mode == "median") buildNodeMedian(0, (int)Order.size());
    else buildNodeSAH(0, (int)Order.size());
}

// buildNodeMedian (renderer.go:506-524): leaf at count <= BVHLeafSize, else sort the range by the
// centroid on the largest centroid-extent axis and split at the middle.  Go sorts with the
// unstable sort.Slice, so the order among equal centroids is not pinned; std::sort is unstable
// too, and any order of equals gives a valid tree of the same shape.
int Scene::buildNodeMedian(int start, int end) {
    const int idx = (int)Nodes.size();
    Nodes.emplace_back();
    AABB box = emptyAABB(), cb = emptyAABB();
    for (int i = start; i < end; i++) {
        const Primitive& p = Primitives[(size_t)Order[(size_t)i]];
        box = expandBox(box, p.Box);
        cb = expandPoint(cb, p.Centroid);
    }
    const int count = end - start;
    if (count <= BVHLeafSize) {
        BVHNode& n = Nodes[(size_t)idx];
        n.Box = box; n.Left = -1; n.Right = -1; n.Start = start; n.Count = count;
        return idx;
    }
    const int axis = largestAxis(cb.Max - cb.Min);
    std::sort(Order.begin() + start, Order.begin() + end, [&](int a, int b) {
        return axisOf(Primitives[(size_t)a].Centroid, axis) < axisOf(Primitives[(size_t)b].Centroid, axis);
    });
    const int mid = (start + end) / 2;
    const int left = buildNodeMedian(start, mid);
    const int right = buildNodeMedian(mid, end);
    BVHNode& n = Nodes[(size_t)idx];
    n.Box = box; n.Left = left; n.Right = right; n.Start = 0; n.Count = 0;
    return idx;
}

int Scene::BuildDevice(bc_ctx* ctx, std::string& err) {  // Build (renderer.go:409-449), tree from the GPU
    if (BVHBins < 4) BVHBins = 16;
    if (BVHBins > 32) BVHBins = 32;
    if (BVHLeafSize < 1) BVHLeafSize = 4;
    if (BVHLeafSize > 16) BVHLeafSize = 16;
    const size_t n = Primitives.size();
    if (n == 0) {
        Nodes.clear();
        Order.clear();
        GPUTopologyRevision++;
        GPUGeometryRevision++;
        return BC_OK;
    }
    std::vector<double> box(6 * n), cen(3 * n), nbox(12 * n);
    std::vector<int32_t> link(8 * n), order(n);
    for (size_t i = 0; i < n; i++) {
        const Primitive& p = Primitives[i];
        const double b[6] = {p.Box.Min.x, p.Box.Min.y, p.Box.Min.z, p.Box.Max.x, p.Box.Max.y, p.Box.Max.z};
        std::copy(b, b + 6, box.begin() + 6 * (ptrdiff_t)i);
        cen[3 * i] = p.Centroid.x; cen[3 * i + 1] = p.Centroid.y; cen[3 * i + 2] = p.Centroid.z;
    }
    int32_t nn = 0;
    const int rc = bc_build_bvh(ctx, (int32_t)n, box.data(), cen.data(), BVHBins, BVHLeafSize, (int32_t)(2 * n), nbox.data(),
                                link.data(), order.data(), &nn);
    if (rc != BC_OK) { err = bc_last_error(ctx); return rc; }
    GPUTopologyRevision++;
    GPUGeometryRevision++;
    Nodes.assign((size_t)nn, BVHNode{});
    for (size_t i = 0; i < (size_t)nn; i++) {
        BVHNode& d = Nodes[i];
        const double* b = &nbox[6 * i];
        d.Box = AABB{V(b[0], b[1], b[2]), V(b[3], b[4], b[5])};
        d.Left = link[4 * i]; d.Right = link[4 * i + 1]; d.Start = link[4 * i + 2]; d.Count = link[4 * i + 3];
    }
    Order.assign(order.begin(), order.end());
    return BC_OK;
}

// buildNodeSAH (renderer.go:526-633): one axis (largest centroid extent), binned sweep,
// leaf when count <= leaf, degenerate extent, no valid split, or the 0.98 rule; in-place
// two-pointer partition; children appended depth first (left subtree first).
int Scene::buildNodeSAH(int start, int end) {
    const int idx = (int)Nodes.size();
    Nodes.emplace_back();
    AABB box = emptyAABB(), cb = emptyAABB();
    for (int i = start; i < end; i++) {
        const Primitive& p = Primitives[(size_t)Order[(size_t)i]];
        box = expandBox(box, p.Box);
        cb = expandPoint(cb, p.Centroid);
    }
    const int count = end - start;
    auto makeLeaf = [&]() {
        BVHNode& n = Nodes[(size_t)idx];
        n.Box = box; n.Left = -1; n.Right = -1; n.Start = start; n.Count = count;
        return idx;
    };
    if (count <= BVHLeafSize) return makeLeaf();
    const int axis = largestAxis(cb.Max - cb.Min);
    const double axisExtent = axisOf(cb.Max, axis) - axisOf(cb.Min, axis);
    if (axisExtent < 1e-12) return makeLeaf();
    const int binsN = BVHBins;
    AABB binBox[32];
    int binCount[32];
    for (int i = 0; i < binsN; i++) { binBox[i] = emptyAABB(); binCount[i] = 0; }
    const double minC = axisOf(cb.Min, axis), invExtent = 1 / axisExtent;
    auto binOf = [&](const Primitive& p) {
        int b = (int)((axisOf(p.Centroid, axis) - minC) * invExtent * (double)binsN);
        return std::min(std::max(b, 0), binsN - 1);
    };
    for (int i = start; i < end; i++) {
        const Primitive& p = Primitives[(size_t)Order[(size_t)i]];
        int b = binOf(p);
        binCount[b]++;
        binBox[b] = expandBox(binBox[b], p.Box);
    }
    int leftCount[31], rightCount[31];
    double leftArea[31], rightArea[31];
    AABB acc = emptyAABB();
    int accCount = 0;
    for (int i = 0; i + 1 < binsN; i++) {
        if (binCount[i] > 0) acc = expandBox(acc, binBox[i]);
        accCount += binCount[i];
        leftCount[i] = accCount;
        leftArea[i] = surfaceArea(acc);
    }
    acc = emptyAABB();
    accCount = 0;
    for (int i = binsN - 1; i > 0; i--) {
        if (binCount[i] > 0) acc = expandBox(acc, binBox[i]);
        accCount += binCount[i];
        rightCount[i - 1] = accCount;
        rightArea[i - 1] = surfaceArea(acc);
    }
    int bestSplit = -1;
    double bestCost = INFINITY;
    for (int i = 0; i + 1 < binsN; i++) {
        if (leftCount[i] == 0 || rightCount[i] == 0) continue;
        double cost = leftArea[i] * (double)leftCount[i] + rightArea[i] * (double)rightCount[i];
        if (cost < bestCost) { bestCost = cost; bestSplit = i; }
    }
    const double leafCost = surfaceArea(box) * (double)count;
    if (bestSplit < 0 || (count <= BVHLeafSize * 2 && bestCost >= leafCost * 0.98)) return makeLeaf();
    int i = start, j = end - 1;
    while (i <= j) {
        if (binOf(Primitives[(size_t)Order[(size_t)i]]) <= bestSplit) {
            i++;
        } else {
            std::swap(Order[(size_t)i], Order[(size_t)j]);
            j--;
        }
    }
    int mid = i;
    if (mid <= start || mid >= end) {  // renderer.go:623-628 (not reached: split bins are non-empty)
        std::sort(Order.begin() + start, Order.begin() + end, [&](int a, int b) {
            double ca = axisOf(Primitives[(size_t)a].Centroid, axis), cb2 = axisOf(Primitives[(size_t)b].Centroid, axis);
            return ca < cb2 || (ca == cb2 && a < b);
        });
        mid = (start + end) / 2;
    }
    const int left = buildNodeSAH(start, mid);
    const int right = buildNodeSAH(mid, end);
    BVHNode& n = Nodes[(size_t)idx];
    n.Box = box; n.Left = left; n.Right = right; n.Start = 0; n.Count = 0;
    return idx;
}

void Scene::Refit() {  // renderer.go:451-473
    GPUGeometryRevision++;
    if (Nodes.empty()) { Build(); return; }
    for (int ni = (int)Nodes.size() - 1; ni >= 0; ni--) {
        BVHNode& n = Nodes[(size_t)ni];
        if (n.Count > 0) {
            AABB box = emptyAABB();
            for (int i = 0; i < n.Count; i++) {
                int pi = Order[(size_t)(n.Start + i)];
                if (pi >= 0 && pi < (int)Primitives.size()) box = expandBox(box, Primitives[(size_t)pi].Box);
            }
            n.Box = box;
        } else if (n.Left >= 0 && n.Right >= 0) {
            n.Box = expandBox(Nodes[(size_t)n.Left].Box, Nodes[(size_t)n.Right].Box);
        }
    }
}

const bc_scene_view& Scene::View(bool nee_triangles) {
    const size_t nn = Nodes.size(), np = Primitives.size(), nm = Materials.size();
    node_box_.resize(nn * 6);
    node_link_.resize(nn * 4);
    for (size_t i = 0; i < nn; i++) {
        const BVHNode& n = Nodes[i];
        double* b = &node_box_[6 * i];
        b[0] = n.Box.Min.x; b[1] = n.Box.Min.y; b[2] = n.Box.Min.z;
        b[3] = n.Box.Max.x; b[4] = n.Box.Max.y; b[5] = n.Box.Max.z;
        int32_t* l = &node_link_[4 * i];
        l[0] = n.Left; l[1] = n.Right; l[2] = n.Start; l[3] = n.Count;
    }
    order_.assign(Order.begin(), Order.end());
    prim_kind_.resize(np);
    prim_mat_.resize(np);
    prim_geom_.assign(np * 10, 0.0);
    for (size_t i = 0; i < np; i++) {
        const Primitive& p = Primitives[i];
        prim_kind_[i] = p.Kind;
        prim_mat_[i] = p.Material;
        double* g = &prim_geom_[10 * i];
        if (p.Kind == PrimSphere) {
            g[0] = p.Center.x; g[1] = p.Center.y; g[2] = p.Center.z; g[3] = p.Radius;
        } else {
            g[0] = p.A.x; g[1] = p.A.y; g[2] = p.A.z;
            g[3] = p.E1.x; g[4] = p.E1.y; g[5] = p.E1.z;
            g[6] = p.E2.x; g[7] = p.E2.y; g[8] = p.E2.z;
        }
    }
    mat_kind_.resize(nm);
    mat_data_.resize(nm * 8);
    for (size_t i = 0; i < nm; i++) {
        const Material& m = Materials[i];
        mat_kind_[i] = m.Kind;
        double* d = &mat_data_[8 * i];
        d[0] = m.Color.x; d[1] = m.Color.y; d[2] = m.Color.z; d[3] = m.Roughness; d[4] = m.IOR;
        d[5] = m.Emission.x; d[6] = m.Emission.y; d[7] = m.Emission.z;
    }
    view_.nnode = (int32_t)nn;
    view_.node_box = node_box_.data();
    view_.node_link = node_link_.data();
    view_.nprim = (int32_t)np;
    view_.order = order_.data();
    view_.prim_kind = prim_kind_.data();
    view_.prim_mat = prim_mat_.data();
    view_.prim_geom = prim_geom_.data();
    view_.nmat = (int32_t)nm;
    view_.mat_kind = mat_kind_.data();
    view_.mat_data = mat_data_.data();
    view_.nee_triangles = nee_triangles ? 1 : 0;
    return view_;
}

void Camera::basis(double aspect, Vec3& origin, Vec3& lower, Vec3& horiz, Vec3& vert) const {  // renderer.go:947-960
    double theta = FOV * M_PI / 180;
    double h = std::tan(theta / 2);
    double vh = 2 * h, vw = aspect * vh;
    Vec3 w = Unit(Position - Target);
    Vec3 u = Unit(Cross(Up, w));
    Vec3 v = Cross(w, u);
    origin = Position;
    horiz = u * vw;
    vert = v * vh;
    lower = origin - horiz * .5 - vert * .5 - w;
}

RenderSettings Quality(const std::string& name) {  // renderer.go:1064-1117 (fields the GPU path reads)
    RenderSettings b;
    b.Backend = "Auto"; b.Sampler = "Halton"; b.MIS = true; b.PowerLightSampling = true;
    b.FireflyClamp = 24; b.RRDepth = 4; b.PublishHz = 30; b.ProgressiveUpdates = true;
    b.AdaptiveSampling = true; b.TAA = 1; b.MSAA = 1; b.FXAA = 1; b.BVHBins = 16; b.BVHLeafSize = 4;
    auto set = [&](int w, int h, int spp, int bounces, bool denoise, const std::string& q) {
        b.Width = w; b.Height = h; b.SPP = spp; b.Bounces = bounces; b.Denoise = denoise; b.Quality = q;
    };
    if (name == "Draft") {
        set(640, 360, 2, 4, false, name); b.FXAA = 1; b.TAA = 0; b.FireflyClamp = 12;
    } else if (name == "Balanced") {
        set(1280, 720, 16, 8, true, name); b.FXAA = 2; b.TAA = 2; b.HBAO = 1;
    } else if (name == "High") {
        set(1600, 900, 48, 10, true, name); b.FXAA = 2; b.TAA = 2; b.TSAA = 1; b.HBAOPlus = 2; b.MSAA = 2; b.FireflyClamp = 32;
    } else if (name == "Ultra") {
        set(1920, 1080, 96, 12, true, name); b.FXAA = 3; b.TAA = 3; b.TSAA = 2; b.TXAA = 1; b.HBAOPlus = 3; b.MSAA = 2; b.FireflyClamp = 48;
    } else if (name == "Ultra Realism") {
        set(2560, 1440, 256, 16, true, name); b.FXAA = 0; b.TAA = 1; b.TSAA = 0; b.TXAA = 0; b.MSAA = 1;
        b.HBAO = 0; b.HBAOPlus = 0; b.FireflyClamp = 0; b.RRDepth = 6; b.ProgressiveUpdates = false;
        b.BVHBins = 32; b.BVHLeafSize = 4;
    } else {
        set(960, 540, 8, 6, true, "Preview"); b.FXAA = 1; b.TAA = 1;
    }
    return b;
}

RenderSettings UltraRealismSettings(const std::string& backend) {  // renderer.go:1119-1135
    RenderSettings q = Quality("Ultra Realism");
    q.Backend = backend;
    if (backend == "OpenCL GPU") {
        q.Sampler = "Random"; q.MIS = false; q.PowerLightSampling = false; q.FireflyClamp = 0;
        q.AdaptiveSampling = false; q.ProgressiveUpdates = false; q.TAA = 1; q.Denoise = true;
    }
    return q;
}

int effectiveSPP(const RenderSettings& s) {  // renderer.go:2395-2417
    int msaa = s.MSAA;
    if (msaa != 2 && msaa != 4 && msaa != 8) msaa = 1;
    int tsaa = 1;
    if (s.TSAA == 1) tsaa = 2; else if (s.TSAA == 2) tsaa = 3; else if (s.TSAA == 3) tsaa = 4;
    int v = s.SPP * msaa * tsaa;
    return v < 1 ? 1 : (v > 4096 ? 4096 : v);
}

void internalSize(const RenderSettings& s, int& w, int& h) {  // renderer.go:2066-2079
    double scale = 1.0;
    if (s.Upscale == "Ultra Quality") scale = 0.77;
    else if (s.Upscale == "Balanced") scale = 0.67;
    else if (s.Upscale == "Performance") scale = 0.50;
    w = std::max(16, (int)std::round((double)s.Width * scale));
    h = std::max(16, (int)std::round((double)s.Height * scale));
}

namespace {
// persistentCL analogue (gpu_opencl_windows.go:94-131): one context, one lock, scene cache.
struct Persistent {
    std::mutex mu;
    bc_ctx* ctx = nullptr;
    const Scene* scene = nullptr;
    uint64_t topo = 0, geom = 0;
};
Persistent g_rt;

Vec3 decodeOctNormal(uint32_t qx, uint32_t qy) {  // gpu_opencl_windows.go:553-563
    double x = (double)qx / 1023.0 * 2 - 1, y = (double)qy / 1023.0 * 2 - 1;
    double z = 1 - std::fabs(x) - std::fabs(y);
    if (z < 0) {
        double ox = x, oy = y;
        x = (1 - std::fabs(oy)) * std::copysign(1.0, ox);
        y = (1 - std::fabs(ox)) * std::copysign(1.0, oy);
    }
    return Unit(V(x, y, z));
}
Vec3 materialPreviewColor(const Material& m) {  // renderer.go:1467-1475
    auto c01 = [](double v) { return std::min(1.0, std::max(0.0, v)); };
    if (m.Kind == MatEmissive) return V(c01(m.Emission.x / 8), c01(m.Emission.y / 8), c01(m.Emission.z / 8));
    if (Dot(m.Color, m.Color) > 0) return m.Color;
    return V(0.8, 0.8, 0.8);
}
}  // namespace

// renderOpenCL (gpu_opencl_windows.go:565-723) over the C ABI: same chunking, output
// decoding, guide decoding and error/cancel behaviour.
bc_render_desc renderDescFor(const Camera& cam, const RenderSettings& settings) {
    int w = 0, h = 0;
    internalSize(settings, w, h);
    Vec3 origin, lower, horiz, vert;
    cam.basis((double)w / (double)h, origin, lower, horiz, vert);
    bc_render_desc d{};
    d.width = w; d.height = h; d.spp = effectiveSPP(settings); d.bounces = settings.Bounces;
    d.rr_depth = settings.RRDepth;
    d.sampler = settings.Sampler == "Halton" ? 1 : (settings.Sampler == "R2" ? 2 : 0);
    d.mis = settings.MIS; d.power_lights = settings.PowerLightSampling;
    d.firefly_clamp = (float)settings.FireflyClamp;
    d.seed = 0x9e3779b9u;  // the reference kernel's seed argument (gpu_opencl_windows.go:606)
    const Vec3* src[4] = {&origin, &lower, &horiz, &vert};
    double* dst[4] = {d.origin, d.lower, d.horiz, d.vert};
    for (int k = 0; k < 4; k++) { dst[k][0] = src[k]->x; dst[k][1] = src[k]->y; dst[k][2] = src[k]->z; }
    return d;
}

void renderHIPForget(const Scene* scene) {
    std::lock_guard<std::mutex> lock(g_rt.mu);
    if (!scene) {
        if (g_rt.ctx) bc_destroy(g_rt.ctx);
        g_rt.ctx = nullptr;
        g_rt.scene = nullptr;
    } else if (g_rt.scene == scene) {
        g_rt.scene = nullptr;
    }
}

int renderHIP(Scene& scene, const Camera& cam, const RenderSettings& settings, const volatile int32_t* cancel,
              const ChunkFn& onChunk, std::vector<Vec3>& linear, std::vector<Guide>& guides, int64_t& rays,
              std::string& deviceName, std::string& err, int* sceneAction) {
    rays = 0;
    if (sceneAction) *sceneAction = 0;
    if (scene.Primitives.empty() || scene.Nodes.empty()) {
        err = "HIP backend requires a non-empty built scene";
        return BC_ERR_SCENE;
    }
    std::lock_guard<std::mutex> lock(g_rt.mu);
    if (!g_rt.ctx) {
        g_rt.ctx = bc_create(0);
        if (!g_rt.ctx) { err = bc_last_error(nullptr); return BC_ERR_DEVICE; }
    }
    bc_ctx* ctx = g_rt.ctx;
    deviceName = bc_device_name(ctx);
    // ensureSceneLocked (gpu_opencl_windows.go:379-445): a geometry-only change refits the
    // resident BVH on the device (updateGeometryLocked); anything else uploads the scene
    bool refitted = false;
    if (g_rt.scene == &scene && g_rt.topo == scene.GPUTopologyRevision && g_rt.geom != scene.GPUGeometryRevision) {
        refitted = bc_scene_refit(ctx, &scene.View(false), scene.GPUGeometryRevision) == BC_OK;
        if (refitted) {
            g_rt.geom = scene.GPUGeometryRevision;
            if (sceneAction) *sceneAction = 1;
        }
    }
    if (!refitted &&
        (g_rt.scene != &scene || g_rt.topo != scene.GPUTopologyRevision || g_rt.geom != scene.GPUGeometryRevision)) {
        g_rt.scene = nullptr;  // whatever was resident is replaced, even if this upload fails
        int rc = bc_scene_upload(ctx, &scene.View(false), scene.GPUTopologyRevision, scene.GPUGeometryRevision);
        if (rc != BC_OK) { err = bc_last_error(ctx); return rc; }
        g_rt.scene = &scene;
        g_rt.topo = scene.GPUTopologyRevision;
        g_rt.geom = scene.GPUGeometryRevision;
        if (sceneAction) *sceneAction = 2;
    }
    const bc_render_desc d = renderDescFor(cam, settings);
    const int w = d.width, h = d.height;
    const bool needGuides = settings.Denoise || settings.HBAO > 0 || settings.HBAOPlus > 0 ||
                            std::max(settings.TAA, std::max(settings.TSAA, settings.TXAA)) > 0 || settings.DebugView != 0;
    const Vec3 origin = V(d.origin[0], d.origin[1], d.origin[2]), lower = V(d.lower[0], d.lower[1], d.lower[2]);
    const Vec3 horiz = V(d.horiz[0], d.horiz[1], d.horiz[2]), vert = V(d.vert[0], d.vert[1], d.vert[2]);
    int chunkRows = 64;
    if (!settings.ProgressiveUpdates) chunkRows = h;
    else if (settings.PublishHz <= 15) chunkRows = 128;
    else if (settings.PublishHz >= 60) chunkRows = 32;
    linear.assign((size_t)w * h, Vec3{});
    guides.assign(needGuides ? (size_t)w * h : 0, Guide{});
    std::vector<float> raw((size_t)w * chunkRows * 4);
    std::vector<uint32_t> rawGuide(needGuides ? (size_t)w * chunkRows * 3 : 0);
    for (int y0 = 0; y0 < h; y0 += chunkRows) {
        if (cancel && *cancel) { err = "render cancelled"; return BC_ERR_CANCELLED; }
        int rows = std::min(chunkRows, h - y0);
        uint64_t chunkRays = 0;
        int rc = bc_render(ctx, &d, y0, rows, raw.data(), needGuides ? rawGuide.data() : nullptr, &chunkRays, nullptr);
        if (rc != BC_OK) { err = bc_last_error(ctx); return rc; }
        const int base = y0 * w, n = rows * w;
        for (int i = 0; i < n; i++) linear[(size_t)(base + i)] = V(raw[4 * (size_t)i], raw[4 * (size_t)i + 1], raw[4 * (size_t)i + 2]);
        rays += (int64_t)chunkRays;
        if (needGuides) {
            const double invW = 1.0 / (double)std::max(w - 1, 1), invH = 1.0 / (double)std::max(h - 1, 1);
            for (int i = 0; i < n; i++) {
                uint32_t packed = rawGuide[3 * (size_t)i + 1];
                if (!(packed & 0x80000000u)) continue;
                float depthf;
                std::memcpy(&depthf, &rawGuide[3 * (size_t)i], 4);
                uint32_t jit = rawGuide[3 * (size_t)i + 2];
                int pix = base + i, y = pix / w, x = pix - y * w;
                double u = ((double)x + (double)(jit & 0xffff) / 65535.0) * invW;
                double v = ((double)(h - 1 - y) + (double)(jit >> 16) / 65535.0) * invH;
                Vec3 dir = Unit(lower + horiz * u + vert * v - origin);
                int mi = (int)((packed >> 20) & 2047);
                Guide g;
                g.Depth = depthf;
                g.P = origin + dir * g.Depth;
                g.N = decodeOctNormal(packed & 1023, (packed >> 10) & 1023);
                g.Albedo = (mi < (int)scene.Materials.size()) ? materialPreviewColor(scene.Materials[(size_t)mi]) : V(0.8, 0.8, 0.8);
                g.Valid = true;
                guides[(size_t)pix] = g;
            }
        }
        if (onChunk) {
            std::vector<Vec3> part(linear.begin() + base, linear.begin() + base + n);
            onChunk(y0, rows, part, (double)(y0 + rows) / (double)h);
        }
    }
    return BC_OK;
}

}  // namespace bcx

// ------------------------------------------------------------------------------ C entry points
namespace {
using bcx::V;
using bcx::Vec3;
bcx::RenderSettings settings_from(const bch_settings& c) {
    bcx::RenderSettings s;
    s.Width = c.width; s.Height = c.height; s.SPP = c.spp; s.Bounces = c.bounces; s.RRDepth = c.rr_depth;
    s.Sampler = c.sampler == 1 ? "Halton" : (c.sampler == 2 ? "R2" : "Random");
    s.MIS = c.mis != 0; s.PowerLightSampling = c.power_lights != 0;
    s.MSAA = c.msaa; s.TSAA = c.tsaa; s.TAA = c.taa; s.TXAA = c.txaa; s.HBAO = c.hbao; s.HBAOPlus = c.hbao_plus;
    s.DebugView = c.debug_view; s.Denoise = c.denoise != 0;
    s.ProgressiveUpdates = c.progressive != 0; s.PublishHz = c.publish_hz;
    static const char* const kUpscale[4] = {"Off", "Ultra Quality", "Balanced", "Performance"};
    s.Upscale = kUpscale[c.upscale >= 0 && c.upscale < 4 ? c.upscale : 0];
    s.FireflyClamp = c.firefly_clamp;
    return s;
}
bcx::Camera camera_from(const double* c10) {
    bcx::Camera cam;
    cam.Position = V(c10[0], c10[1], c10[2]);
    cam.Target = V(c10[3], c10[4], c10[5]);
    cam.Up = V(c10[6], c10[7], c10[8]);
    cam.FOV = c10[9];
    return cam;
}
void put(char* dst, int cap, const std::string& s) {
    if (dst && cap > 0) std::snprintf(dst, (size_t)cap, "%s", s.c_str());
}
}  // namespace

extern "C" {

int bch_settings_quality(const char* name, bch_settings* o) {
    if (!o) return BC_ERR_ARG;
    const bcx::RenderSettings q = bcx::Quality(name ? name : "");
    std::memset(o, 0, sizeof *o);
    o->width = q.Width; o->height = q.Height; o->spp = q.SPP; o->bounces = q.Bounces; o->rr_depth = q.RRDepth;
    o->sampler = q.Sampler == "Halton" ? 1 : (q.Sampler == "R2" ? 2 : 0);
    o->mis = q.MIS; o->power_lights = q.PowerLightSampling;
    o->msaa = q.MSAA; o->tsaa = q.TSAA; o->taa = q.TAA; o->txaa = q.TXAA; o->hbao = q.HBAO; o->hbao_plus = q.HBAOPlus;
    o->debug_view = q.DebugView; o->denoise = q.Denoise;
    o->progressive = q.ProgressiveUpdates; o->publish_hz = q.PublishHz;
    o->upscale = q.Upscale == "Ultra Quality" ? 1 : (q.Upscale == "Balanced" ? 2 : (q.Upscale == "Performance" ? 3 : 0));
    o->firefly_clamp = q.FireflyClamp;
    return BC_OK;
}

int bch_render_desc_for(const double* camera10, const bch_settings* st, bc_render_desc* out) {
    if (!camera10 || !st || !out) return BC_ERR_ARG;
    *out = bcx::renderDescFor(camera_from(camera10), settings_from(*st));
    return BC_OK;
}

int bch_render_hip(bch_scene* s, const double* camera10, const bch_settings* st, const volatile int32_t* cancel,
                   bch_chunk_fn on_chunk, void* user, double* linear, bch_guide* guides, int64_t* rays_out,
                   int32_t* scene_action, char* device_name, int name_cap, char* err, int err_cap) {
    if (!s || !camera10 || !st || !linear) {
        put(err, err_cap, "invalid arguments");
        return BC_ERR_ARG;
    }
    try {
        const bcx::RenderSettings rs = settings_from(*st);
        bcx::ChunkFn fn;
        int w = 0, h = 0;
        bcx::internalSize(rs, w, h);
        std::vector<double> flat;
        if (on_chunk)
            fn = [&](int y0, int rows, const std::vector<Vec3>& data, double progress) {
                flat.resize(data.size() * 3);
                for (size_t i = 0; i < data.size(); i++) {
                    flat[3 * i] = data[i].x; flat[3 * i + 1] = data[i].y; flat[3 * i + 2] = data[i].z;
                }
                on_chunk(user, y0, rows, flat.data(), progress);
            };
        std::vector<Vec3> lin;
        std::vector<bcx::Guide> gd;
        int64_t rays = 0;
        int action = 0;
        std::string name, msg;
        const int rc = bcx::renderHIP(s->s, camera_from(camera10), rs, cancel, fn, lin, gd, rays, name, msg, &action);
        put(device_name, name_cap, name);
        if (scene_action) *scene_action = action;
        if (rc != BC_OK) {
            put(err, err_cap, msg);
            return rc;
        }
        for (size_t i = 0; i < lin.size(); i++) {
            linear[3 * i] = lin[i].x; linear[3 * i + 1] = lin[i].y; linear[3 * i + 2] = lin[i].z;
        }
        if (guides) {
            for (size_t i = 0; i < (size_t)w * (size_t)h; i++) {
                bch_guide& o = guides[i];
                std::memset(&o, 0, sizeof o);
                if (i >= gd.size() || !gd[i].Valid) continue;
                const bcx::Guide& g = gd[i];
                const Vec3 v3[3] = {g.P, g.N, g.Albedo};
                double* d3[3] = {o.p, o.n, o.albedo};
                for (int k = 0; k < 3; k++) { d3[k][0] = v3[k].x; d3[k][1] = v3[k].y; d3[k][2] = v3[k].z; }
                o.depth = g.Depth;
                o.valid = 1;
            }
        }
        if (rays_out) *rays_out = rays;
        return BC_OK;
    } catch (const std::bad_alloc&) {
        put(err, err_cap, "out of host memory");
        return BC_ERR_NOMEM;
    } catch (const std::exception& e) {
        put(err, err_cap, e.what());
        return BC_ERR_DEVICE;
    }
}

void bch_render_hip_release(void) { bcx::renderHIPForget(nullptr); }

}  // extern "C"
