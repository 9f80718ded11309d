// host_scene.hpp -- host-side mirror of the Go renderer's scene, settings and GPU render
// entry (windows_portable/renderer.go).  In the shipped product the Go code owns these
// types and calls the C ABI through cgo; this C++ mirror stands in for it where no Go
// toolchain exists, with the same names, argument meaning and error behaviour.
#pragma once
#include <atomic>
#include <cstdint>
#include <functional>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/beamcast_hip.h"

namespace bcx {

struct Vec3 {
    double x = 0, y = 0, z = 0;
};
inline Vec3 V(double x, double y, double z) { return Vec3{x, y, z}; }
inline Vec3 operator+(Vec3 a, Vec3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
inline double Dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 Cross(Vec3 a, Vec3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
Vec3 Unit(Vec3 a);
// math.Min / math.Max: NaN wins, -0 is below +0
inline double GoMin(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a == 0 && b == 0) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
inline double GoMax(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a == 0 && b == 0) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}

enum MaterialKind { MatDiffuse = 0, MatMetal = 1, MatGlass = 2, MatEmissive = 3 };  // renderer.go:86-91
enum PrimKind { PrimSphere = 0, PrimTriangle = 1 };                               // renderer.go:101-104

struct Material {  // renderer.go:93-99
    int Kind = MatDiffuse;
    Vec3 Color, Emission;
    double Roughness = 0, IOR = 0;
};
struct AABB {
    Vec3 Min, Max;
};
struct Primitive {  // renderer.go:106-116
    int Kind = PrimTriangle, Material = 0;
    Vec3 Center;
    double Radius = 0;
    Vec3 A, B, C, E1, E2, Normal;
    AABB Box;
    Vec3 Centroid;
};
struct BVHNode {  // renderer.go:212-216
    AABB Box;
    int Left = -1, Right = -1, Start = 0, Count = 0;
};

// Scene (renderer.go:225-241) with Build / Refit / revision tracking.
class Scene {
public:
    std::vector<Material> Materials;
    std::vector<Primitive> Primitives;
    std::vector<BVHNode> Nodes;
    std::vector<int> Order;
    std::string Name;
    std::string BVHMode = "SAH";  // "SAH" or "Median" (case-insensitive), as Scene.BVHMode
    int BVHBins = 0, BVHLeafSize = 0;
    uint64_t GPUTopologyRevision = 0, GPUGeometryRevision = 0;

    int AddMaterial(const Material& m);
    void AddSphere(Vec3 c, double r, int mat);
    void AddTriangle(Vec3 a, Vec3 b, Vec3 c, int mat);
    void Build();
    // Build with the SAH tree made on the GPU (bc_build_bvh): the same Nodes and Order as
    // Build(); returns the status of the call (the scene is unchanged on an error).
    int BuildDevice(bc_ctx* ctx, std::string& err);
    void Refit();
    // Flat views for the C ABI; valid until the scene is mutated.
    const bc_scene_view& View(bool nee_triangles);

private:
    int buildNodeSAH(int start, int end);
    int buildNodeMedian(int start, int end);
    std::vector<double> node_box_, prim_geom_, mat_data_;
    std::vector<int32_t> node_link_, order_, prim_kind_, prim_mat_, mat_kind_;
    bc_scene_view view_{};
};

struct Camera {  // renderer.go:941-960
    Vec3 Position, Target, Up;
    double FOV = 40;
    void basis(double aspect, Vec3& origin, Vec3& lower, Vec3& horiz, Vec3& vert) const;
};

// RenderSettings subset read by the GPU path (renderer.go:1027-1062) and the presets.
struct RenderSettings {
    int Width = 0, Height = 0, SPP = 1, Bounces = 4;
    bool Denoise = false;
    std::string Quality, Backend = "Auto", Sampler = "Halton", Upscale = "Off";
    int MSAA = 1, TSAA = 0, TAA = 0, TXAA = 0, FXAA = 1, HBAO = 0, HBAOPlus = 0, DebugView = 0;
    bool MIS = true, PowerLightSampling = true, AdaptiveSampling = true, ProgressiveUpdates = true;
    double FireflyClamp = 24;
    int RRDepth = 4, PublishHz = 30;
    int BVHBins = 16, BVHLeafSize = 4;
};
RenderSettings Quality(const std::string& name);            // renderer.go:1064-1117
RenderSettings UltraRealismSettings(const std::string& backend);  // renderer.go:1119-1135
int effectiveSPP(const RenderSettings& s);                  // renderer.go:2395-2417
void internalSize(const RenderSettings& s, int& w, int& h); // renderer.go:2066-2079

struct Guide {  // renderer.go:726-731
    Vec3 P, N, Albedo;
    double Depth = 0;
    bool Valid = false;
};

// renderOpenCL's contract (gpu_opencl_windows.go:565-566) over the HIP backend.  cancel (may be
// null) is the atomic.Bool the Go caller sets, polled before every chunk.  sceneAction (may be
// null) reports what ensureSceneLocked did: 0 the resident scene was reused, 1 refit on the
// device (only the geometry revision changed), 2 full upload.
using ChunkFn = std::function<void(int y0, int rows, const std::vector<Vec3>& data, double progress)>;
int renderHIP(Scene& scene, const Camera& cam, const RenderSettings& settings, const volatile int32_t* cancel,
              const ChunkFn& onChunk, std::vector<Vec3>& linear, std::vector<Guide>& guides,
              int64_t& rays, std::string& deviceName, std::string& err, int* sceneAction = nullptr);
// The bc_render_desc renderHIP renders a frame with: internalSize, effectiveSPP, the sampler and
// light settings, the camera basis and the reference kernel's seed.
bc_render_desc renderDescFor(const Camera& cam, const RenderSettings& settings);
// The persistent runtime forgets `scene` (it is being freed), so a later scene at the same
// address is uploaded rather than taken for the resident one; null releases the runtime.
void renderHIPForget(const Scene* scene);

}  // namespace bcx

// The opaque scene handle of the C entry points (beamcast_host.h).
struct bch_scene {
    bcx::Scene s;
};
