//go:build linux && cgo && beamcast_hip

// gpu_hip_linux.go -- cgo shim that binds the Go renderer's GPU entry points to
// libbeamcast_hip.so.  Dropped into windows_portable/ (package main) it provides the
// functions that gpu_opencl_windows.go / gpu_opencl_stub.go provide today:
//
//	openCLGPUAvailable() (bool, string)                                       gpu_opencl_windows.go:172
//	renderOpenCL(scene, cam, settings, cancel, onChunk) (linear, guides, rays, device, err)  :565-723
//
// Build with: go build -tags beamcast_hip (CGO_LDFLAGS=-L<dir> -lbeamcast_hip, and the
// include/ directory of this repository on CGO_CFLAGS).  The stub must then carry the
// matching build tag "!windows && !beamcast_hip".  diff_hip_linux.go (same tag) adds the
// differentiable render on the same context.
package main

/*
#cgo LDFLAGS: -lbeamcast_hip
#include <stdlib.h>
#include "beamcast_hip.h"
*/
import "C"

import (
	"errors"
	"math"
	"sync"
	"sync/atomic"
	"unsafe"
)

// hipRuntime mirrors persistentCL (gpu_opencl_windows.go:94-131): one context, one lock,
// scene residency keyed by pointer and revisions.
type hipRuntime struct {
	mu          sync.Mutex
	ctx         *C.bc_ctx
	scene       *Scene
	topologyRev uint64
	geometryRev uint64
	// pinned C copies of the scene view (kept until the next upload)
	cbuf []unsafe.Pointer
}

var persistentHIP hipRuntime

// gpuFullTransport tells RenderState.Start's GPU gate (renderer.go:2649, widened by
// renderer_gate.patch) that this backend renders the CPU integrator's whole path transport:
// MIS, power-weighted light selection, Halton and R2 sampling and the firefly clamp.  The
// OpenCL kernel and the stub do not (gpu_transport_opencl.go).
const gpuFullTransport = true

func openCLGPUAvailable() (bool, string) {
	buf := (*C.char)(C.malloc(256))
	defer C.free(unsafe.Pointer(buf))
	ok := C.bc_device_available(buf, 256)
	return ok != 0, C.GoString(buf)
}

func hipErr(ctx *C.bc_ctx) error { return errors.New(C.GoString(C.bc_last_error(ctx))) }

func (rt *hipRuntime) cAlloc(n int) unsafe.Pointer {
	p := C.malloc(C.size_t(n))
	rt.cbuf = append(rt.cbuf, p)
	return p
}

func (rt *hipRuntime) freeView() {
	for _, p := range rt.cbuf {
		C.free(p)
	}
	rt.cbuf = rt.cbuf[:0]
}

// ensureSceneLocked packs Scene.Nodes / Order / Primitives / Materials into C memory and
// uploads them (the BVH8 flattening happens inside the library).
func (rt *hipRuntime) ensureSceneLocked(scene *Scene) error {
	if rt.scene == scene && rt.topologyRev == scene.GPUTopologyRevision && rt.geometryRev == scene.GPUGeometryRevision {
		return nil
	}
	rt.freeView()
	nn, np, nm := len(scene.Nodes), len(scene.Primitives), len(scene.Materials)
	box := unsafe.Slice((*C.double)(rt.cAlloc(8*6*nn+8)), 6*nn)
	link := unsafe.Slice((*C.int32_t)(rt.cAlloc(4*4*nn+4)), 4*nn)
	for i, n := range scene.Nodes {
		box[6*i], box[6*i+1], box[6*i+2] = C.double(n.Box.Min.X), C.double(n.Box.Min.Y), C.double(n.Box.Min.Z)
		box[6*i+3], box[6*i+4], box[6*i+5] = C.double(n.Box.Max.X), C.double(n.Box.Max.Y), C.double(n.Box.Max.Z)
		link[4*i], link[4*i+1], link[4*i+2], link[4*i+3] = C.int32_t(n.Left), C.int32_t(n.Right), C.int32_t(n.Start), C.int32_t(n.Count)
	}
	order := unsafe.Slice((*C.int32_t)(rt.cAlloc(4*np+4)), np)
	kind := unsafe.Slice((*C.int32_t)(rt.cAlloc(4*np+4)), np)
	pmat := unsafe.Slice((*C.int32_t)(rt.cAlloc(4*np+4)), np)
	geom := unsafe.Slice((*C.double)(rt.cAlloc(8*10*np+8)), 10*np)
	for i, o := range scene.Order {
		order[i] = C.int32_t(o)
	}
	for i := range scene.Primitives {
		p := &scene.Primitives[i]
		pmat[i] = C.int32_t(p.Material)
		g := geom[10*i : 10*i+10]
		if p.Kind == PrimSphere {
			kind[i] = 0
			g[0], g[1], g[2], g[3] = C.double(p.Center.X), C.double(p.Center.Y), C.double(p.Center.Z), C.double(p.Radius)
		} else {
			kind[i] = 1
			g[0], g[1], g[2] = C.double(p.A.X), C.double(p.A.Y), C.double(p.A.Z)
			g[3], g[4], g[5] = C.double(p.E1.X), C.double(p.E1.Y), C.double(p.E1.Z)
			g[6], g[7], g[8] = C.double(p.E2.X), C.double(p.E2.Y), C.double(p.E2.Z)
		}
	}
	mkind := unsafe.Slice((*C.int32_t)(rt.cAlloc(4*nm+4)), nm)
	mdata := unsafe.Slice((*C.double)(rt.cAlloc(8*8*nm+8)), 8*nm)
	fillMaterials(mkind, mdata, scene.Materials)
	view := (*C.bc_scene_view)(rt.cAlloc(int(unsafe.Sizeof(C.bc_scene_view{}))))
	*view = C.bc_scene_view{
		nnode: C.int32_t(nn), node_box: &box[0], node_link: &link[0],
		nprim: C.int32_t(np), order: &order[0], prim_kind: &kind[0], prim_mat: &pmat[0], prim_geom: &geom[0],
		nmat: C.int32_t(nm), mat_kind: &mkind[0], mat_data: &mdata[0], nee_triangles: 0,
	}
	// geometry-only change (updateGeometryLocked): refit the resident BVH8 on the device
	if rt.scene == scene && rt.topologyRev == scene.GPUTopologyRevision {
		if C.bc_scene_refit(rt.ctx, view, C.uint64_t(scene.GPUGeometryRevision)) == 0 {
			rt.geometryRev = scene.GPUGeometryRevision
			return nil
		}
	}
	if rc := C.bc_scene_upload(rt.ctx, view, C.uint64_t(scene.GPUTopologyRevision), C.uint64_t(scene.GPUGeometryRevision)); rc != 0 {
		return hipErr(rt.ctx)
	}
	rt.scene, rt.topologyRev, rt.geometryRev = scene, scene.GPUTopologyRevision, scene.GPUGeometryRevision
	return nil
}

// fillMaterials packs Material records as bc_scene_view's mat_kind / mat_data (8 doubles each:
// Color, Roughness, IOR, Emission), the material part of uploadFullSceneLocked
// (gpu_opencl_windows.go:384-396).
func fillMaterials(mkind []C.int32_t, mdata []C.double, mats []Material) {
	for i, m := range mats {
		mkind[i] = C.int32_t(m.Kind)
		d := mdata[8*i : 8*i+8]
		d[0], d[1], d[2], d[3], d[4] = C.double(m.Color.X), C.double(m.Color.Y), C.double(m.Color.Z), C.double(m.Roughness), C.double(m.IOR)
		d[5], d[6], d[7] = C.double(m.Emission.X), C.double(m.Emission.Y), C.double(m.Emission.Z)
	}
}

// lockedContext takes the runtime lock and creates the context on first use; the caller unlocks.
func (rt *hipRuntime) lockedContext() error {
	rt.mu.Lock()
	if rt.ctx == nil {
		rt.ctx = C.bc_create(0)
		if rt.ctx == nil {
			rt.mu.Unlock()
			return hipErr(nil)
		}
	}
	return nil
}

// hipRenderDesc fills the render description of one frame from the camera and settings, as
// renderOpenCL's kernel arguments are filled (gpu_opencl_windows.go:580-611), and returns the
// internal image size and the camera basis.
func hipRenderDesc(cam Camera, settings RenderSettings) (d C.bc_render_desc, w, h int, origin, lower, horiz, vert Vec3) {
	w, h = internalSize(settings)
	origin, lower, horiz, vert = cam.basis(float64(w) / float64(h))
	d.width, d.height = C.int32_t(w), C.int32_t(h)
	d.spp, d.bounces, d.rr_depth = C.int32_t(effectiveSPP(&settings)), C.int32_t(settings.Bounces), C.int32_t(settings.RRDepth)
	switch settings.Sampler {
	case SamplerHalton:
		d.sampler = 1
	case SamplerR2:
		d.sampler = 2
	}
	if settings.MIS {
		d.mis = 1
	}
	if settings.PowerLightSampling {
		d.power_lights = 1
	}
	d.firefly_clamp = C.float(settings.FireflyClamp)
	d.seed = 0x9e3779b9
	for k, v := range [4]Vec3{origin, lower, horiz, vert} {
		arr := [4]*[3]C.double{&d.origin, &d.lower, &d.horiz, &d.vert}[k]
		arr[0], arr[1], arr[2] = C.double(v.X), C.double(v.Y), C.double(v.Z)
	}
	return
}

func renderOpenCL(scene *Scene, cam Camera, settings RenderSettings, cancel *atomic.Bool,
	onChunk func(y0, rows int, data []Vec3, progress float64)) ([]Vec3, []Guide, int64, string, error) {
	if scene == nil || len(scene.Primitives) == 0 || len(scene.Nodes) == 0 {
		return nil, nil, 0, "", errors.New("HIP backend requires a non-empty built scene")
	}
	rt := &persistentHIP
	if err := rt.lockedContext(); err != nil {
		return nil, nil, 0, "", err
	}
	defer rt.mu.Unlock()
	device := C.GoString(C.bc_device_name(rt.ctx))
	if err := rt.ensureSceneLocked(scene); err != nil {
		return nil, nil, 0, device, err
	}
	d, w, h, origin, lower, horiz, vert := hipRenderDesc(cam, settings)
	needGuides := settings.Denoise || settings.HBAO > 0 || settings.HBAOPlus > 0 || temporalWeight(settings) > 0 || settings.DebugView != DebugBeauty
	chunkRows := 64
	if !settings.ProgressiveUpdates {
		chunkRows = h
	} else if settings.PublishHz <= 15 {
		chunkRows = 128
	} else if settings.PublishHz >= 60 {
		chunkRows = 32
	}
	raw := (*C.float)(C.malloc(C.size_t(16 * w * chunkRows)))
	defer C.free(unsafe.Pointer(raw))
	var rawGuide *C.uint32_t
	if needGuides {
		rawGuide = (*C.uint32_t)(C.malloc(C.size_t(12 * w * chunkRows)))
		defer C.free(unsafe.Pointer(rawGuide))
	}
	linear := make([]Vec3, w*h)
	var guides []Guide
	if needGuides {
		guides = make([]Guide, w*h)
	}
	var rays int64
	for y0 := 0; y0 < h; y0 += chunkRows {
		if cancel != nil && cancel.Load() {
			return nil, nil, 0, device, errors.New("render cancelled")
		}
		rows := chunkRows
		if y0+rows > h {
			rows = h - y0
		}
		var chunkRays C.uint64_t
		if rc := C.bc_render(rt.ctx, &d, C.int(y0), C.int(rows), raw, rawGuide, &chunkRays, nil); rc != 0 {
			return nil, nil, 0, device, hipErr(rt.ctx)
		}
		px := unsafe.Slice((*float32)(unsafe.Pointer(raw)), 4*w*rows)
		base := y0 * w
		for i := 0; i < w*rows; i++ {
			linear[base+i] = V(float64(px[4*i]), float64(px[4*i+1]), float64(px[4*i+2]))
		}
		rays += int64(chunkRays)
		if needGuides {
			gd := unsafe.Slice((*uint32)(unsafe.Pointer(rawGuide)), 3*w*rows)
			invW, invH := 1.0/float64(iMax(w-1, 1)), 1.0/float64(iMax(h-1, 1))
			for i := 0; i < w*rows; i++ {
				packed := gd[3*i+1]
				if packed&0x80000000 == 0 {
					continue
				}
				depth := float64(math.Float32frombits(gd[3*i]))
				jitter := gd[3*i+2]
				pix := base + i
				y := pix / w
				x := pix - y*w
				u := (float64(x) + float64(jitter&0xffff)/65535.0) * invW
				v := (float64(h-1-y) + float64(jitter>>16)/65535.0) * invH
				dir := lower.Add(horiz.Mul(u)).Add(vert.Mul(v)).Sub(origin).Unit()
				mi := int((packed >> 20) & 2047)
				albedo := V(0.8, 0.8, 0.8)
				if mi < len(scene.Materials) {
					albedo = materialPreviewColor(&scene.Materials[mi])
				}
				guides[pix] = Guide{P: origin.Add(dir.Mul(depth)), N: decodeOctNormal(packed&1023, (packed>>10)&1023), Albedo: albedo, Depth: depth, Valid: true}
			}
		}
		if onChunk != nil {
			onChunk(y0, rows, linear[base:base+w*rows], float64(y0+rows)/float64(h))
		}
	}
	return linear, guides, rays, device, nil
}

// decodeOctNormal as gpu_opencl_windows.go:553-563 (that file is windows-only).
func decodeOctNormal(qx, qy uint32) Vec3 {
	x := float64(qx)/1023.0*2 - 1
	y := float64(qy)/1023.0*2 - 1
	z := 1 - math.Abs(x) - math.Abs(y)
	if z < 0 {
		ox, oy := x, y
		x = (1 - math.Abs(oy)) * math.Copysign(1, ox)
		y = (1 - math.Abs(ox)) * math.Copysign(1, oy)
	}
	return V(x, y, z).Unit()
}
