//go:build linux && cgo && beamcast_hip

// diff_hip_linux.go -- differentiable rendering over libbeamcast_hip.so (no counterpart in the
// reference renderer): a Forward/Backward function object that gives the gradient of an image
// loss with respect to every material's Color and Emission by path-replay backpropagation
// (bc_forward keeps each sample's path radiance on the device, bc_backward replays every sample
// from its stored key; there is no per-bounce tape).  It shares the process-wide HIP context,
// its lock and the scene residency with renderOpenCL (gpu_hip_linux.go).
package main

/*
#include <stdlib.h>
#include "beamcast_hip.h"
*/
import "C"

import (
	"errors"
	"unsafe"
)

// DiffRender renders one scene from one view with fixed settings.  A typical loop:
//
//	f, _ := NewDiffRender(scene, cam, settings)
//	img, _, _ := f.Forward()
//	dColor, dEmission, _ := f.Backward(dLoss(img))   // d loss / d pixel, same layout as img
//	... update scene.Materials[i].Color / Emission ...
//	f.SetMaterials()                                  // new parameters, BVH kept
//
// The derivative holds Russian roulette and the power-weighted light choice fixed (they are
// constants of the parameters) and needs FireflyClamp <= 0 and the one-sample NEE integrator.
type DiffRender struct {
	scene   *Scene
	desc    C.bc_render_desc
	w, h    int
	forward bool // a Forward of the current parameters precedes Backward
}

func NewDiffRender(scene *Scene, cam Camera, settings RenderSettings) (*DiffRender, error) {
	if scene == nil || len(scene.Primitives) == 0 || len(scene.Nodes) == 0 {
		return nil, errors.New("HIP backend requires a non-empty built scene")
	}
	if settings.FireflyClamp > 0 {
		return nil, errors.New("differentiable render needs FireflyClamp <= 0 (the clamp has no derivative)")
	}
	d, w, h, _, _, _, _ := hipRenderDesc(cam, settings)
	return &DiffRender{scene: scene, desc: d, w: w, h: h}, nil
}

// Forward renders the image for the scene's current materials: linear radiance per pixel, row
// major, and the path-segment count.
func (f *DiffRender) Forward() ([]Vec3, int64, error) {
	rt := &persistentHIP
	if err := rt.lockedContext(); err != nil {
		return nil, 0, err
	}
	defer rt.mu.Unlock()
	if err := rt.ensureSceneLocked(f.scene); err != nil {
		return nil, 0, err
	}
	n := f.w * f.h
	raw := (*C.float)(C.malloc(C.size_t(16 * n)))
	defer C.free(unsafe.Pointer(raw))
	var rays C.uint64_t
	if C.bc_forward(rt.ctx, &f.desc, 0, C.int(f.h), raw, &rays) != 0 {
		return nil, 0, hipErr(rt.ctx)
	}
	px := unsafe.Slice((*float32)(unsafe.Pointer(raw)), 4*n)
	img := make([]Vec3, n)
	for i := range img {
		img[i] = V(float64(px[4*i]), float64(px[4*i+1]), float64(px[4*i+2]))
	}
	f.forward = true
	return img, int64(rays), nil
}

// Backward returns d loss / d Material.Color and d loss / d Material.Emission for every
// material of the scene, given dLdI[pixel] = d loss / d (pixel radiance of the last Forward).
func (f *DiffRender) Backward(dLdI []Vec3) (dColor, dEmission []Vec3, err error) {
	if !f.forward {
		return nil, nil, errors.New("backward needs a forward of the current parameters first")
	}
	n := f.w * f.h
	if len(dLdI) != n {
		return nil, nil, errors.New("dLdI needs one entry per pixel")
	}
	rt := &persistentHIP
	if err := rt.lockedContext(); err != nil {
		return nil, nil, err
	}
	defer rt.mu.Unlock()
	nm := len(f.scene.Materials)
	grad := (*C.float)(C.malloc(C.size_t(12 * n)))
	defer C.free(unsafe.Pointer(grad))
	g := unsafe.Slice((*float32)(unsafe.Pointer(grad)), 3*n)
	for i, v := range dLdI {
		g[3*i], g[3*i+1], g[3*i+2] = float32(v.X), float32(v.Y), float32(v.Z)
	}
	// bc_backward accumulates into its outputs, so they start at zero
	dc := (*C.float)(C.calloc(C.size_t(3*nm+1), 4))
	defer C.free(unsafe.Pointer(dc))
	de := (*C.float)(C.calloc(C.size_t(3*nm+1), 4))
	defer C.free(unsafe.Pointer(de))
	if C.bc_backward(rt.ctx, &f.desc, 0, C.int(f.h), grad, dc, de) != 0 {
		return nil, nil, hipErr(rt.ctx)
	}
	gc := unsafe.Slice((*float32)(unsafe.Pointer(dc)), 3*nm)
	ge := unsafe.Slice((*float32)(unsafe.Pointer(de)), 3*nm)
	dColor, dEmission = make([]Vec3, nm), make([]Vec3, nm)
	for i := 0; i < nm; i++ {
		dColor[i] = V(float64(gc[3*i]), float64(gc[3*i+1]), float64(gc[3*i+2]))
		dEmission[i] = V(float64(ge[3*i]), float64(ge[3*i+1]), float64(ge[3*i+2]))
	}
	return dColor, dEmission, nil
}

// SetMaterials sends the scene's current Materials (new colours or emission, same count and the
// same emissive materials) to the device without rebuilding the BVH; the next Backward needs a
// new Forward.
func (f *DiffRender) SetMaterials() error {
	rt := &persistentHIP
	if err := rt.lockedContext(); err != nil {
		return err
	}
	defer rt.mu.Unlock()
	f.forward = false
	if rt.scene != f.scene {
		return rt.ensureSceneLocked(f.scene) // a full upload carries the materials
	}
	nm := len(f.scene.Materials)
	kinds := (*C.int32_t)(C.malloc(C.size_t(4*nm + 4)))
	defer C.free(unsafe.Pointer(kinds))
	data := (*C.double)(C.malloc(C.size_t(64*nm + 8)))
	defer C.free(unsafe.Pointer(data))
	fillMaterials(unsafe.Slice(kinds, nm), unsafe.Slice(data, 8*nm), f.scene.Materials)
	if C.bc_update_materials(rt.ctx, C.int32_t(nm), kinds, data) != 0 {
		return hipErr(rt.ctx)
	}
	return nil
}
