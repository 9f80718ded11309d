//go:build !(linux && cgo && beamcast_hip)

// gpu_transport_opencl.go -- the GPU gate's capability flag where the OpenCL kernel
// (gpu_opencl_windows.go) or the stub (gpu_opencl_stub.go) provides renderOpenCL: the
// kernel traces uniform random samples with one-sample light selection and no clamp, so
// RenderState.Start keeps sending MIS, power-weighted, Halton/R2 and clamped renders to the
// CPU there.  Copied into windows_portable/ with gpu_hip_linux.go.
package main

const gpuFullTransport = false
