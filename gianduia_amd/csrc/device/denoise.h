// denoise.h — launcher of the film denoiser kernels (denoise.hip).
#pragma once
#include <hip/hip_runtime.h>

// a-trous passes (steps 1, 2, 4, 8, 16)
constexpr int kDenoisePasses = 5;

// scratch: 4 * width * height float4 (two radiance buffers, two guide
// buffers); color / albedo / normal / out: width * height * 3 floats, albedo
// and normal may be null, out may be color.  passBegin / passEnd (optional)
// bracket the first a-trous pass.
hipError_t gndDenoiseLaunch(hipStream_t stream, const float* color, const float* albedo, const float* normal, int W,
                            int H, float* out, float4* scratch, hipEvent_t passBegin, hipEvent_t passEnd);
