// gi_device.hpp — device-side arithmetic shared by the gfx950 kernels.
//
// Numerics contract: every kernel is compiled with -ffp-contract=off, so each
// a*b+c below is a rounded multiply then a rounded add, in the order written;
// FP32 division and sqrt are IEEE correctly rounded (the HIP default).  GLSL
// built-ins are implemented with their GLSL-spec definitions so results are
// reproducible on any IEEE machine:
//   mix(x,y,a) = x*(1-a) + y*a     max(x,y) = x < y ? y : x
//   min(x,y)   = y < x ? y : x     clamp(x,a,b) = min(max(x,a),b)
//   fract(x)   = x - floor(x)      normalize(v) = v * (1/sqrt(dot(v,v)))
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gi {

__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gclamp(float x, float a, float b) { return gmin(gmax(x, a), b); }
__device__ __forceinline__ float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float gfract(float x) { return x - floorf(x); }
__device__ __forceinline__ float le01(float z, float d) { return (z <= d) ? 1.0f : 0.0f; }

// Shadow-map texel at (s, t): GL_NEAREST, CLAMP_TO_BORDER with border depth 0
// (ShadowMapping/src/Viewers/MyGLTextureViewer.cpp:3-28).  Out-of-range
// coordinates (and NaN) read the border, i.e. count as shadowed.
struct DepthMap {
    const float* __restrict__ texels;
    int w, h;
    float fw, fh;  // (float)w, (float)h

    __device__ __forceinline__ float fetch(float s, float t) const {
        float u = s * fw;
        float v = t * fh;
        if (!(u >= 0.0f && u < fw && v >= 0.0f && v < fh)) return 0.0f;
        int i = (int)floorf(u);
        int j = (int)floorf(v);
        return __ldg(texels + (size_t)j * (size_t)w + (size_t)i);
    }
};

// M * (x, y, z, w) with GLSL column-major order, summed left to right.
__device__ __forceinline__ void mat4_vec4(const float* m, float x, float y, float z, float w,
                                          float o[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float acc = m[0 * 4 + i] * x;
        acc = acc + m[1 * 4 + i] * y;
        acc = acc + m[2 * 4 + i] * z;
        acc = acc + m[3 * 4 + i] * w;
        o[i] = acc;
    }
}

__device__ __forceinline__ void normalize3(const float v[3], float o[3]) {
    float sqr = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float inv = 1.0f / sqrtf(sqr);
    o[0] = v[0] * inv;
    o[1] = v[1] * inv;
    o[2] = v[2] * inv;
}

// Kernel-argument copies of small matrices (passed by value).
struct Mat4 { float m[16]; };
struct Mat3 { float m[9]; };

}  // namespace gi
