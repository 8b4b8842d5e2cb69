// diff.hip — differentiable area-light soft shadows for gfx950.
//
// No reference counterpart: the reference's shadow tests are step functions.
// The continuous model keeps the structure of AccurateSoftShadow.frag (one
// depth map per light sample, the per-pixel mean over samples) and makes each
// sample's visibility smooth:
//   texel depth d  = window depth of the plane of the triangle that covers
//                    the texel centre (visibility raster, IDs held fixed),
//                    through the triangle's unsnapped projected vertices;
//   q              = sum over the 4 texels around (s*W - .5, t*H - .5) of
//                    smoothstep-bilinear weights * sigmoid((d - z + bias)/tau);
//   V_i            = si + (1 - si) * q;   S = mean_i V_i.
//
// A texel's depth and barycentrics are affine in the texel centre, so the
// per-triangle part is set up once per (map, triangle) (k_diff_setup) and a
// texel evaluates two affine forms.  The adjoint of the texel depths with
// respect to the triangle's clip-space vertices is linear in G * lambda_j
// (G = dL/dd, lambda_j the texel's barycentrics), so the texel adjoint only
// sums w_j = sum G * lambda_j per (map, triangle) and k_diff_tri_bwd turns the
// three sums into vertex and matrix gradients.
//
// Forward:  visibility raster, k_diff_setup, k_diff_depth, k_diff_shade.
// Backward: the visibility raster and k_diff_setup again (recomputed, nothing
//           is kept from the forward pass), then k_diff_shade_bwd (pixel
//           adjoint, tap depths from keys and planes: dL/d(B*M'_i) per block,
//           dL/dd per texel by scatter-add), k_diff_texel_bwd,
//           k_diff_tri_bwd, and k_diff_reduce (fixed-order reduction of the
//           per-block matrix partials).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"
#include "gi_internal.hpp"

namespace gi {
namespace {

constexpr int kThreads = 256;
constexpr int kTexPerThread = 4;   // texels per thread in the texel adjoint
constexpr int kGatherMax = 256;    // bbox texels up to which a triangle gathers its own adjoint
constexpr int kLargeSlots = 4;     // large triangles a texel block sums in LDS
constexpr int kMapsPerThread = 1;  // maps a triangle-adjoint thread walks (vertex gradients summed over them)
constexpr int kGatherBatch = 8;    // bounding-box texels whose keys a gather loads together
constexpr int kBwdPixPerThread = 4;  // pixels a pixel-adjoint thread sums before the block reduction

// Per-block sum of one large triangle's G * lambda_j (tri = -1: unused).
struct LargePart {
    double w[3];
    int tri;
    int pad;
};

// Geometry (projections, barycentrics, window depths, the sigmoid argument)
// is evaluated in FP64: the argument (d - z + bias) / tau divides the
// rounding of two window depths near 0.9 by tau ~ 2e-4, so FP32 geometry
// would leave ~1e-3 noise in every term.  Texel depths and the per-triangle
// sums are FP64; the texel-depth adjoint and the vertex gradients are FP32.
struct DiffCfg {
    int smw, smh, S;
    int gw, gh;
    double tau, bias, si;
};

// Affine form of one (map, triangle): with dx = x - X0, dy = y - Y0,
//   l1 = a1*dx + b1*dy,  l2 = a2*dx + b2*dy,  d = Z0 + l1*dZ1 + l2*dZ2.
// X0 = NaN marks a triangle without a plane (a vertex with w <= 0, or a
// degenerate projection): its texels read depth 1 and get no gradient.
// [bx0, bx1] x [by0, by1]: texels that can hold the triangle: texel x is
// covered only if x + .5 lies in the triangle's snapped window extent, and
// snapping moves a vertex by at most 1/512 px (bound taken as 1/256), clamped
// to the map; empty when bx0 > bx1.
struct TriPlane {
    double X0, Y0;
    double a1, b1, a2, b2;
    double Z0, dZ1, dZ2;
    short bx0, bx1, by0, by1;
};

__device__ __forceinline__ bool plane_ok(const TriPlane& p) { return p.X0 == p.X0; }
__device__ __forceinline__ int bbox_area(const TriPlane& p) {
    return (p.bx1 < p.bx0 || p.by1 < p.by0) ? 0 : (p.bx1 - p.bx0 + 1) * (p.by1 - p.by0 + 1);
}
// Triangles whose adjoint the texel pass scatters (k_diff_texel_bwd); the
// others gather theirs in k_diff_tri_bwd.
__device__ __forceinline__ bool plane_large(const TriPlane& p) { return bbox_area(p) > kGatherMax; }

__device__ __forceinline__ double sigmoid_d(double x) { return 1.0 / (1.0 + exp(-x)); }
__device__ __forceinline__ double smoothw(double f) { return f * f * (3.0 - 2.0 * f); }
__device__ __forceinline__ double dsmoothw(double f) { return 6.0 * f * (1.0 - f); }

__device__ __forceinline__ void load_m(const double* __restrict__ src, double m[16]) {
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = src[i];
}

// FP32 copies of the FP64 matrices for the visibility raster (round to
// nearest even, as a host-side float cast).
__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = __double2float_rn(src[i]);
}

__device__ __forceinline__ void mat4_vec4_d(const double m[16], double x, double y, double z, double w,
                                            double out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

// Window coordinates of a triangle's unsnapped projected vertices
// (oracle: or_diff_depth).  False when a vertex has w <= 0.
__device__ __forceinline__ bool tri_screen(const float* __restrict__ pos, const int* __restrict__ idx, int tri,
                                           const double m[16], int w, int h, double X[3], double Y[3], double Z[3],
                                           double C[3][4]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        int v = idx[3 * tri + j];
        mat4_vec4_d(m, pos[3 * v], pos[3 * v + 1], pos[3 * v + 2], 1.0, C[j]);
        ok = ok && (C[j][3] > 0.0);
        X[j] = (C[j][0] / C[j][3]) * (0.5 * w) + 0.5 * w;
        Y[j] = (C[j][1] / C[j][3]) * (0.5 * h) + 0.5 * h;
        Z[j] = (C[j][2] / C[j][3]) * 0.5 + 0.5;
    }
    return ok;
}

// One thread per (triangle, map).  grid (triangle blocks, S).
__global__ __launch_bounds__(kThreads) void k_diff_setup(const float* __restrict__ pos, const int* __restrict__ idx,
                                                         int ntris, const double* __restrict__ mvps, DiffCfg c,
                                                         TriPlane* __restrict__ planes,
                                                         unsigned char* __restrict__ large) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    const int i = blockIdx.y;
    if (t >= ntris) return;
    double m[16];
    load_m(mvps + 16 * i, m);
    double X[3], Y[3], Z[3], C[3][4];
    TriPlane p;
    const bool ok = tri_screen(pos, idx, t, m, c.smw, c.smh, X, Y, Z, C);
    const double D = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
    p.bx0 = p.by0 = 1;
    p.bx1 = p.by1 = 0;
    if (ok && D != 0.0) {
        const double e = 0.5 + 1.0 / 256.0;
        const double x0 = ceil(fmin(X[0], fmin(X[1], X[2])) - e), x1 = floor(fmax(X[0], fmax(X[1], X[2])) - 0.5 + 1.0 / 256.0);
        const double y0 = ceil(fmin(Y[0], fmin(Y[1], Y[2])) - e), y1 = floor(fmax(Y[0], fmax(Y[1], Y[2])) - 0.5 + 1.0 / 256.0);
        p.bx0 = (short)fmax(0.0, fmin(x0, (double)c.smw));
        p.bx1 = (short)fmax(-1.0, fmin(x1, (double)(c.smw - 1)));
        p.by0 = (short)fmax(0.0, fmin(y0, (double)c.smh));
        p.by1 = (short)fmax(-1.0, fmin(y1, (double)(c.smh - 1)));
        p.X0 = X[0];
        p.Y0 = Y[0];
        p.a1 = (Y[2] - Y[0]) / D;
        p.b1 = -(X[2] - X[0]) / D;
        p.a2 = -(Y[1] - Y[0]) / D;
        p.b2 = (X[1] - X[0]) / D;
        p.Z0 = Z[0];
        p.dZ1 = Z[1] - Z[0];
        p.dZ2 = Z[2] - Z[0];
    } else {
        p.X0 = __builtin_nan("");
        p.Y0 = p.a1 = p.b1 = p.a2 = p.b2 = p.Z0 = p.dZ1 = p.dZ2 = 0.0;
    }
    planes[(size_t)i * ntris + t] = p;
    large[(size_t)i * ntris + t] = (plane_ok(p) && plane_large(p)) ? 1 : 0;
}

// Triangle of a visibility key (a piece id is tri * 8 + piece), or -1.
__device__ __forceinline__ int key_tri(unsigned long long key, int ntris) {
    if (key == ~0ull) return -1;
    const int t = (int)((key & 0xffffffffull) >> 3);
    return t < ntris ? t : -1;
}

// Depth of texel (kx, ky) of one map from its visibility key and the plane
// record of the triangle the key names; 1 for background.
__device__ __forceinline__ double texel_depth(const unsigned long long* __restrict__ vk,
                                              const TriPlane* __restrict__ pl, int ntris, uint32_t smw, uint32_t kx,
                                              uint32_t ky) {
    const int t = key_tri(vk[(size_t)ky * smw + kx], ntris);
    double d = 1.0;
    if (t >= 0) {
        const TriPlane& p = pl[t];
        if (p.X0 == p.X0) {
            const double dx = ((double)kx + 0.5) - p.X0, dy = ((double)ky + 0.5) - p.Y0;
            const double l1 = p.a1 * dx + p.b1 * dy, l2 = p.a2 * dx + p.b2 * dy;
            d = p.Z0 + l1 * p.dZ1 + l2 * p.dZ2;
        }
    }
    return d;
}

__global__ __launch_bounds__(kThreads) void k_diff_depth(const unsigned long long* __restrict__ vis,
                                                         const TriPlane* __restrict__ planes, int ntris, DiffCfg c,
                                                         double* __restrict__ depth) {
    const uint32_t plane = (uint32_t)c.smw * (uint32_t)c.smh;  // < 2^31 (edges <= 16384)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const int map = blockIdx.y;
    if (k >= plane) return;
    const uint32_t ky = k / (uint32_t)c.smw, kx = k - ky * (uint32_t)c.smw;
    depth[(size_t)map * plane + k] =
        texel_depth(vis + (size_t)map * plane, planes + (size_t)map * ntris, ntris, (uint32_t)c.smw, kx, ky);
}

struct Tap {
    double z, cw, cx, cy, cz;
    int a0, b0;
    double fu, fv;
};

__device__ __forceinline__ Tap project(const double m[16], double x, double y, double z, const DiffCfg& c) {
    double cc[4];
    mat4_vec4_d(m, x, y, z, 1.0, cc);
    Tap t;
    t.cx = cc[0]; t.cy = cc[1]; t.cz = cc[2]; t.cw = cc[3];
    const double s = cc[0] / cc[3], tt = cc[1] / cc[3];
    t.z = cc[2] / cc[3];
    const double u = s * (double)c.smw - 0.5, v = tt * (double)c.smh - 0.5;
    const double fa = floor(u), fb = floor(v);
    // clamp before the int conversion; the texel reads are bounds-checked
    t.a0 = (int)fmin(fmax(fa, -4.0), (double)c.smw + 4.0);
    t.b0 = (int)fmin(fmax(fb, -4.0), (double)c.smh + 4.0);
    t.fu = u - fa;
    t.fv = v - fb;
    return t;
}

__device__ __forceinline__ double texel(const double* __restrict__ d, int a, int b, const DiffCfg& c) {
    return (a >= 0 && a < c.smw && b >= 0 && b < c.smh) ? d[(size_t)b * c.smw + a] : 0.0;
}

__global__ __launch_bounds__(kThreads) void k_diff_shade(const float* __restrict__ gpos,
                                                         const double* __restrict__ biased,
                                                         const double* __restrict__ depth, DiffCfg c,
                                                         float* __restrict__ out) {
    const size_t plane = (size_t)c.gw * c.gh;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= plane) return;
    const float xf = gpos[k];
    if (xf == 0.0f) {
        out[k] = 0.0f;
        return;
    }
    const double x = xf, y = gpos[plane + k], z = gpos[2 * plane + k];
    const size_t msz = (size_t)c.smw * c.smh;
    double acc = 0.0;
    for (int i = 0; i < c.S; ++i) {
        double m[16];
        load_m(biased + 16 * i, m);
        const Tap t = project(m, x, y, z, c);
        const double* d = depth + (size_t)i * msz;
        const double wu[2] = {1.0 - smoothw(t.fu), smoothw(t.fu)};
        const double wv[2] = {1.0 - smoothw(t.fv), smoothw(t.fv)};
        double q = 0.0;
#pragma unroll
        for (int db = 0; db < 2; ++db)
#pragma unroll
            for (int da = 0; da < 2; ++da)
                q += wu[da] * wv[db] * sigmoid_d((texel(d, t.a0 + da, t.b0 + db, c) - t.z + c.bias) / c.tau);
        acc += c.si + (1.0 - c.si) * q;
    }
    out[k] = (float)(acc / (double)c.S);
}

// Block sum of 16 per-thread values into partial[16] (fixed order).
__device__ __forceinline__ void block_sum16(const double v[16], double* __restrict__ partial) {
    __shared__ double red[kThreads / 64][16];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        double x = v[e];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[wave][e] = x;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
        partial[threadIdx.x] = s;
    }
    __syncthreads();
}

// Pixel adjoint.  grid (pixel blocks of kBwdPixPerThread * kThreads pixels, S):
// each thread adds the matrix terms of kBwdPixPerThread pixels (strided by
// the block size, so every step reads coalesced) before the one block
// reduction.  The tap depths are evaluated from the recomputed visibility
// keys and plane records (texel_depth, the arithmetic k_diff_depth stores in
// the forward pass), so backward never materialises the depth maps.
__global__ __launch_bounds__(kThreads) void k_diff_shade_bwd(const float* __restrict__ gpos,
                                                             const double* __restrict__ biased,
                                                             const unsigned long long* __restrict__ vis,
                                                             const TriPlane* __restrict__ planes, int ntris,
                                                             const float* __restrict__ grad_out, DiffCfg c,
                                                             float* __restrict__ grad_depth,
                                                             double* __restrict__ partial) {
    const size_t plane = (size_t)c.gw * c.gh;
    const int i = blockIdx.y;
    double gm[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) gm[e] = 0.0;
    double m[16];
    load_m(biased + 16 * i, m);
    for (int p = 0; p < kBwdPixPerThread; ++p) {
        const size_t k = ((size_t)blockIdx.x * kBwdPixPerThread + p) * blockDim.x + threadIdx.x;
        if (k < plane && gpos[k] != 0.0f) {
            const double x = gpos[k], y = gpos[plane + k], z = gpos[2 * plane + k];
            const double gq = (double)grad_out[k] / (double)c.S * (1.0 - c.si);
            const Tap t = project(m, x, y, z, c);
            const size_t msz = (size_t)c.smw * c.smh;
            const unsigned long long* vk = vis + (size_t)i * msz;
            const TriPlane* pl = planes + (size_t)i * ntris;
            const double su = smoothw(t.fu), sv = smoothw(t.fv);
            const double wu[2] = {1.0 - su, su}, wv[2] = {1.0 - sv, sv};
            const double dwu[2] = {-dsmoothw(t.fu), dsmoothw(t.fu)}, dwv[2] = {-dsmoothw(t.fv), dsmoothw(t.fv)};
            double gz = 0.0, gfu = 0.0, gfv = 0.0;
#pragma unroll
            for (int db = 0; db < 2; ++db)
#pragma unroll
                for (int da = 0; da < 2; ++da) {
                    const int a = t.a0 + da, b = t.b0 + db;
                    const bool in = a >= 0 && a < c.smw && b >= 0 && b < c.smh;
                    const double dd = in ? texel_depth(vk, pl, ntris, (uint32_t)c.smw, (uint32_t)a, (uint32_t)b) : 0.0;
                    const double sg = sigmoid_d((dd - t.z + c.bias) / c.tau);
                    const double ds = sg * (1.0 - sg) / c.tau;
                    const double w = wu[da] * wv[db];
                    if (in && gq != 0.0)
                        unsafeAtomicAdd(&grad_depth[(size_t)i * msz + (size_t)b * c.smw + a], (float)(gq * w * ds));
                    gz -= w * ds;
                    gfu += dwu[da] * wv[db] * sg;
                    gfv += wu[da] * dwv[db] * sg;
                }
            const double gs = gq * gfu * (double)c.smw, gt = gq * gfv * (double)c.smh;
            gz *= gq;
            const double iw = 1.0 / t.cw;
            const double gc[4] = {gs * iw, gt * iw, gz * iw, -(gs * t.cx + gt * t.cy + gz * t.cz) * iw * iw};
            const double xp[4] = {x, y, z, 1.0};
#pragma unroll
            for (int col = 0; col < 4; ++col)
#pragma unroll
                for (int row = 0; row < 4; ++row) gm[col * 4 + row] += gc[row] * xp[col];
        }
    }
    block_sum16(gm, partial + ((size_t)i * gridDim.x + blockIdx.x) * 16);
}

// Texel adjoint of the large triangles (bounding box over kGatherMax
// texels; a ground plane covers most texels of every map): the sums of
// G * lambda_j over each large triangle's texels.  grid (texel blocks, S);
// each thread takes kTexPerThread texels strided by the block size
// (coalesced).  Device-scope atomics on one triangle's cache line from every
// wave would serialize, so a wave first reduces: its texels are 64
// consecutive texels of a row, in which each triangle covers a contiguous
// run, and a segmented scan leaves every run's sum in the run's last lane;
// a wave that is one run keeps a running sum in lane 0 while consecutive rows
// stay on that triangle.  Runs are added into kLargeSlots LDS slots per
// block, written out as the block's LargePart entries that
// k_diff_large_reduce adds into wsum; a block meeting more large triangles
// than slots adds the rest to wsum with device atomics.
template <int kSlots>
__device__ __forceinline__ void add_slots(int* slot_tri, double (*slot_w)[3], double* __restrict__ wsum, int tri,
                                          const double w[3]) {
    if (tri < 0) return;
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const int old = atomicCAS(&slot_tri[q], -1, tri);
        if (old == -1 || old == tri) {
#pragma unroll
            for (int j = 0; j < 3; ++j) atomicAdd(&slot_w[q][j], w[j]);
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) unsafeAtomicAdd(&wsum[3 * (size_t)tri + j], w[j]);
}

__global__ __launch_bounds__(kThreads) void k_diff_texel_bwd(const unsigned long long* __restrict__ vis,
                                                             const TriPlane* __restrict__ planes,
                                                             const unsigned char* __restrict__ large, int ntris,
                                                             const float* __restrict__ grad_depth, DiffCfg c,
                                                             double* __restrict__ wsum_all,
                                                             LargePart* __restrict__ parts,
                                                             unsigned char* __restrict__ touched) {
    __shared__ int slot_tri[kLargeSlots];
    __shared__ double slot_w[kLargeSlots][3];
    if (threadIdx.x < kLargeSlots) {
        slot_tri[threadIdx.x] = -1;
        slot_w[threadIdx.x][0] = slot_w[threadIdx.x][1] = slot_w[threadIdx.x][2] = 0.0;
    }
    __syncthreads();
    const uint32_t msz = (uint32_t)c.smw * (uint32_t)c.smh;  // < 2^31 (edges <= 16384)
    const int i = blockIdx.y;
    const float* __restrict__ gd = grad_depth + (size_t)i * msz;
    const unsigned long long* __restrict__ vk = vis + (size_t)i * msz;
    double* __restrict__ wsum = wsum_all + (size_t)i * ntris * 3;
    const TriPlane* __restrict__ pl = planes + (size_t)i * ntris;
    const unsigned char* __restrict__ lg = large + (size_t)i * ntris;
    unsigned char* __restrict__ tch = touched + (size_t)i * ntris;
    const int lane = threadIdx.x % 64;
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const uint32_t base = (uint32_t)blockIdx.x * (kTexPerThread * kThreads) + threadIdx.x;
    int cur = -1;  // lane 0: triangle of the running sum
    double run[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < kTexPerThread; ++r) {
        const uint32_t k = base + (uint32_t)r * kThreads;
        double w[3] = {0.0, 0.0, 0.0};
        int tri = -1;
        const float G = (k < msz) ? gd[k] : 0.0f;
        if (G != 0.0f) {
            const int t = key_tri(vk[k], ntris);
            if (t >= 0 && !lg[t]) tch[t] = 1;  // a small triangle with gradient: k_diff_tri_bwd gathers it
            if (t >= 0 && lg[t]) {
                const TriPlane& p = pl[t];
                {
                    tri = t;
                    const uint32_t ky = k / (uint32_t)c.smw, kx = k - ky * (uint32_t)c.smw;
                    const double dx = ((double)kx + 0.5) - p.X0, dy = ((double)ky + 0.5) - p.Y0;
                    const double l1 = p.a1 * dx + p.b1 * dy, l2 = p.a2 * dx + p.b2 * dy;
                    w[0] = (double)G * (1.0 - l1 - l2);
                    w[1] = (double)G * l1;
                    w[2] = (double)G * l2;
                }
            }
        }
        // every lane of the wave reaches this point
        const unsigned long long valid = __ballot(tri >= 0);
        if (valid == 0ull) continue;
        // lanes without a contribution (w = 0) join the run of the nearest
        // contributing lane at or below them
        const unsigned long long below = valid & upto;
        const int src = below ? 63 - __clzll((long long)below) : __ffsll((long long)valid) - 1;
        tri = __shfl(tri, src, 64);
        const int prev = __shfl_up(tri, 1, 64);
        const unsigned long long heads = __ballot(lane == 0 || tri != prev);
        if (heads == 1ull) {  // one run: the whole wave is one triangle
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double x = w[j];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
                w[j] = x;
            }
            if (lane == 0) {
                if (cur != tri) {
                    add_slots<kLargeSlots>(slot_tri, slot_w, wsum, cur, run);
                    cur = tri;
                    run[0] = run[1] = run[2] = 0.0;
                }
                run[0] += w[0];
                run[1] += w[1];
                run[2] += w[2];
            }
            continue;
        }
        // segmented inclusive scan: lane sums its run from the run's head
        const int head = 63 - __clzll((long long)(heads & upto));
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double y = __shfl_up(w[j], off, 64);
                if (lane - off >= head) w[j] += y;
            }
        }
        const bool tail = (lane == 63) || ((heads >> (lane + 1)) & 1ull);
        if (tail) add_slots<kLargeSlots>(slot_tri, slot_w, wsum, tri, w);
    }
    if (lane == 0) add_slots<kLargeSlots>(slot_tri, slot_w, wsum, cur, run);
    __syncthreads();
    if (threadIdx.x < kLargeSlots) {
        LargePart lp;
        lp.tri = slot_tri[threadIdx.x];
        lp.pad = 0;
        lp.w[0] = slot_w[threadIdx.x][0];
        lp.w[1] = slot_w[threadIdx.x][1];
        lp.w[2] = slot_w[threadIdx.x][2];
        parts[((size_t)i * gridDim.x + blockIdx.x) * kLargeSlots + threadIdx.x] = lp;
    }
}

// Adds every block's LargePart entries of map blockIdx.x into wsum: the
// entries first meet in LDS slots (few large triangles per map), then one
// device atomic per slot.  grid (S).
constexpr int kMapSlots = 32;

__global__ __launch_bounds__(kThreads) void k_diff_large_reduce(const LargePart* __restrict__ parts, int tx_blocks,
                                                                int ntris, double* __restrict__ wsum_all) {
    __shared__ int slot_tri[kMapSlots];
    __shared__ double slot_w[kMapSlots][3];
    if (threadIdx.x < kMapSlots) {
        slot_tri[threadIdx.x] = -1;
        slot_w[threadIdx.x][0] = slot_w[threadIdx.x][1] = slot_w[threadIdx.x][2] = 0.0;
    }
    __syncthreads();
    const int i = blockIdx.x;
    double* __restrict__ wsum = wsum_all + (size_t)i * ntris * 3;
    const LargePart* __restrict__ bp = parts + (size_t)i * tx_blocks * kLargeSlots;
    for (int b = threadIdx.x; b < tx_blocks * kLargeSlots; b += kThreads) {
        const LargePart lp = bp[b];
        if (lp.tri >= 0) add_slots<kMapSlots>(slot_tri, slot_w, wsum, lp.tri, lp.w);
    }
    __syncthreads();
    if (threadIdx.x < kMapSlots && slot_tri[threadIdx.x] >= 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            unsafeAtomicAdd(&wsum[3 * (size_t)slot_tri[threadIdx.x] + j], slot_w[threadIdx.x][j]);
    }
}

// Triangle adjoint.  grid (triangle blocks, S / kMapsPerThread): a thread
// walks kMapsPerThread maps of one triangle.  w_j = sum of G * lambda_j
// over the triangle's texels in map i: gathered here over the triangle's
// bounding box (texels whose visibility key names it) for small triangles,
// read from the texel pass's sums for large ones.  The texel depths then move
// by
//   dL = sum_j [ w_j/2 dz_j - w_j (dd/dX) dX_j - w_j (dd/dY) dY_j ]
// (a texel's depth through the plane moves by -lambda_j * grad when vertex j
// moves in the window, and by lambda_j times vertex j's window depth change),
// chained through the perspective divide to vertex j's clip coordinates, then
// to dL/dM'_i (block partials) and dL/dvertex (scatter-add).
__global__ __launch_bounds__(kThreads) void k_diff_tri_bwd(const float* __restrict__ pos, const int* __restrict__ idx,
                                                           int ntris, const double* __restrict__ mvps,
                                                           const TriPlane* __restrict__ planes,
                                                           const unsigned long long* __restrict__ vis,
                                                           const float* __restrict__ grad_depth,
                                                           const double* __restrict__ wsum_all,
                                                           const unsigned char* __restrict__ touched, DiffCfg c,
                                                           float* __restrict__ grad_pos,
                                                           double* __restrict__ partial) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    // vertex gradients summed over this thread's maps, added once at the end
    double gvs[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) gvs[j][0] = gvs[j][1] = gvs[j][2] = 0.0;
    const int i0 = blockIdx.y * kMapsPerThread;
    const int i1 = min(i0 + kMapsPerThread, c.S);
    for (int i = i0; i < i1; ++i) {
        double gm[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) gm[e] = 0.0;
        double w[3] = {0.0, 0.0, 0.0};
        if (t < ntris) {
            const TriPlane p = planes[(size_t)i * ntris + t];
            if (plane_ok(p)) {
                if (plane_large(p)) {
                    const double* ws = wsum_all + ((size_t)i * ntris + t) * 3;
                    w[0] = ws[0];
                    w[1] = ws[1];
                    w[2] = ws[2];
                } else if (touched[(size_t)i * ntris + t]) {
                    // bounding-box texels in batches whose loads are issued together
                    const size_t moff = (size_t)i * c.smw * c.smh;
                    const int bw = p.bx1 - p.bx0 + 1, area = bbox_area(p);
                    for (int q0 = 0; q0 < area; q0 += kGatherBatch) {
                        unsigned long long key[kGatherBatch];
                        float G[kGatherBatch];
#pragma unroll
                        for (int u = 0; u < kGatherBatch; ++u) {
                            const int q = q0 + u;
                            const size_t k = moff + (size_t)(p.by0 + q / bw) * c.smw + (p.bx0 + q % bw);
                            key[u] = q < area ? vis[k] : ~0ull;
                            G[u] = q < area ? grad_depth[k] : 0.0f;
                        }
#pragma unroll
                        for (int u = 0; u < kGatherBatch; ++u) {
                            if (G[u] == 0.0f || key_tri(key[u], ntris) != t) continue;
                            const int q = q0 + u;
                            const double dx = ((double)(p.bx0 + q % bw) + 0.5) - p.X0;
                            const double dy = ((double)(p.by0 + q / bw) + 0.5) - p.Y0;
                            const double l1 = p.a1 * dx + p.b1 * dy, l2 = p.a2 * dx + p.b2 * dy;
                            w[0] += (double)G[u] * (1.0 - l1 - l2);
                            w[1] += (double)G[u] * l1;
                            w[2] += (double)G[u] * l2;
                        }
                    }
                }
            }
        }
        if (w[0] != 0.0 || w[1] != 0.0 || w[2] != 0.0) {
            double m[16];
            load_m(mvps + 16 * i, m);
            double X[3], Y[3], Z[3], C[3][4];
            const bool ok = tri_screen(pos, idx, t, m, c.smw, c.smh, X, Y, Z, C);
            const double D = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
            if (ok && D != 0.0) {
                const double gx = ((Z[1] - Z[0]) * (Y[2] - Y[0]) - (Z[2] - Z[0]) * (Y[1] - Y[0])) / D;
                const double gy = ((X[1] - X[0]) * (Z[2] - Z[0]) - (X[2] - X[0]) * (Z[1] - Z[0])) / D;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double gX = -w[j] * gx * (0.5 * c.smw);
                    const double gY = -w[j] * gy * (0.5 * c.smh);
                    const double gZ = w[j] * 0.5;
                    const double iw = 1.0 / C[j][3];
                    const double gc[4] = {gX * iw, gY * iw, gZ * iw,
                                          -(gX * C[j][0] + gY * C[j][1] + gZ * C[j][2]) * iw * iw};
                    const int v = idx[3 * t + j];
                    const double vp[4] = {pos[3 * v], pos[3 * v + 1], pos[3 * v + 2], 1.0};
#pragma unroll
                    for (int col = 0; col < 4; ++col)
#pragma unroll
                        for (int row = 0; row < 4; ++row) gm[col * 4 + row] += gc[row] * vp[col];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        double g = 0.0;
#pragma unroll
                        for (int row = 0; row < 4; ++row) g += m[a * 4 + row] * gc[row];
                        gvs[j][a] += g;
                    }
                }
            }
        }
        block_sum16(gm, partial + ((size_t)i * gridDim.x + blockIdx.x) * 16);
    }
    if (t < ntris && (gvs[0][0] != 0.0 || gvs[0][1] != 0.0 || gvs[0][2] != 0.0 || gvs[1][0] != 0.0 ||
                      gvs[1][1] != 0.0 || gvs[1][2] != 0.0 || gvs[2][0] != 0.0 || gvs[2][1] != 0.0 ||
                      gvs[2][2] != 0.0)) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int v = idx[3 * t + j];
#pragma unroll
            for (int a = 0; a < 3; ++a) unsafeAtomicAdd(&grad_pos[3 * v + a], (float)gvs[j][a]);
        }
    }
}

// out[i][e] = sum over blocks of partial[i][block][e]: one workgroup per
// sample, each thread sums a fixed stride of blocks, then a fixed-order block
// reduction (the result does not depend on scheduling).
__global__ __launch_bounds__(kThreads) void k_diff_reduce(const double* __restrict__ partial, int nblocks,
                                                          double* __restrict__ out) {
    __shared__ double res[16];
    const int i = blockIdx.x;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kThreads) {
        const double2* p = reinterpret_cast<const double2*>(partial + ((size_t)i * nblocks + b) * 16);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double2 v = p[q];
            acc[2 * q] += v.x;
            acc[2 * q + 1] += v.y;
        }
    }
    block_sum16(acc, res);
    if (threadIdx.x < 16) out[16 * (size_t)i + threadIdx.x] = res[threadIdx.x];
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double* p, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct DiffWs {
    float* mvps32;
    unsigned long long* vis;
    double* depth;
    float* grad_depth;
    TriPlane* planes;
    double* wsum;
    double* part_px;
    double* part_tri;
    LargePart* parts;
    unsigned char* large;
    unsigned char* touched;
    void* raster;
    size_t raster_bytes;
    int px_blocks, pxb_blocks, tri_blocks, tx_blocks;
};

size_t diff_layout(int ntris, int S, int smw, int smh, int gw, int gh, DiffWs* ws, char* base) {
    const size_t msz = (size_t)smw * smh;
    const int px_blocks = (int)(((size_t)gw * gh + kThreads - 1) / kThreads);
    const int pxb_blocks = (int)(((size_t)gw * gh + (size_t)kThreads * kBwdPixPerThread - 1) /
                                 ((size_t)kThreads * kBwdPixPerThread));
    const int tri_blocks = (ntris + kThreads - 1) / kThreads;
    const int tx_blocks = (int)((msz + (size_t)kThreads * kTexPerThread - 1) / ((size_t)kThreads * kTexPerThread));
    size_t off = 0;
    size_t o_m32 = off; off = align256(off + sizeof(float) * 16 * (size_t)S);
    size_t o_vis = off; off = align256(off + sizeof(unsigned long long) * msz * S);
    size_t o_depth = off; off = align256(off + sizeof(double) * msz * S);
    size_t o_gd = off; off = align256(off + sizeof(float) * msz * S);
    size_t o_pl = off; off = align256(off + sizeof(TriPlane) * (size_t)ntris * S);
    size_t o_ws = off; off = align256(off + sizeof(double) * 3 * (size_t)ntris * S);
    size_t o_ppx = off; off = align256(off + sizeof(double) * 16 * (size_t)px_blocks * S);
    size_t o_ptr = off; off = align256(off + sizeof(double) * 16 * (size_t)(tri_blocks > 0 ? tri_blocks : 1) * S);
    size_t o_lp = off; off = align256(off + sizeof(LargePart) * kLargeSlots * (size_t)tx_blocks * S);
    size_t o_lg = off; off = align256(off + (size_t)ntris * S);
    size_t o_tc = off; off = align256(off + (size_t)ntris * S);
    size_t rbytes = internal::vis_raster_workspace_size(ntris, S, smw, smh);
    size_t o_r = off; off = align256(off + rbytes);
    if (ws) {
        ws->mvps32 = reinterpret_cast<float*>(base + o_m32);
        ws->vis = reinterpret_cast<unsigned long long*>(base + o_vis);
        ws->depth = reinterpret_cast<double*>(base + o_depth);
        ws->grad_depth = reinterpret_cast<float*>(base + o_gd);
        ws->planes = reinterpret_cast<TriPlane*>(base + o_pl);
        ws->wsum = reinterpret_cast<double*>(base + o_ws);
        ws->part_px = reinterpret_cast<double*>(base + o_ppx);
        ws->part_tri = reinterpret_cast<double*>(base + o_ptr);
        ws->parts = reinterpret_cast<LargePart*>(base + o_lp);
        ws->large = reinterpret_cast<unsigned char*>(base + o_lg);
        ws->touched = reinterpret_cast<unsigned char*>(base + o_tc);
        ws->raster = base + o_r;
        ws->raster_bytes = rbytes;
        ws->px_blocks = px_blocks;
        ws->pxb_blocks = pxb_blocks;
        ws->tri_blocks = tri_blocks;
        ws->tx_blocks = tx_blocks;
    }
    return rbytes ? off : 0;
}

bool finite_all(const double* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

int check_common(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                 const double* mvps_dev, const double* biased_dev, int S, int smw, int smh, const gi_gbuffer* gb,
                 const gi_diff_params* p, void* ws, DiffCfg& c) {
    if (!positions_dev || (!indices_dev && ntris > 0) || !mvps_dev || !biased_dev || !gb || !gb->position_dev || !p ||
        !ws)
        return GI_EINVAL;
    if (nverts <= 0 || ntris < 0 || S <= 0 || S > GI_MAX_SAMPLES) return GI_EINVAL;
    if (smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE || gb->width <= 0 || gb->height <= 0 ||
        gb->width > GI_MAX_EDGE || gb->height > GI_MAX_EDGE)
        return GI_EINVAL;
    double sc[3] = {p->tau, p->bias, p->shadow_intensity};
    if (!finite_all(sc, 3)) return GI_ENONFINITE;
    if (!(p->tau > 0.0)) return GI_EINVAL;
    c = DiffCfg{smw, smh, S, gb->width, gb->height, p->tau, p->bias, p->shadow_intensity};
    return GI_OK;
}

// Visibility raster, per-triangle planes and (forward) texel depths.
int recompute(const float* pos, int nverts, const int32_t* idx, int ntris, const double* mvps, const DiffCfg& c,
              const DiffWs& w, hipStream_t s, bool with_depth) {
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((16 * c.S + 255) / 256)), dim3(256), 0, s, mvps, w.mvps32, 16 * c.S);
    int rc = internal::vis_raster(pos, nverts, idx, ntris, w.mvps32, c.S, c.smw, c.smh, w.vis, w.raster,
                                  w.raster_bytes, s);
    if (rc != GI_OK) return rc;
    if (ntris > 0)
        hipLaunchKernelGGL(k_diff_setup, dim3((unsigned)w.tri_blocks, c.S), dim3(kThreads), 0, s, pos, idx, ntris,
                           mvps, c, w.planes, w.large);
    const size_t msz = (size_t)c.smw * c.smh;
    if (with_depth)
        hipLaunchKernelGGL(k_diff_depth, dim3((unsigned)((msz + kThreads - 1) / kThreads), c.S), dim3(kThreads), 0, s,
                           w.vis, w.planes, ntris, c, w.depth);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // namespace
}  // namespace gi

using namespace gi;

extern "C" {

size_t gi_diff_workspace_size(int ntris, int nsamples, int smw, int smh, int view_w, int view_h) {
    if (ntris < 0 || nsamples <= 0 || nsamples > GI_MAX_SAMPLES || smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE ||
        smh > GI_MAX_EDGE || view_w <= 0 || view_h <= 0 || view_w > GI_MAX_EDGE || view_h > GI_MAX_EDGE)
        return 0;
    return diff_layout(ntris, nsamples, smw, smh, view_w, view_h, nullptr, nullptr);
}

int gi_diff_soft_forward(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                         const double* mvps_dev, const double* biased_mvps_dev, int nsamples, int smw, int smh,
                         const gi_gbuffer* gbuffer, const gi_diff_params* params, float* out_dev,
                         void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    DiffCfg c;
    int rc = check_common(positions_dev, nverts, indices_dev, ntris, mvps_dev, biased_mvps_dev, nsamples, smw, smh,
                          gbuffer, params, workspace_dev, c);
    if (rc != GI_OK) return rc;
    if (!out_dev) return GI_EINVAL;
    DiffWs w;
    size_t need = diff_layout(ntris, nsamples, smw, smh, c.gw, c.gh, &w, reinterpret_cast<char*>(workspace_dev));
    if (need == 0) return GI_ENOTSUP;
    if (workspace_bytes < need) return GI_ENOSPACE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    rc = recompute(positions_dev, nverts, indices_dev, ntris, mvps_dev, c, w, s, true);
    if (rc != GI_OK) return rc;
    hipLaunchKernelGGL(k_diff_shade, dim3(w.px_blocks), dim3(kThreads), 0, s, gbuffer->position_dev, biased_mvps_dev,
                       w.depth, c, out_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_diff_soft_backward(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                          const double* mvps_dev, const double* biased_mvps_dev, int nsamples, int smw, int smh,
                          const gi_gbuffer* gbuffer, const gi_diff_params* params, const float* grad_out_dev,
                          double* grad_biased_mvps_dev, double* grad_mvps_dev, float* grad_positions_dev,
                          void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    DiffCfg c;
    int rc = check_common(positions_dev, nverts, indices_dev, ntris, mvps_dev, biased_mvps_dev, nsamples, smw, smh,
                          gbuffer, params, workspace_dev, c);
    if (rc != GI_OK) return rc;
    if (!grad_out_dev || !grad_biased_mvps_dev || !grad_mvps_dev || !grad_positions_dev) return GI_EINVAL;
    DiffWs w;
    size_t need = diff_layout(ntris, nsamples, smw, smh, c.gw, c.gh, &w, reinterpret_cast<char*>(workspace_dev));
    if (need == 0) return GI_ENOTSUP;
    if (workspace_bytes < need) return GI_ENOSPACE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // the forward intermediates are recomputed, not stored; the pixel adjoint
    // evaluates the texel depths it reads from the keys and planes
    rc = recompute(positions_dev, nverts, indices_dev, ntris, mvps_dev, c, w, s, false);
    if (rc != GI_OK) return rc;
    const size_t msz = (size_t)smw * smh * nsamples;
    hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(kThreads), 0, s, w.grad_depth, msz, 0.0f);
    hipLaunchKernelGGL(k_diff_shade_bwd, dim3(w.pxb_blocks, nsamples), dim3(kThreads), 0, s, gbuffer->position_dev,
                       biased_mvps_dev, w.vis, w.planes, ntris, grad_out_dev, c, w.grad_depth, w.part_px);
    hipLaunchKernelGGL(k_diff_reduce, dim3(nsamples), dim3(kThreads), 0, s, w.part_px, w.pxb_blocks,
                       grad_biased_mvps_dev);
    if (ntris > 0) {
        hipLaunchKernelGGL(k_fill_f64, dim3(256), dim3(kThreads), 0, s, w.wsum, (size_t)3 * ntris * nsamples, 0.0);
        // touched flags: the region is 256-byte aligned and padded, so whole words clear it
        hipLaunchKernelGGL(k_fill_f32, dim3(64), dim3(kThreads), 0, s, reinterpret_cast<float*>(w.touched),
                           ((size_t)ntris * nsamples + 3) / 4, 0.0f);
        hipLaunchKernelGGL(k_diff_texel_bwd, dim3(w.tx_blocks, nsamples), dim3(kThreads), 0, s, w.vis, w.planes,
                           w.large, ntris, w.grad_depth, c, w.wsum, w.parts, w.touched);
        hipLaunchKernelGGL(k_diff_large_reduce, dim3(nsamples), dim3(kThreads), 0, s, w.parts, w.tx_blocks, ntris,
                           w.wsum);
        hipLaunchKernelGGL(k_diff_tri_bwd, dim3(w.tri_blocks, (nsamples + kMapsPerThread - 1) / kMapsPerThread),
                           dim3(kThreads), 0, s, positions_dev,
                           indices_dev, ntris, mvps_dev, w.planes, w.vis, w.grad_depth, w.wsum, w.touched, c,
                           grad_positions_dev, w.part_tri);
        hipLaunchKernelGGL(k_diff_reduce, dim3(nsamples), dim3(kThreads), 0, s, w.part_tri, w.tri_blocks,
                           grad_mvps_dev);
    } else {
        hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(kThreads), 0, s, grad_mvps_dev, (size_t)16 * nsamples, 0.0);
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // extern "C"
