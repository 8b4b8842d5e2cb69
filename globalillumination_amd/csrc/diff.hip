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
// The backward pass recomputes the visibility raster and texel depths (they
// are not stored by the forward pass), then
//   k_diff_shade_bwd  pixel adjoint: dL/d(B*M'_i) per block and dL/dd per
//                     texel (scatter-add),
//   k_diff_texel_bwd  texel adjoint: dL/dM'_i per block, dL/dvertex
//                     (scatter-add),
//   k_diff_reduce     fixed-order reduction of the per-block partials.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"
#include "gi_internal.hpp"

namespace gi {
namespace {

constexpr int kThreads = 256;

// Geometry (projections, barycentrics, window depths, the sigmoid argument)
// is evaluated in FP64: the argument (d - z + bias) / tau divides the
// rounding of two window depths near 0.9 by tau ~ 2e-4, so FP32 geometry
// would leave ~1e-3 noise in every term.  Texel depths are stored in FP64;
// the scatter-added adjoints (texel depth, vertex position) are FP32.
struct DiffCfg {
    int smw, smh, S;
    int gw, gh;
    double tau, bias, si;
};

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

__global__ __launch_bounds__(kThreads) void k_diff_depth(const float* __restrict__ pos, const int* __restrict__ idx,
                                                         int ntris, const double* __restrict__ mvps,
                                                         const unsigned long long* __restrict__ vis, DiffCfg c,
                                                         double* __restrict__ depth) {
    const size_t plane = (size_t)c.smw * c.smh;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int map = blockIdx.y;
    if (k >= plane) return;
    const unsigned long long key = vis[(size_t)map * plane + k];
    double d = 1.0;
    if (key != ~0ull) {
        const int tri = (int)((key & 0xffffffffull) >> 3);
        double m[16];
        load_m(mvps + 16 * map, m);
        double X[3], Y[3], Z[3], C[3][4];
        if (tri < ntris && tri_screen(pos, idx, tri, m, c.smw, c.smh, X, Y, Z, C)) {
            double D = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
            if (D != 0.0) {
                double xc = (double)(k % c.smw) + 0.5, yc = (double)(k / c.smw) + 0.5;
                double l1 = ((xc - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (yc - Y[0])) / D;
                double l2 = ((X[1] - X[0]) * (yc - Y[0]) - (xc - X[0]) * (Y[1] - Y[0])) / D;
                d = Z[0] + l1 * (Z[1] - Z[0]) + l2 * (Z[2] - Z[0]);
            }
        }
    }
    depth[(size_t)map * plane + k] = d;
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

// Pixel adjoint.  grid (pixel blocks, S).
__global__ __launch_bounds__(kThreads) void k_diff_shade_bwd(const float* __restrict__ gpos,
                                                             const double* __restrict__ biased,
                                                             const double* __restrict__ depth,
                                                             const float* __restrict__ grad_out, DiffCfg c,
                                                             float* __restrict__ grad_depth,
                                                             double* __restrict__ partial) {
    const size_t plane = (size_t)c.gw * c.gh;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    double gm[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) gm[e] = 0.0;
    if (k < plane && gpos[k] != 0.0f) {
        const double x = gpos[k], y = gpos[plane + k], z = gpos[2 * plane + k];
        const double gq = (double)grad_out[k] / (double)c.S * (1.0 - c.si);
        double m[16];
        load_m(biased + 16 * i, m);
        const Tap t = project(m, x, y, z, c);
        const size_t msz = (size_t)c.smw * c.smh;
        const double* d = depth + (size_t)i * msz;
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
                const double dd = in ? d[(size_t)b * c.smw + a] : 0.0;
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
            for (int row = 0; row < 4; ++row) gm[col * 4 + row] = gc[row] * xp[col];
    }
    block_sum16(gm, partial + ((size_t)i * gridDim.x + blockIdx.x) * 16);
}

// Texel adjoint.  grid (texel blocks, S); each thread takes kTexPerThread
// texels (strided by the block size, so loads stay coalesced).
//
// Vertex gradients are scatter-adds, and a large triangle (a ground plane
// covers most texels of every map) would send every wave's adds to the same
// three vertices, one cache line that device-scope atomics from all XCDs then
// serialize on.  So: a wave whose active texels all belong to one triangle
// sums its vertex gradients across lanes, and lane 0 keeps a running sum while
// consecutive texel rows stay on that triangle, flushing once when the
// triangle changes; the flushes of block b go to copy b % ncopies of the
// vertex-gradient array, and k_diff_vertex_reduce adds the copies in a fixed
// order.  Waves spanning several triangles add lane by lane.
constexpr int kTexPerThread = 4;

__device__ __forceinline__ void flush_vertex(const int* __restrict__ idx, int tri, const double acc[9],
                                             float* __restrict__ dst) {
    if (tri < 0) return;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int v = idx[3 * tri + j];
#pragma unroll
        for (int a = 0; a < 3; ++a) unsafeAtomicAdd(&dst[3 * v + a], (float)acc[3 * j + a]);
    }
}

__global__ __launch_bounds__(kThreads) void k_diff_texel_bwd(const float* __restrict__ pos, const int* __restrict__ idx,
                                                             int ntris, int nverts, const double* __restrict__ mvps,
                                                             const unsigned long long* __restrict__ vis,
                                                             const float* __restrict__ grad_depth, DiffCfg c,
                                                             float* __restrict__ copies, int ncopies,
                                                             double* __restrict__ partial) {
    const size_t msz = (size_t)c.smw * c.smh;
    const int i = blockIdx.y;
    float* __restrict__ dst = copies + (size_t)(blockIdx.x % ncopies) * nverts * 3;
    double m[16];
    load_m(mvps + 16 * i, m);
    double gm[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) gm[e] = 0.0;
    const int lane = threadIdx.x % 64;
    int cur = -1;  // lane 0: triangle of the running sum
    double run[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) run[e] = 0.0;
    for (int r = 0; r < kTexPerThread; ++r) {
        const size_t k = ((size_t)blockIdx.x * kTexPerThread + r) * kThreads + threadIdx.x;
        double gv[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) gv[e] = 0.0;
        int tri = -1;
        const double G = (k < msz) ? (double)grad_depth[(size_t)i * msz + k] : 0.0;
        const unsigned long long key = (k < msz) ? vis[(size_t)i * msz + k] : ~0ull;
        if (G != 0.0 && key != ~0ull) {
            const int t = (int)((key & 0xffffffffull) >> 3);
            double X[3], Y[3], Z[3], C[3][4];
            if (t < ntris && tri_screen(pos, idx, t, m, c.smw, c.smh, X, Y, Z, C)) {
                const double D = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
                if (D != 0.0) {
                    tri = t;
                    const double xc = (double)(k % c.smw) + 0.5, yc = (double)(k / c.smw) + 0.5;
                    const double l1 = ((xc - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (yc - Y[0])) / D;
                    const double l2 = ((X[1] - X[0]) * (yc - Y[0]) - (xc - X[0]) * (Y[1] - Y[0])) / D;
                    const double lam[3] = {1.0 - l1 - l2, l1, l2};
                    // plane gradient in window x, y; d moves by -lambda_j * grad when vertex j moves
                    const double gx = ((Z[1] - Z[0]) * (Y[2] - Y[0]) - (Z[2] - Z[0]) * (Y[1] - Y[0])) / D;
                    const double gy = ((X[1] - X[0]) * (Z[2] - Z[0]) - (X[2] - X[0]) * (Z[1] - Z[0])) / D;
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double gX = -G * lam[j] * gx * (0.5 * c.smw);
                        const double gY = -G * lam[j] * gy * (0.5 * c.smh);
                        const double gZ = G * lam[j] * 0.5;
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
                            gv[3 * j + a] = g;
                        }
                    }
                }
            }
        }
        // every lane of the wave reaches this point
        const unsigned long long act = __ballot(tri >= 0);
        if (act == 0ull) continue;
        const int first = __ffsll((long long)act) - 1;
        const int t0 = __shfl(tri, first, 64);
        if (__ballot(tri >= 0 && tri != t0) == 0ull) {
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                double x = gv[e];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
                gv[e] = x;
            }
            if (lane == 0) {
                if (cur != t0) {
                    flush_vertex(idx, cur, run, dst);
                    cur = t0;
#pragma unroll
                    for (int e = 0; e < 9; ++e) run[e] = 0.0;
                }
#pragma unroll
                for (int e = 0; e < 9; ++e) run[e] += gv[e];
            }
        } else if (tri >= 0) {
            flush_vertex(idx, tri, gv, dst);
        }
    }
    if (lane == 0) flush_vertex(idx, cur, run, dst);
    block_sum16(gm, partial + ((size_t)i * gridDim.x + blockIdx.x) * 16);
}

// grad_pos[e] += sum over copies c of copies[c][e], in copy order.
__global__ void k_diff_vertex_reduce(const float* __restrict__ copies, int ncopies, size_t n,
                                     float* __restrict__ grad_pos) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float s = 0.0f;
    for (int c = 0; c < ncopies; ++c) s += copies[(size_t)c * n + e];
    grad_pos[e] += s;
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

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Copies of the vertex-gradient array the texel adjoint spreads its adds
// over (capped so the copies stay within kCopyBudget bytes).
constexpr int kVertCopies = 64;
constexpr size_t kCopyBudget = size_t(256) << 20;

int vert_copies(int nverts) {
    size_t per = (size_t)nverts * 3 * sizeof(float);
    size_t n = per ? kCopyBudget / per : kVertCopies;
    return (int)(n < 1 ? 1 : (n > (size_t)kVertCopies ? kVertCopies : n));
}

struct DiffWs {
    float* copies;
    int ncopies;
    float* mvps32;
    unsigned long long* vis;
    double* depth;
    float* grad_depth;
    double* part_px;
    double* part_tx;
    void* raster;
    size_t raster_bytes;
    int px_blocks, tx_blocks;
};

size_t diff_layout(int nverts, int ntris, int S, int smw, int smh, int gw, int gh, DiffWs* ws, char* base) {
    const size_t msz = (size_t)smw * smh;
    const int px_blocks = (int)(((size_t)gw * gh + kThreads - 1) / kThreads);
    const int tx_blocks = (int)((msz + (size_t)kThreads * kTexPerThread - 1) / ((size_t)kThreads * kTexPerThread));
    const int ncopies = vert_copies(nverts);
    size_t off = 0;
    size_t o_cp = off; off = align256(off + sizeof(float) * 3 * (size_t)nverts * ncopies);
    size_t o_m32 = off; off = align256(off + sizeof(float) * 16 * (size_t)S);
    size_t o_vis = off; off = align256(off + sizeof(unsigned long long) * msz * S);
    size_t o_depth = off; off = align256(off + sizeof(double) * msz * S);
    size_t o_gd = off; off = align256(off + sizeof(float) * msz * S);
    size_t o_ppx = off; off = align256(off + sizeof(double) * 16 * (size_t)px_blocks * S);
    size_t o_ptx = off; off = align256(off + sizeof(double) * 16 * (size_t)tx_blocks * S);
    size_t rbytes = internal::vis_raster_workspace_size(ntris, S, smw, smh);
    size_t o_r = off; off = align256(off + rbytes);
    if (ws) {
        ws->copies = reinterpret_cast<float*>(base + o_cp);
        ws->ncopies = ncopies;
        ws->mvps32 = reinterpret_cast<float*>(base + o_m32);
        ws->vis = reinterpret_cast<unsigned long long*>(base + o_vis);
        ws->depth = reinterpret_cast<double*>(base + o_depth);
        ws->grad_depth = reinterpret_cast<float*>(base + o_gd);
        ws->part_px = reinterpret_cast<double*>(base + o_ppx);
        ws->part_tx = reinterpret_cast<double*>(base + o_ptx);
        ws->raster = base + o_r;
        ws->raster_bytes = rbytes;
        ws->px_blocks = px_blocks;
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
                 const double* mvps_dev, const double* biased_dev, int S, int smw, int smh, const gi_gbuffer* gb, const gi_diff_params* p,
                 void* ws, DiffCfg& c) {
    if (!positions_dev || (!indices_dev && ntris > 0) || !mvps_dev || !biased_dev || !gb || !gb->position_dev || !p ||
        !ws)
        return GI_EINVAL;
    if (nverts <= 0 || ntris < 0 || S <= 0 || S > GI_MAX_SAMPLES) return GI_EINVAL;
    if (smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE || gb->width <= 0 || gb->height <= 0 ||
        gb->width > GI_MAX_EDGE || gb->height > GI_MAX_EDGE)
        return GI_EINVAL;
    double sc[3] = {p->tau, p->bias, p->shadow_intensity};
    if (!finite_all(sc, 3)) return GI_ENONFINITE;
    if (!(p->tau > 0.0f)) return GI_EINVAL;
    c = DiffCfg{smw, smh, S, gb->width, gb->height, p->tau, p->bias, p->shadow_intensity};
    return GI_OK;
}

int recompute(const float* pos, int nverts, const int32_t* idx, int ntris, const double* mvps, const DiffCfg& c,
              const DiffWs& w, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((16 * c.S + 255) / 256)), dim3(256), 0, s, mvps, w.mvps32, 16 * c.S);
    int rc = internal::vis_raster(pos, nverts, idx, ntris, w.mvps32, c.S, c.smw, c.smh, w.vis, w.raster,
                                  w.raster_bytes, s);
    if (rc != GI_OK) return rc;
    const size_t msz = (size_t)c.smw * c.smh;
    hipLaunchKernelGGL(k_diff_depth, dim3((unsigned)((msz + kThreads - 1) / kThreads), c.S), dim3(kThreads), 0, s,
                       pos, idx, ntris, mvps, w.vis, c, w.depth);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // namespace
}  // namespace gi

using namespace gi;

extern "C" {

size_t gi_diff_workspace_size(int nverts, int ntris, int nsamples, int smw, int smh, int view_w, int view_h) {
    if (nverts <= 0 || ntris < 0 || nsamples <= 0 || nsamples > GI_MAX_SAMPLES || smw <= 0 || smh <= 0 ||
        smw > GI_MAX_EDGE || smh > GI_MAX_EDGE || view_w <= 0 || view_h <= 0 || view_w > GI_MAX_EDGE ||
        view_h > GI_MAX_EDGE)
        return 0;
    return diff_layout(nverts, ntris, nsamples, smw, smh, view_w, view_h, nullptr, nullptr);
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
    size_t need =
        diff_layout(nverts, ntris, nsamples, smw, smh, c.gw, c.gh, &w, reinterpret_cast<char*>(workspace_dev));
    if (need == 0) return GI_ENOTSUP;
    if (workspace_bytes < need) return GI_ENOSPACE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    rc = recompute(positions_dev, nverts, indices_dev, ntris, mvps_dev, c, w, s);
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
    size_t need =
        diff_layout(nverts, ntris, nsamples, smw, smh, c.gw, c.gh, &w, reinterpret_cast<char*>(workspace_dev));
    if (need == 0) return GI_ENOTSUP;
    if (workspace_bytes < need) return GI_ENOSPACE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // the forward intermediates are recomputed, not stored
    rc = recompute(positions_dev, nverts, indices_dev, ntris, mvps_dev, c, w, s);
    if (rc != GI_OK) return rc;
    const size_t msz = (size_t)smw * smh * nsamples;
    hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(kThreads), 0, s, w.grad_depth, msz, 0.0f);
    hipLaunchKernelGGL(k_diff_shade_bwd, dim3(w.px_blocks, nsamples), dim3(kThreads), 0, s, gbuffer->position_dev,
                       biased_mvps_dev, w.depth, grad_out_dev, c, w.grad_depth, w.part_px);
    const size_t nv3 = (size_t)nverts * 3;
    hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(kThreads), 0, s, w.copies, nv3 * w.ncopies, 0.0f);
    hipLaunchKernelGGL(k_diff_texel_bwd, dim3(w.tx_blocks, nsamples), dim3(kThreads), 0, s, positions_dev,
                       indices_dev, ntris, nverts, mvps_dev, w.vis, w.grad_depth, c, w.copies, w.ncopies, w.part_tx);
    hipLaunchKernelGGL(k_diff_vertex_reduce, dim3((unsigned)((nv3 + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       w.copies, w.ncopies, nv3, grad_positions_dev);
    hipLaunchKernelGGL(k_diff_reduce, dim3(nsamples), dim3(kThreads), 0, s, w.part_px, w.px_blocks,
                       grad_biased_mvps_dev);
    hipLaunchKernelGGL(k_diff_reduce, dim3(nsamples), dim3(kThreads), 0, s, w.part_tx, w.tx_blocks, grad_mvps_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // extern "C"
