// raster_kernels.hpp — device side of the binned raster (see raster.hip for
// the design and the raster rules).  Included by raster.hip (single-instance
// passes) and batch.hip (batched passes); everything here has internal
// linkage, so each translation unit holds its own instantiations.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"

namespace gi {
namespace {


// Tile edge of the binned raster: a template parameter of every tile-shaped
// kernel.  The single-instance passes use kTile; the batched path picks 32 or
// 64 per batch.
constexpr int kTile = 64;
constexpr int kRasterThreads = 256;
constexpr float kGuard = 4.0f;
constexpr int kMaxPoly = 9;
// Launch-shape constants; a build may override them (-DGI_SELF_RASTER_MAX_PX=...).
#ifndef GI_SELF_RASTER_MAX_PX
#define GI_SELF_RASTER_MAX_PX 256
#endif
// 63 setup records (64 B each) beside the 16 KiB tile and the record count
// keep a workgroup's LDS at 20,420 bytes, so 8 workgroups fit in a CU's
// 160 KiB (64 records made it 20,484 bytes: 7 workgroups, k_raster 3% slower).
#ifndef GI_REC_QUEUE
#define GI_REC_QUEUE 63
#endif
constexpr int kSelfRasterMaxPx = GI_SELF_RASTER_MAX_PX;  // bbox∩tile area a lone thread rasterizes itself
constexpr int kScanThreads = 1024;
constexpr int kRecQueue = GI_REC_QUEUE;  // LDS setup records a workgroup rasterizes together per chunk

struct MeshView {
    const float* __restrict__ pos;  // nverts x 3
    const float* __restrict__ nrm;  // nverts x 3 or null
    const int* __restrict__ idx;    // ntris x 3
    int nt;
};

// Per-map transform: a device array of 16-float matrices, or one by value.
struct MvpSource {
    const float* __restrict__ dev;
    Mat4 single;
    __device__ __forceinline__ void load(int map, float m[16]) const {
        if (dev) {
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = dev[16 * map + i];
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = single.m[i];
        }
    }
};

struct RasterCfg {
    int w, h;
    int tiles_x, tiles_y;
    float off_factor, off_units;
};

template <int NA>
struct CV {
    float c[4];
    float a[NA > 0 ? NA : 1];
};

template <int NA>
struct Tri {
    // 24.8 fixed-point window coordinates: within the guard band (|x_ndc| <=
    // kGuard) and a view of at most GI_MAX_EDGE pixels they stay below 2^24
    // in magnitude, so they, their differences and pixel-centre offsets fit
    // in int32; edge functions and areas are formed as 32 x 32 -> 64-bit
    // products.
    int X[3], Y[3];
    float w[3];
    float a[3][NA > 0 ? NA : 1];
    int front;
    int px0, px1, py0, py1;
    int tl[3];
    float z0, A, B, zoff;  // depth plane about vertex 0, per 1/256 pixel
};


__device__ __forceinline__ float plane_dist(const float c[4], int p) {
    switch (p) {
    case 0: return c[2] + c[3];
    case 1: return c[3] - c[2];
    case 2: return c[0] + kGuard * c[3];
    case 3: return kGuard * c[3] - c[0];
    case 4: return c[1] + kGuard * c[3];
    default: return kGuard * c[3] - c[1];
    }
}

template <int NA>
__device__ void load_tri(const MeshView& m, const float mvp[16], int tri, CV<NA> v[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int vi = m.idx[3 * tri + k];
        float x = m.pos[3 * vi + 0], y = m.pos[3 * vi + 1], z = m.pos[3 * vi + 2];
        mat4_vec4(mvp, x, y, z, 1.0f, v[k].c);
        if constexpr (NA == 6) {
            v[k].a[0] = x;
            v[k].a[1] = y;
            v[k].a[2] = z;
            v[k].a[3] = m.nrm[3 * vi + 0];
            v[k].a[4] = m.nrm[3 * vi + 1];
            v[k].a[5] = m.nrm[3 * vi + 2];
        }
    }
}

template <int NA>
__device__ __forceinline__ void lerp_cv(const CV<NA>& in, const CV<NA>& out, float din, float dout,
                                        CV<NA>& r) {
    float t = din / (din - dout);
#pragma unroll
    for (int k = 0; k < 4; ++k) r.c[k] = in.c[k] + t * (out.c[k] - in.c[k]);
#pragma unroll
    for (int k = 0; k < NA; ++k) r.a[k] = in.a[k] + t * (out.a[k] - in.a[k]);
}

// Fills `t` from three clip-space vertices.  Returns false for an empty piece.
template <int NA>
__device__ bool setup_piece(const CV<NA>& v0, const CV<NA>& v1, const CV<NA>& v2, const RasterCfg& cfg,
                            Tri<NA>& t) {
    const CV<NA>* vs[3] = {&v0, &v1, &v2};
    float hw = (float)cfg.w * 0.5f, hh = (float)cfg.h * 0.5f;
    float z[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float* c = vs[i]->c;
        if (!(c[3] > 0.0f)) return false;
        float xn = c[0] / c[3], yn = c[1] / c[3], zn = c[2] / c[3];
        float xw = xn * hw + hw;
        float yw = yn * hh + hh;
        z[i] = zn * 0.5f + 0.5f;
        t.X[i] = (int)rintf(xw * 256.0f);
        t.Y[i] = (int)rintf(yw * 256.0f);
        t.w[i] = c[3];
#pragma unroll
        for (int q = 0; q < NA; ++q) t.a[i][q] = vs[i]->a[q];
    }
    long long area = (long long)(t.X[1] - t.X[0]) * (t.Y[2] - t.Y[0]) - (long long)(t.Y[1] - t.Y[0]) * (t.X[2] - t.X[0]);
    if (area == 0) return false;
    t.front = area > 0;
    if (area < 0) {
        int sx = t.X[1]; t.X[1] = t.X[2]; t.X[2] = sx;
        int sy = t.Y[1]; t.Y[1] = t.Y[2]; t.Y[2] = sy;
        float sz = z[1]; z[1] = z[2]; z[2] = sz;
        float sw = t.w[1]; t.w[1] = t.w[2]; t.w[2] = sw;
#pragma unroll
        for (int q = 0; q < NA; ++q) { float sa = t.a[1][q]; t.a[1][q] = t.a[2][q]; t.a[2][q] = sa; }
        area = -area;
    }
    int minX = min(t.X[0], min(t.X[1], t.X[2])), maxX = max(t.X[0], max(t.X[1], t.X[2]));
    int minY = min(t.Y[0], min(t.Y[1], t.Y[2])), maxY = max(t.Y[0], max(t.Y[1], t.Y[2]));
    int px0 = (minX - 128 + 255) >> 8, px1 = (maxX - 128) >> 8;
    int py0 = (minY - 128 + 255) >> 8, py1 = (maxY - 128) >> 8;
    if (px0 < 0) px0 = 0;
    if (py0 < 0) py0 = 0;
    if (px1 > cfg.w - 1) px1 = cfg.w - 1;
    if (py1 > cfg.h - 1) py1 = cfg.h - 1;
    if (px0 > px1 || py0 > py1) return false;
    t.px0 = (int)px0; t.px1 = (int)px1; t.py0 = (int)py0; t.py1 = (int)py1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int a = (i + 1) % 3, b = (i + 2) % 3;
        int dy = t.Y[b] - t.Y[a], dx = t.X[b] - t.X[a];
        t.tl[i] = (dy < 0 || (dy == 0 && dx < 0)) ? 1 : 0;
    }
    const double dz1 = (double)z[1] - (double)z[0];
    const double dz2 = (double)z[2] - (double)z[0];
    const double inv_area = 1.0 / (double)area;
    const double a = ((double)(t.Y[2] - t.Y[0]) * dz1 + (double)(t.Y[0] - t.Y[1]) * dz2) * inv_area;
    const double b = ((double)(t.X[0] - t.X[2]) * dz1 + (double)(t.X[1] - t.X[0]) * dz2) * inv_area;
    t.z0 = z[0];
    t.A = (float)a;
    t.B = (float)b;
    t.zoff = 0.0f;
    if (cfg.off_factor != 0.0f || cfg.off_units != 0.0f) {
        double m = fmax(fabs(a), fabs(b)) * 256.0;
        float zmax = fmaxf(fmaxf(fabsf(z[0]), fabsf(z[1])), fabsf(z[2]));
        double r = 0.0;
        if (zmax > 0.0f) {
            int e;
            frexpf(zmax, &e);
            r = ldexp(1.0, (e - 1) - 23);
        }
        t.zoff = (float)((double)cfg.off_factor * m + (double)cfg.off_units * r);
    }
    return true;
}

// Sutherland-Hodgman over the six clip planes, from the triangle in
// poly[0..2]; `poly` and `tmp` hold kMaxPoly vertices each and swap roles
// after every plane.  Sets up fan piece `piece` of the result and returns the
// number of fan pieces; `ok` tells whether `piece` is non-empty.
template <int NA>
__device__ __forceinline__ int clip_fan(CV<NA>* poly, CV<NA>* tmp, int piece, const RasterCfg& cfg, Tri<NA>& t,
                                        bool& ok) {
    int n = 3;
    ok = false;
    for (int p = 0; p < 6; ++p) {
        int m = 0;
        for (int k = 0; k < n; ++k) {
            const CV<NA>& prev = poly[k == 0 ? n - 1 : k - 1];
            const CV<NA>& cur = poly[k];
            float dp = plane_dist(prev.c, p), dc = plane_dist(cur.c, p);
            if (dc >= 0.0f) {
                if (dp < 0.0f) lerp_cv<NA>(cur, prev, dc, dp, tmp[m++]);
                tmp[m++] = cur;
            } else if (dp >= 0.0f) {
                lerp_cv<NA>(prev, cur, dp, dc, tmp[m++]);
            }
        }
        n = m;
        if (n < 3) return 0;
        CV<NA>* x = poly;
        poly = tmp;
        tmp = x;
    }
    if (piece + 2 < n) ok = setup_piece<NA>(poly[0], poly[piece + 1], poly[piece + 2], cfg, t);
    return n - 2;
}

// clip_fan with the polygons in private memory.
template <int NA>
__device__ __noinline__ int clip_and_setup(const CV<NA> v[3], int piece, const RasterCfg& cfg,
                                           Tri<NA>& t, bool& ok) {
    CV<NA> poly[kMaxPoly], tmp[kMaxPoly];
    poly[0] = v[0];
    poly[1] = v[1];
    poly[2] = v[2];
    return clip_fan<NA>(poly, tmp, piece, cfg, t, ok);
}

// Number of fan pieces of `tri` (0 when trivially rejected) and the setup of
// piece `piece`; `ok` is false when that piece covers no pixel centre.
template <int NA>
__device__ int get_piece(const MeshView& m, const float mvp[16], int tri, int piece, const RasterCfg& cfg,
                         Tri<NA>& t, bool& ok) {
    CV<NA> v[3];
    load_tri<NA>(m, mvp, tri, v);
    bool inside_all = true;
    for (int p = 0; p < 6; ++p) {
        float d0 = plane_dist(v[0].c, p), d1 = plane_dist(v[1].c, p), d2 = plane_dist(v[2].c, p);
        int nin = (d0 >= 0.0f) + (d1 >= 0.0f) + (d2 >= 0.0f);
        if (nin == 0) { ok = false; return 0; }
        if (nin != 3) inside_all = false;
    }
    if (inside_all) {
        ok = (piece == 0) && setup_piece<NA>(v[0], v[1], v[2], cfg, t);
        return 1;
    }
    return clip_and_setup<NA>(v, piece, cfg, t, ok);
}

// Clip-free classification used by the hot kernels: kReject (outside a clip
// plane), kClip (crosses a plane: handled by k_clip through setup records) or
// kInside, in which case `t` is set up and `ok` tells whether it covers a
// pixel centre.
enum { kReject = 0, kClip = 1, kInside = 2 };

template <int NA>
__device__ __forceinline__ int classify(const MeshView& m, const float mvp[16], int tri, const RasterCfg& cfg,
                                        Tri<NA>& t, bool& ok) {
    CV<NA> v[3];
    load_tri<NA>(m, mvp, tri, v);
    bool inside_all = true;
    ok = false;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        float d0 = plane_dist(v[0].c, p), d1 = plane_dist(v[1].c, p), d2 = plane_dist(v[2].c, p);
        int nin = (d0 >= 0.0f) + (d1 >= 0.0f) + (d2 >= 0.0f);
        if (nin == 0) return kReject;
        if (nin != 3) inside_all = false;
    }
    if (!inside_all) return kClip;
    ok = setup_piece<NA>(v[0], v[1], v[2], cfg, t);
    return kInside;
}

// Edge functions at the centre of pixel (i, j); false when not covered.
template <int NA>
__device__ __forceinline__ bool eval_pixel(const Tri<NA>& t, int i, int j, long long e[3], float& zout) {
    const int Px = i * 256 + 128, Py = j * 256 + 128;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        e[k] = (long long)(t.X[b] - t.X[a]) * (Py - t.Y[a]) - (long long)(t.Y[b] - t.Y[a]) * (Px - t.X[a]);
    }
    if (e[0] + t.tl[0] <= 0 || e[1] + t.tl[1] <= 0 || e[2] + t.tl[2] <= 0) return false;
    float dx = (float)(Px - t.X[0]), dy = (float)(Py - t.Y[0]);
    float z = t.z0 + (t.A * dx + t.B * dy);
    z = z + t.zoff;
    if (!(z > 0.0f)) z = 0.0f;
    if (z > 1.0f) z = 1.0f;
    zout = z;
    return true;
}

// Conservative test: can any pixel centre of [x0,x1]x[y0,y1] be inside?
template <int NA>
__device__ __forceinline__ bool rect_overlap(const Tri<NA>& t, int x0, int x1, int y0, int y1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        const int ex = t.X[b] - t.X[a], ey = t.Y[b] - t.Y[a];
        const int Px = ((ey < 0) ? x1 : x0) * 256 + 128;
        const int Py = ((ex > 0) ? y1 : y0) * 256 + 128;
        const long long e = (long long)ex * (Py - t.Y[a]) - (long long)ey * (Px - t.X[a]);
        if (e + t.tl[k] <= 0) return false;
    }
    return true;
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

// ---------------------------------------------------------------------------
// Compact setup record of one screen-space piece.  Wide pieces get one record
// per map in HBM (written once by the binning pass); pieces a whole workgroup
// rasterizes are shared through LDS.
// ---------------------------------------------------------------------------
struct Setup {
    int X[3], Y[3];            // 24.8 fixed-point window coordinates
    int px0, px1, py0, py1;    // pixel-centre bounding box (clamped)
    int tl;                    // bit k: top-left tie bit of the edge opposite vertex k
    uint32_t id;               // tri * 8 + fan piece
    float z0, A, B, zoff;      // depth plane about vertex 0
};

__device__ __forceinline__ Setup to_setup(const Tri<0>& t, uint32_t id) {
    Setup s;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s.X[i] = t.X[i];
        s.Y[i] = t.Y[i];
    }
    s.px0 = t.px0; s.px1 = t.px1; s.py0 = t.py0; s.py1 = t.py1;
    s.tl = t.tl[0] | (t.tl[1] << 1) | (t.tl[2] << 2);
    s.id = id;
    s.z0 = t.z0; s.A = t.A; s.B = t.B; s.zoff = t.zoff;
    return s;
}

// The count pass's record of a directly binned triangle, 48 bytes: the list
// entry supplies the id, and the tie bits follow from the edges.  Pixel
// bounds fit 16 bits (map edges are at most GI_MAX_EDGE).
struct alignas(16) SetupRec {
    int X[3], Y[3];
    float z0, A, B, zoff;
    uint32_t bx, by;  // px0 | px1 << 16, py0 | py1 << 16
};
static_assert(sizeof(SetupRec) == 48, "three 16-byte loads");
static_assert(GI_MAX_EDGE <= 65536, "pixel bounds are packed in 16 bits");

__device__ __forceinline__ SetupRec to_rec(const Tri<0>& t) {
    SetupRec r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.X[i] = t.X[i];
        r.Y[i] = t.Y[i];
    }
    r.z0 = t.z0; r.A = t.A; r.B = t.B; r.zoff = t.zoff;
    r.bx = (uint32_t)t.px0 | ((uint32_t)t.px1 << 16);
    r.by = (uint32_t)t.py0 | ((uint32_t)t.py1 << 16);
    return r;
}

// Same tie rule as setup_piece, on the same 24.8 coordinates.
__device__ __forceinline__ Setup from_rec(const SetupRec& r, uint32_t id) {
    Setup s;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s.X[i] = r.X[i];
        s.Y[i] = r.Y[i];
    }
    s.px0 = (int)(r.bx & 0xFFFFu); s.px1 = (int)(r.bx >> 16);
    s.py0 = (int)(r.by & 0xFFFFu); s.py1 = (int)(r.by >> 16);
    s.tl = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int a = (i + 1) % 3, b = (i + 2) % 3;
        const int dy = r.Y[b] - r.Y[a], dx = r.X[b] - r.X[a];
        s.tl |= (dy < 0 || (dy == 0 && dx < 0)) ? (1 << i) : 0;
    }
    s.id = id;
    s.z0 = r.z0; s.A = r.A; s.B = r.B; s.zoff = r.zoff;
    return s;
}

// Same arithmetic as eval_pixel on the compact record.
__device__ __forceinline__ bool eval_setup(const Setup& s, int i, int j, float& zout) {
    const int Px = i * 256 + 128, Py = j * 256 + 128;
    long long e0 = (long long)(s.X[2] - s.X[1]) * (Py - s.Y[1]) - (long long)(s.Y[2] - s.Y[1]) * (Px - s.X[1]);
    long long e1 = (long long)(s.X[0] - s.X[2]) * (Py - s.Y[2]) - (long long)(s.Y[0] - s.Y[2]) * (Px - s.X[2]);
    long long e2 = (long long)(s.X[1] - s.X[0]) * (Py - s.Y[0]) - (long long)(s.Y[1] - s.Y[0]) * (Px - s.X[0]);
    if (e0 + (s.tl & 1) <= 0 || e1 + ((s.tl >> 1) & 1) <= 0 || e2 + ((s.tl >> 2) & 1) <= 0) return false;
    float dx = (float)(Px - s.X[0]), dy = (float)(Py - s.Y[0]);
    float z = s.z0 + (s.A * dx + s.B * dy);
    z = z + s.zoff;
    if (!(z > 0.0f)) z = 0.0f;
    if (z > 1.0f) z = 1.0f;
    zout = z;
    return true;
}

__device__ __forceinline__ bool rect_overlap_s(const Setup& s, int x0, int x1, int y0, int y1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        const int ex = s.X[b] - s.X[a], ey = s.Y[b] - s.Y[a];
        const int Px = ((ey < 0) ? x1 : x0) * 256 + 128;
        const int Py = ((ex > 0) ? y1 : y0) * 256 + 128;
        const long long e = (long long)ex * (Py - s.Y[a]) - (long long)ey * (Px - s.X[a]);
        if (e + ((s.tl >> k) & 1) <= 0) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Binning
// ---------------------------------------------------------------------------
// Pieces overlapping more than wide_span<TS>() tiles are "wide" (binned from
// setup records by k_bin_wide): 4 tiles of 64x64 texels, the same texel area
// for other tile edges, so the set of wide pieces does not grow as tiles
// shrink.
template <int TS>
constexpr int wide_span() {
    return 4 * (64 / TS) * (64 / TS);
}
constexpr uint32_t kWideBit = 0x80000000u;
constexpr int kWideBlocks = 64;         // workgroups per map binning wide pieces
constexpr int kClipBlocks = 16;         // workgroups per map clipping listed triangles
constexpr int kClipThreads = 64;        // threads per clipping workgroup

struct BinBufs {
    uint32_t* counts;    // nmaps x ntiles (scatter cursors after k_scan)
    uint32_t* offsets;   // nmaps x (ntiles + 1)
    int* overflow;       // nmaps: lists or wide records exceed the workspace
    int* wide_over;      // nmaps
    uint32_t* nwide;     // nmaps
    Setup* wide;         // nmaps x wcap
    uint32_t* nclip;     // nmaps
    uint32_t* clip;      // nmaps x wcap: triangles crossing a clip plane
    uint32_t* entries;   // nmaps x cap
    uint32_t* tiles;     // nmaps x tstride: the count pass's tiles of each directly binned triangle
    SetupRec* setups;    // nmaps x tstride: its setup record (64-texel tiles only)
    uint32_t cap, wcap, tstride;
};

// The tiles a directly binned triangle overlaps, as the count pass found
// them, for the scatter pass: bits 0-15 the first tile of its tile box,
// bits 16-17 the box width in tiles minus 1, bits 18-21 which of the box's
// (at most 4) tiles it overlaps.  0: none (rejected, clipped, wide or empty).
__device__ __forceinline__ uint32_t tiles_code(int base, int ntx, uint32_t mask) {
    return mask ? ((uint32_t)base | ((uint32_t)(ntx - 1) << 16) | (mask << 18)) : 0u;
}

__device__ __forceinline__ void bin_tile(const BinBufs& b, int map, int ntiles, int tile, uint32_t entry,
                                         bool scatter) {
    uint32_t* cnt = b.counts + (size_t)map * ntiles;
    if (!scatter) {
        atomicAdd(&cnt[tile], 1u);
    } else {
        uint32_t pos = b.offsets[(size_t)map * (ntiles + 1) + tile] + atomicAdd(&cnt[tile], 1u);
        if (pos < b.cap) b.entries[(size_t)map * b.cap + pos] = entry;
    }
}

// A workgroup barrier that orders LDS only.  __syncthreads() also waits for
// the workgroup's outstanding global stores and atomics (its release fence
// covers global memory); the binning kernels' barriers guard their LDS
// histograms and prefix sums, and what they write to global memory is read
// only by later kernels (an atomic whose result goes to LDS has returned
// before that LDS write).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Appends a setup record (count pass only); records past the capacity flag
// the map for the fallback raster.
__device__ __forceinline__ void push_record(const BinBufs& b, int map, const Tri<0>& t, uint32_t entry) {
    uint32_t w = atomicAdd(&b.nwide[map], 1u);
    if (w < b.wcap) b.wide[(size_t)map * b.wcap + w] = to_setup(t, entry);
    else b.wide_over[map] = 1;
}

// One thread per (triangle, map).  Triangles inside every clip plane whose
// piece spans at most kWideSpan tiles are binned here; wider ones get a setup
// record (binned by k_bin_wide), and triangles crossing a clip plane are
// listed for k_clip.
//
// Tile counts are first aggregated in an LDS histogram, so the workgroup
// issues one device-scope atomic per touched tile instead of one per
// triangle: consecutive triangles of a mesh are neighbours and share tiles,
// and device-scope atomics on the same line serialize outside the XCD's L2.
constexpr int kBinHist = 4096;  // largest tile grid aggregated in LDS (64 x 64 tiles)

template <int TS, bool kScatter>
__device__ __forceinline__ void bin_body(const MeshView& mesh, const MvpSource& mvps, const RasterCfg& cfg,
                                         const BinBufs& b, int bx, int map) {
    __shared__ uint32_t hist[kBinHist];
    const int tid = threadIdx.x;
    const int tri = bx * kRasterThreads + tid;
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    const bool lds = ntiles <= kBinHist;
    if (kScatter && b.overflow[map]) return;  // uniform per workgroup
    if (lds)
        for (int p = tid; p < ntiles; p += blockDim.x) hist[p] = 0;
    // Up to wide_span<TS>() tiles per thread, kept in registers (static indices).
    constexpr int kWideSpan = wide_span<TS>();
    // 64-texel tiles: the count pass stores each triangle's tiles and the
    // scatter pass reads them back; smaller tiles (a wider span than the code
    // holds) classify again in the scatter pass
    constexpr bool kCache = kWideSpan <= 4;
    int tile_of[kWideSpan];
    uint32_t rank_of[kWideSpan];
#pragma unroll
    for (int k = 0; k < kWideSpan; ++k) tile_of[k] = -1;
    const uint32_t entry = (uint32_t)tri * 8u;
    if (kScatter && kCache) {
        // the count pass classified this (triangle, map) already
        if (tri < mesh.nt) {
            const uint32_t code = b.tiles[(size_t)map * b.tstride + tri];
            const int base = (int)(code & 0xFFFFu), ntx = (int)((code >> 16) & 3u) + 1;
            const uint32_t mask = code >> 18;
#pragma unroll
            for (int k = 0; k < kWideSpan; ++k)
                if (mask & (1u << k)) tile_of[k] = base + (k / ntx) * cfg.tiles_x + (k % ntx);
        }
    } else if (tri < mesh.nt) {
        float mvp[16];
        mvps.load(map, mvp);
        uint32_t code = 0;
        Tri<0> t;
        bool ok;
        int cls = classify<0>(mesh, mvp, tri, cfg, t, ok);
        if (cls == kClip) {
            if (!kScatter) {
                uint32_t c = atomicAdd(&b.nclip[map], 1u);
                if (c < b.wcap) b.clip[(size_t)map * b.wcap + c] = (uint32_t)tri;
                else b.wide_over[map] = 1;
            }
        } else if (ok) {
            int tx0 = t.px0 / TS, tx1 = t.px1 / TS, ty0 = t.py0 / TS, ty1 = t.py1 / TS;
            int ntx = tx1 - tx0 + 1;
            int span = ntx * (ty1 - ty0 + 1);
            if (span > kWideSpan) {
                if (!kScatter) push_record(b, map, t, entry);
            } else {
                uint32_t mask = 0;
#pragma unroll
                for (int k = 0; k < kWideSpan; ++k) {
                    if (k >= span) continue;
                    int tx = tx0 + k % ntx, ty = ty0 + k / ntx;
                    if (span > 1) {
                        int x0 = max(tx * TS, t.px0), x1 = min(tx * TS + TS - 1, t.px1);
                        int y0 = max(ty * TS, t.py0), y1 = min(ty * TS + TS - 1, t.py1);
                        if (!rect_overlap<0>(t, x0, x1, y0, y1)) continue;
                    }
                    tile_of[k] = ty * cfg.tiles_x + tx;
                    mask |= 1u << k;
                }
                code = tiles_code(ty0 * cfg.tiles_x + tx0, ntx, mask);
                if (!kScatter && kCache && mask) b.setups[(size_t)map * b.tstride + tri] = to_rec(t);
            }
        }
        if (!kScatter && kCache) b.tiles[(size_t)map * b.tstride + tri] = code;
    }
    if (!lds) {
#pragma unroll
        for (int k = 0; k < kWideSpan; ++k)
            if (tile_of[k] >= 0) bin_tile(b, map, ntiles, tile_of[k], entry, kScatter);
        return;
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < kWideSpan; ++k)
        if (tile_of[k] >= 0) rank_of[k] = atomicAdd(&hist[tile_of[k]], 1u);
    lds_barrier();
    uint32_t* cnt = b.counts + (size_t)map * ntiles;
    for (int p = tid; p < ntiles; p += blockDim.x) {
        uint32_t c = hist[p];
        if (c == 0) continue;
        if (!kScatter) atomicAdd(&cnt[p], c);
        else hist[p] = atomicAdd(&cnt[p], c);  // this workgroup's base within the tile's list
    }
    if (!kScatter) return;
    lds_barrier();
    const uint32_t* off = b.offsets + (size_t)map * (ntiles + 1);
#pragma unroll
    for (int k = 0; k < kWideSpan; ++k) {
        if (tile_of[k] < 0) continue;
        uint32_t pos = off[tile_of[k]] + hist[tile_of[k]] + rank_of[k];
        if (pos < b.cap) b.entries[(size_t)map * b.cap + pos] = entry;
    }
}

template <int TS, bool kScatter>
__global__ __launch_bounds__(256) void k_bin(MeshView mesh, MvpSource mvps, RasterCfg cfg, BinBufs b) {
    bin_body<TS, kScatter>(mesh, mvps, cfg, b, blockIdx.x, blockIdx.y);
}


// Clips the listed triangles; every non-empty fan piece becomes a setup
// record.  Runs between the count passes, before k_bin_wide.  One thread per
// (listed triangle, fan piece): each clips the triangle and keeps its own
// piece, so the pieces of a triangle are set up in parallel.  Workgroup
// `part` of `nparts` takes every nparts-th stride of the map's items.
constexpr int kMaxPieces = kMaxPoly - 2;

// Each thread keeps its two polygons in LDS (kClipThreads x 360 bytes):
// private arrays indexed at run time would live in scratch memory, and the
// clip's chain of dependent reads and writes would wait on each one.
__device__ __forceinline__ void clip_body(const MeshView& mesh, const MvpSource& mvps, const RasterCfg& cfg,
                                          const BinBufs& b, int part, int nparts, int map) {
    __shared__ CV<0> pool[kClipThreads][2][kMaxPoly];
    const uint32_t n = min(b.nclip[map], b.wcap) * (uint32_t)kMaxPieces;
    float mvp[16];
    mvps.load(map, mvp);
    CV<0>* poly = pool[threadIdx.x][0];
    CV<0>* tmp = pool[threadIdx.x][1];
    for (uint32_t item = part * kClipThreads + threadIdx.x; item < n; item += nparts * kClipThreads) {
        const uint32_t c = item / kMaxPieces;
        const int piece = (int)(item % kMaxPieces);
        const int tri = (int)b.clip[(size_t)map * b.wcap + c];
        load_tri<0>(mesh, mvp, tri, poly);
        Tri<0> t;
        bool ok;
        clip_fan<0>(poly, tmp, piece, cfg, t, ok);
        if (ok) push_record(b, map, t, (uint32_t)tri * 8u + (uint32_t)piece);
    }
}

__global__ __launch_bounds__(kClipThreads) void k_clip(MeshView mesh, MvpSource mvps, RasterCfg cfg, BinBufs b) {
    clip_body(mesh, mvps, cfg, b, blockIdx.x, gridDim.x, blockIdx.y);
}

// Wide pieces (setup records), kWideThreads threads per workgroup.  Part p
// of nparts takes records p, p + nparts, p + 2 nparts, ..., one per thread,
// kWideThreads records a round.  The round's (record, tile) pairs are
// numbered through a prefix sum of the records' tile spans, and each thread
// takes every kWideThreads-th pair, so the pairs' appends (whose scatter
// atomics return the list slot) are all in flight at once whether the round
// holds many records of a few tiles (a triangle soup on a large map) or one
// record over the whole map (a ground plane's clipped piece); a map's few
// records still go to different parts.
constexpr int kWideThreads = 256;

template <int TS, bool kScatter>
__device__ __forceinline__ void wide_tile(const RasterCfg& cfg, const BinBufs& b, int map, int ntiles, const Setup& s,
                                          uint32_t w, int tx0, int ty0, int ntx, int q) {
    const int tx = tx0 + q % ntx, ty = ty0 + q / ntx;
    const int x0 = max(tx * TS, s.px0), x1 = min(tx * TS + TS - 1, s.px1);
    const int y0 = max(ty * TS, s.py0), y1 = min(ty * TS + TS - 1, s.py1);
    if (rect_overlap_s(s, x0, x1, y0, y1)) bin_tile(b, map, ntiles, ty * cfg.tiles_x + tx, kWideBit | w, kScatter);
}

template <int TS, bool kScatter>
__device__ __forceinline__ void bin_wide_body(const RasterCfg& cfg, const BinBufs& b, int part, int nparts, int map) {
    __shared__ int incl[kWideThreads];  // inclusive prefix of the round's spans
    __shared__ int wave_sum[kWideThreads / 64];
    if (kScatter && b.overflow[map]) return;  // uniform per workgroup
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    const uint32_t n = min(b.nwide[map], b.wcap);
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t stride = (uint32_t)nparts;
    const Setup* recs = b.wide + (size_t)map * b.wcap;
    for (uint32_t r0 = 0; (uint32_t)part + stride * r0 < n; r0 += kWideThreads) {  // uniform
        const uint32_t w = (uint32_t)part + stride * (r0 + (uint32_t)tid);
        int span = 0;
        if (w < n) {
            const Setup& s = recs[w];
            span = (s.px1 / TS - s.px0 / TS + 1) * (s.py1 / TS - s.py0 / TS + 1);
        }
        int v = span;  // inclusive scan in the wave, then across the waves
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(v, d, 64);
            if (lane >= d) v += u;
        }
        if (lane == 63) wave_sum[wave] = v;
        lds_barrier();
        int base = 0, total = 0;
#pragma unroll
        for (int k = 0; k < kWideThreads / 64; ++k) {
            base += k < wave ? wave_sum[k] : 0;
            total += wave_sum[k];
        }
        incl[tid] = base + v;
        lds_barrier();
        for (int p = tid; p < total; p += kWideThreads) {
            // the round's record holding pair p: the first k with incl[k] > p
            int lo = 0, hi = kWideThreads - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (incl[mid] > p) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t wk = (uint32_t)part + stride * (r0 + (uint32_t)lo);
            const Setup s = recs[wk];
            const int tx0 = s.px0 / TS, ty0 = s.py0 / TS;
            const int ntx = s.px1 / TS - tx0 + 1;
            const int q = p - (lo == 0 ? 0 : incl[lo - 1]);
            wide_tile<TS, kScatter>(cfg, b, map, ntiles, s, wk, tx0, ty0, ntx, q);
        }
        lds_barrier();  // incl and wave_sum are rewritten by the next round
    }
}

template <int TS, bool kScatter>
__global__ __launch_bounds__(256) void k_bin_wide(RasterCfg cfg, BinBufs b) {
    bin_wide_body<TS, kScatter>(cfg, b, blockIdx.x, gridDim.x, blockIdx.y);
}

// The scatter pass when the count pass stored each triangle's tile code
// (64-texel tiles): kT triangles per thread.  A workgroup's appends cost two
// dependent global round trips (the atomics that reserve its slice of each
// tile's list, then the list offsets), so covering kT x 256 triangles per
// workgroup divides those rounds by kT.  Entries land in the same lists as
// bin_body's scatter (their order within a list does not matter: the raster
// takes the minimum).  `block` covers triangles [block kT 256, (block + 1) kT 256)
// of map `map`.
template <int TS, int kT>
__device__ __forceinline__ void bin_scatter_body(const MeshView& mesh, const RasterCfg& cfg, const BinBufs& b,
                                                 int block, int map) {
    constexpr int kWideSpan = wide_span<TS>();
    static_assert(kWideSpan <= 4, "tile codes are stored for 64-texel tiles");
    __shared__ uint32_t hist[kBinHist];
    const int tid = (int)threadIdx.x;
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    if (b.overflow[map]) return;  // uniform per workgroup
    const bool lds = ntiles <= kBinHist;
    if (lds)
        for (int p = tid; p < ntiles; p += 256) hist[p] = 0;
    int tile_of[kT][kWideSpan];
    uint32_t rank_of[kT][kWideSpan];
#pragma unroll
    for (int t = 0; t < kT; ++t) {
        const int tri = (block * kT + t) * 256 + tid;
#pragma unroll
        for (int k = 0; k < kWideSpan; ++k) tile_of[t][k] = -1;
        if (tri < mesh.nt) {
            const uint32_t code = b.tiles[(size_t)map * b.tstride + tri];
            const int base = (int)(code & 0xFFFFu), ntx = (int)((code >> 16) & 3u) + 1;
            const uint32_t mask = code >> 18;
#pragma unroll
            for (int k = 0; k < kWideSpan; ++k)
                if (mask & (1u << k)) tile_of[t][k] = base + (k / ntx) * cfg.tiles_x + (k % ntx);
        }
    }
    if (!lds) {
#pragma unroll
        for (int t = 0; t < kT; ++t)
#pragma unroll
            for (int k = 0; k < kWideSpan; ++k)
                if (tile_of[t][k] >= 0)
                    bin_tile(b, map, ntiles, tile_of[t][k], (uint32_t)((block * kT + t) * 256 + tid) * 8u,
                             true);
        return;
    }
    lds_barrier();
#pragma unroll
    for (int t = 0; t < kT; ++t)
#pragma unroll
        for (int k = 0; k < kWideSpan; ++k)
            if (tile_of[t][k] >= 0) rank_of[t][k] = atomicAdd(&hist[tile_of[t][k]], 1u);
    lds_barrier();
    uint32_t* cnt = b.counts + (size_t)map * ntiles;
    for (int p = tid; p < ntiles; p += 256) {
        const uint32_t c = hist[p];
        if (c != 0) hist[p] = atomicAdd(&cnt[p], c);  // this workgroup's base within the tile's list
    }
    lds_barrier();
    const uint32_t* off = b.offsets + (size_t)map * (ntiles + 1);
#pragma unroll
    for (int t = 0; t < kT; ++t) {
        const uint32_t entry = (uint32_t)((block * kT + t) * 256 + tid) * 8u;
#pragma unroll
        for (int k = 0; k < kWideSpan; ++k) {
            if (tile_of[t][k] < 0) continue;
            const uint32_t pos = off[tile_of[t][k]] + hist[tile_of[t][k]] + rank_of[t][k];
            if (pos < b.cap) b.entries[(size_t)map * b.cap + pos] = entry;
        }
    }
}
// The single-instance scatter: workgroups past `tri_blocks` bin the wide
// pieces' records instead (bin_wide_body's scatter, kWideBlocks of them per
// map), so the two appends share one launch and overlap.
template <int TS, int kT>
__global__ __launch_bounds__(256) void k_bin_scatter(MeshView mesh, RasterCfg cfg, BinBufs b, int tri_blocks) {
    if ((int)blockIdx.x >= tri_blocks) {  // uniform per workgroup
        bin_wide_body<TS, true>(cfg, b, (int)blockIdx.x - tri_blocks, kWideBlocks, (int)blockIdx.y);
        return;
    }
    bin_scatter_body<TS, kT>(mesh, cfg, b, (int)blockIdx.x, (int)blockIdx.y);
}

constexpr int kScatterTris = 4;  // triangles per thread in k_bin_scatter

// Exclusive scan of one map's tile counts; resets counts to 0 (they become
// the scatter cursors) and flags maps whose lists exceed the workspace.
__device__ __forceinline__ void scan_body(const BinBufs& b, int ntiles, int map) {
    __shared__ unsigned long long part[kScanThreads];
    uint32_t* c = b.counts + (size_t)map * ntiles;
    uint32_t* o = b.offsets + (size_t)map * (ntiles + 1);
    int per = (ntiles + kScanThreads - 1) / kScanThreads;
    int lo = threadIdx.x * per, hi = min(lo + per, ntiles);
    unsigned long long s = 0;
    for (int i = lo; i < hi; ++i) s += c[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        unsigned long long v = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0ull;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - s;  // exclusive prefix
    for (int i = lo; i < hi; ++i) {
        o[i] = (uint32_t)run;
        run += c[i];
        c[i] = 0;
    }
    if (threadIdx.x == kScanThreads - 1) {
        unsigned long long total = part[kScanThreads - 1];
        bool over = total > b.cap || b.wide_over[map] != 0;
        o[ntiles] = (uint32_t)(total > b.cap ? b.cap : total);
        b.overflow[map] = over ? 1 : 0;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan(BinBufs b, int ntiles) { scan_body(b, ntiles, blockIdx.x); }

// ---------------------------------------------------------------------------
// Tile raster
// ---------------------------------------------------------------------------
template <bool kVis>
struct KeyOf;
template <>
struct KeyOf<false> {
    using T = uint32_t;
    static constexpr uint32_t clear = 0x3F800000u;  // 1.0f
};
template <>
struct KeyOf<true> {
    using T = unsigned long long;
    static constexpr unsigned long long clear = ~0ull;
};

// Edge functions of `s` at pixel centre (i, j) and their per-pixel steps.
// Stepping adds exact integers, so every value equals eval_setup's.
struct Edges {
    long long e[3];
    long long dx[3], dy[3];  // change per +1 pixel in x / y
};

__device__ __forceinline__ void edges_at(const Setup& s, int i, int j, Edges& E) {
    const int Px = i * 256 + 128, Py = j * 256 + 128;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        const int ex = s.X[b] - s.X[a], ey = s.Y[b] - s.Y[a];
        E.e[k] = (long long)ex * (Py - s.Y[a]) - (long long)ey * (Px - s.X[a]);
        E.dx[k] = (long long)ey * -256;
        E.dy[k] = (long long)ex * 256;
    }
}

// Depth of a covered pixel centre at fixed-point offset (dx, dy) from vertex 0.
__device__ __forceinline__ uint32_t plane_bits(const Setup& s, int dx, int dy) {
    float z = s.z0 + (s.A * (float)dx + s.B * (float)dy);
    z = z + s.zoff;
    if (!(z > 0.0f)) z = 0.0f;  // NaN and -0 become +0
    // z is +0 or positive here (never NaN), where min(z, 1) is `z > 1 ? 1 : z`
    return fbits(__builtin_fminf(z, 1.0f));
}

template <bool kVis>
__device__ __forceinline__ void store_min(const Setup& s, uint32_t zb, int slot, typename KeyOf<kVis>::T* tile) {
    if constexpr (kVis) {
        atomicMin(&tile[slot], ((unsigned long long)zb << 32) | s.id);
    } else {
        atomicMin(&tile[slot], zb);
    }
}

template <bool kVis, typename E>
__device__ __forceinline__ void plot(const Setup& s, const E e[3], int i, int j, int slot,
                                     typename KeyOf<kVis>::T* tile) {
    if (e[0] + (E)(s.tl & 1) <= 0 || e[1] + (E)((s.tl >> 1) & 1) <= 0 || e[2] + (E)((s.tl >> 2) & 1) <= 0) return;
    store_min<kVis>(s, plane_bits(s, i * 256 + 128 - s.X[0], j * 256 + 128 - s.Y[0]), slot, tile);
}

// True when every pixel centre of [rx0, rx1] x [ry0, ry1] passes plot()'s
// three edge tests.  Each edge value is affine in the pixel indices, so its
// minimum over the rectangle is at a corner, and the test `e + tl > 0` holds
// everywhere once it holds at the four corners.
__device__ __forceinline__ bool rect_inside(const Setup& s, int rx0, int rx1, int ry0, int ry1) {
    const int cx[2] = {rx0 * 256 + 128, rx1 * 256 + 128};
    const int cy[2] = {ry0 * 256 + 128, ry1 * 256 + 128};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        const int ex = s.X[b] - s.X[a], ey = s.Y[b] - s.Y[a];
        const long long tl = (s.tl >> k) & 1;
        // the corner minimising ex * Py - ey * Px
        const int px = ey > 0 ? cx[1] : cx[0];
        const int py = ex > 0 ? cy[0] : cy[1];
        if ((long long)ex * (py - s.Y[a]) - (long long)ey * (px - s.X[a]) + tl <= 0) return false;
    }
    return true;
}

// One thread covers the whole (small) rectangle, row by row.  When every
// edge value inside the triangle's bounding box fits in 32 bits, the exact
// stepping runs in int32.
template <int TS, bool kVis, typename E>
__device__ __forceinline__ void raster_rect_self_t(const Setup& s, const Edges& E64, int x0, int y0, int rx0, int rx1,
                                                   int ry0, int ry1, typename KeyOf<kVis>::T* tile) {
    E row[3], dx[3], dy[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        row[k] = (E)E64.e[k];
        dx[k] = (E)E64.dx[k];
        dy[k] = (E)E64.dy[k];
    }
    for (int j = ry0; j <= ry1; ++j) {
        E e[3] = {row[0], row[1], row[2]};
        for (int i = rx0; i <= rx1; ++i) {
            plot<kVis, E>(s, e, i, j, (j - y0) * TS + (i - x0), tile);
#pragma unroll
            for (int k = 0; k < 3; ++k) e[k] += dx[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) row[k] += dy[k];
    }
}

template <int TS, bool kVis>
__device__ __forceinline__ void raster_rect_self(const Setup& s, int x0, int y0, int rx0, int rx1, int ry0, int ry1,
                                                 typename KeyOf<kVis>::T* tile) {
    Edges E;
    edges_at(s, rx0, ry0, E);
    const long long bw = (long long)max(s.X[0], max(s.X[1], s.X[2])) - min(s.X[0], min(s.X[1], s.X[2]));
    const long long bh = (long long)max(s.Y[0], max(s.Y[1], s.Y[2])) - min(s.Y[0], min(s.Y[1], s.Y[2]));
    // |e| <= 2*bw*bh inside the box and one step past it adds at most
    // 256*max(bw, bh), so with these bounds every value stays below 2^30.
    if (bw * bh < (1ll << 28) && bw < (1 << 20) && bh < (1 << 20))
        raster_rect_self_t<TS, kVis, int>(s, E, x0, y0, rx0, rx1, ry0, ry1, tile);
    else
        raster_rect_self_t<TS, kVis, long long>(s, E, x0, y0, rx0, rx1, ry0, ry1, tile);
}

// The workgroup covers the rectangle in row groups of cw columns, cw the
// smallest power of two (8 to TS) not below the rectangle's width: thread t
// owns column rx0 + t % cw and every (kRasterThreads / cw)-th row from
// ry0 + t / cw, so a wave covers 64 / cw rows and at least half its lanes
// hold a pixel column.
template <int TS, bool kVis>
__device__ __forceinline__ void raster_rect_group(const Setup& s, int x0, int y0, int rx0, int rx1, int ry0, int ry1,
                                                  int tid, typename KeyOf<kVis>::T* tile) {
    const int rw = rx1 - rx0 + 1;
    int lg = 3;
    while ((1 << lg) < rw) ++lg;  // rw <= TS
    const int i = rx0 + (tid & ((1 << lg) - 1));
    if (i > rx1) return;
    int j = ry0 + (tid >> lg);
    if (j > ry1) return;
    const int row_step = kRasterThreads >> lg;
    if (rect_inside(s, rx0, rx1, ry0, ry1)) {
        // every pixel centre of the rectangle passes the three edge tests
#pragma unroll 4
        for (; j <= ry1; j += row_step)
            store_min<kVis>(s, plane_bits(s, i * 256 + 128 - s.X[0], j * 256 + 128 - s.Y[0]),
                            (j - y0) * TS + (i - x0), tile);
        return;
    }
    Edges E;
    edges_at(s, i, j, E);
    long long e[3] = {E.e[0], E.e[1], E.e[2]};
    long long step[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) step[k] = E.dy[k] * row_step;
    for (; j <= ry1; j += row_step) {
        plot<kVis, long long>(s, e, i, j, (j - y0) * TS + (i - x0), tile);
#pragma unroll
        for (int k = 0; k < 3; ++k) e[k] += step[k];
    }
}

// One (tile, map) of the raster.  kFallback: the map's lists overflowed the
// workspace, so every (triangle, piece) pair is visited with full clipping.
template <int TS, bool kVis, bool kFallback>
__device__ __forceinline__ void raster_tile(const MeshView& mesh, const MvpSource& mvps, const RasterCfg& cfg,
                                            const BinBufs& b, void* out, int map, int tile_id,
                                            typename KeyOf<kVis>::T* tile, Setup* rec, int& nrec) {
    const int tid = threadIdx.x;
    const int tx = tile_id % cfg.tiles_x, ty = tile_id / cfg.tiles_x;
    constexpr int kTilePx = TS * TS;
    const int x0 = tx * TS, y0 = ty * TS;
    const int x1 = min(x0 + TS - 1, cfg.w - 1), y1 = min(y0 + TS - 1, cfg.h - 1);
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    uint32_t begin = 0, end;
    if constexpr (!kFallback) {
        begin = b.offsets[(size_t)map * (ntiles + 1) + tile_id];
        end = b.offsets[(size_t)map * (ntiles + 1) + tile_id + 1];
    } else {
        end = (uint32_t)mesh.nt * 8u;
    }
    if constexpr (!kFallback && !kVis) {
        // A whole tile whose list is one wide record covering every pixel
        // centre (most tiles of a large ground plane): each depth is that
        // plane's, written straight from registers.  The LDS path would store
        // the same bits: min(clear, z) is z for every clamped depth.
        if (end - begin == 1 && x1 - x0 + 1 == TS && y1 - y0 + 1 == TS && (cfg.w & 3) == 0) {
            const uint32_t entry = b.entries[(size_t)map * b.cap + begin];
            if (entry & kWideBit) {
                const Setup& w = b.wide[(size_t)map * b.wcap + (entry & ~kWideBit)];
                const Setup s = w;
                if (s.px0 <= x0 && s.px1 >= x1 && s.py0 <= y0 && s.py1 >= y1 && rect_inside(s, x0, x1, y0, y1)) {
                    typedef float f4 __attribute__((ext_vector_type(4)));
                    float* dst = reinterpret_cast<float*>(out) + (size_t)map * ((size_t)cfg.w * cfg.h);
                    for (int q = tid; q < kTilePx / 4; q += kRasterThreads) {
                        const int r = q / (TS / 4), cidx = (q % (TS / 4)) * 4;
                        const int i = x0 + cidx, j = y0 + r;
                        const int dy = j * 256 + 128 - s.Y[0];
                        const f4 v = {__uint_as_float(plane_bits(s, i * 256 + 128 - s.X[0], dy)),
                                      __uint_as_float(plane_bits(s, (i + 1) * 256 + 128 - s.X[0], dy)),
                                      __uint_as_float(plane_bits(s, (i + 2) * 256 + 128 - s.X[0], dy)),
                                      __uint_as_float(plane_bits(s, (i + 3) * 256 + 128 - s.X[0], dy))};
                        __builtin_nontemporal_store(v, reinterpret_cast<f4*>(dst + (size_t)j * cfg.w + i));
                    }
                    return;
                }
            }
        }
    }
    if constexpr (!kVis) {  // 16-byte LDS writes
        constexpr uint32_t c = KeyOf<kVis>::clear;
        for (int q = tid; q < kTilePx / 4; q += kRasterThreads) reinterpret_cast<uint4*>(tile)[q] = make_uint4(c, c, c, c);
    } else {
        for (int p = tid; p < kTilePx; p += kRasterThreads) tile[p] = KeyOf<kVis>::clear;
    }
    if (tid == 0) nrec = 0;
    float mvp[16];
    mvps.load(map, mvp);
    __syncthreads();
    for (uint32_t base = begin; base < end; base += kRasterThreads) {
        uint32_t k = base + tid;
        if (k < end) {
            uint32_t entry = kFallback ? k : b.entries[(size_t)map * b.cap + k];
            if (!kFallback && (entry & kWideBit)) {
                const Setup& w = b.wide[(size_t)map * b.wcap + (entry & ~kWideBit)];
                int slot = atomicAdd(&nrec, 1);
                if (slot < kRecQueue) {
                    rec[slot] = w;
                } else {
                    int rx0 = max(x0, w.px0), rx1 = min(x1, w.px1), ry0 = max(y0, w.py0), ry1 = min(y1, w.py1);
                    if (rx0 <= rx1 && ry0 <= ry1) raster_rect_self<TS, kVis>(w, x0, y0, rx0, rx1, ry0, ry1, tile);
                }
            } else if constexpr (!kFallback && wide_span<TS>() <= 4) {
                // a directly binned triangle: the count pass stored its setup
                const Setup s = from_rec(b.setups[(size_t)map * b.tstride + (entry >> 3)], entry);
                int rx0 = max(x0, s.px0), rx1 = min(x1, s.px1), ry0 = max(y0, s.py0), ry1 = min(y1, s.py1);
                if (rx0 <= rx1 && ry0 <= ry1) {
                    int slot = kRecQueue;
                    if ((rx1 - rx0 + 1) * (ry1 - ry0 + 1) > kSelfRasterMaxPx) slot = atomicAdd(&nrec, 1);
                    if (slot < kRecQueue) rec[slot] = s;
                    else raster_rect_self<TS, kVis>(s, x0, y0, rx0, rx1, ry0, ry1, tile);
                }
            } else {
                Tri<0> t;
                bool ok;
                if constexpr (kFallback)
                    get_piece<0>(mesh, mvp, (int)(entry >> 3), (int)(entry & 7u), cfg, t, ok);
                else
                    classify<0>(mesh, mvp, (int)(entry >> 3), cfg, t, ok);
                if (ok) {
                    int rx0 = max(x0, t.px0), rx1 = min(x1, t.px1), ry0 = max(y0, t.py0), ry1 = min(y1, t.py1);
                    if (rx0 <= rx1 && ry0 <= ry1) {
                        Setup s = to_setup(t, entry);
                        int slot = kRecQueue;
                        if ((rx1 - rx0 + 1) * (ry1 - ry0 + 1) > kSelfRasterMaxPx) slot = atomicAdd(&nrec, 1);
                        if (slot < kRecQueue) rec[slot] = s;
                        else raster_rect_self<TS, kVis>(s, x0, y0, rx0, rx1, ry0, ry1, tile);
                    }
                }
            }
        }
        __syncthreads();
        const int nq = min(nrec, kRecQueue);
        for (int q = 0; q < nq; ++q) {
            const Setup s = rec[q];
            int rx0 = max(x0, s.px0), rx1 = min(x1, s.px1), ry0 = max(y0, s.py0), ry1 = min(y1, s.py1);
            if (rx0 > rx1 || ry0 > ry1) continue;
            raster_rect_group<TS, kVis>(s, x0, y0, rx0, rx1, ry0, ry1, tid, tile);
        }
        __syncthreads();
        if (tid == 0) nrec = 0;
        __syncthreads();
    }
    // Write the tile once.  Depth maps go out with non-temporal stores: a
    // frame writes far more map bytes than the caches hold (64 x 16 MiB), and
    // with ordinary stores the last-written lines stay dirty in the MALL and
    // are written back while the screen pass reads the maps.  On the
    // 64-sample 1080p frame this took the raster from 0.363 to 0.346 ms, the
    // soft pass that follows from 0.374 to 0.341 ms and the frame from 0.991
    // to 0.940 ms.
    const size_t plane = (size_t)cfg.w * cfg.h;
    if constexpr (!kVis) {
        float* dst = reinterpret_cast<float*>(out) + (size_t)map * plane;
        const bool full = (x1 - x0 + 1 == TS) && ((cfg.w & 3) == 0);
        if (full) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            for (int q = tid; q < kTilePx / 4; q += kRasterThreads) {
                int r = q / (TS / 4), cidx = (q % (TS / 4)) * 4;
                int j = y0 + r;
                if (j > y1) continue;
                const f4 v = {__uint_as_float(tile[r * TS + cidx + 0]), __uint_as_float(tile[r * TS + cidx + 1]),
                              __uint_as_float(tile[r * TS + cidx + 2]), __uint_as_float(tile[r * TS + cidx + 3])};
                __builtin_nontemporal_store(v, reinterpret_cast<f4*>(dst + (size_t)j * cfg.w + x0 + cidx));
            }
        } else {
            for (int p = tid; p < kTilePx; p += kRasterThreads) {
                int i = x0 + p % TS, j = y0 + p / TS;
                if (i > x1 || j > y1) continue;
                __builtin_nontemporal_store(__uint_as_float(tile[p]), dst + (size_t)j * cfg.w + i);
            }
        }
    } else {
        unsigned long long* dst = reinterpret_cast<unsigned long long*>(out) + (size_t)map * plane;
        for (int p = tid; p < kTilePx; p += kRasterThreads) {
            int i = x0 + p % TS, j = y0 + p / TS;
            if (i > x1 || j > y1) continue;
            dst[(size_t)j * cfg.w + i] = tile[p];
        }
    }
}

// Binned raster: one workgroup per (tile, map); overflowed maps are left to
// k_raster_fallback.  Workgroups go out tile-major (the map_count maps of
// tile 0, then of tile 1, ...).  The hardware deals workgroups to CUs in
// turn, about equally many to each, whatever their cost; in map-major order
// the costly tiles (those under the occluder, about 7x the time of a tile
// holding only the ground) arrive in runs that pile onto a few CUs while the
// rest idle.  Tile-major order spreads each costly tile's maps over the XCDs
// and CUs.
template <int TS, bool kVis>
__global__ __launch_bounds__(kRasterThreads) void k_raster(MeshView mesh, MvpSource mvps, RasterCfg cfg, BinBufs b,
                                                           void* out, int map_base, int map_count) {
    __shared__ __align__(16) typename KeyOf<kVis>::T tile[TS * TS];
    __shared__ Setup rec[kRecQueue];
    __shared__ int nrec;
    const int map = map_base + (int)(blockIdx.x % (unsigned)map_count);
    const int tile_id = (int)(blockIdx.x / (unsigned)map_count);
    if (b.overflow[map]) return;
    raster_tile<TS, kVis, false>(mesh, mvps, cfg, b, out, map, tile_id, tile, rec, nrec);
}

// One workgroup per tile, looping over the maps whose lists overflowed.
template <int TS, bool kVis>
__global__ __launch_bounds__(kRasterThreads) void k_raster_fallback(MeshView mesh, MvpSource mvps, RasterCfg cfg,
                                                                    BinBufs b, void* out, int nmaps) {
    __shared__ __align__(16) typename KeyOf<kVis>::T tile[TS * TS];
    __shared__ Setup rec[kRecQueue];
    __shared__ int nrec;
    // The usual case is no overflowed map: the threads read the flags
    // together and the workgroup returns after one barrier.
    bool any = false;
    for (int m = (int)threadIdx.x; m < nmaps; m += (int)blockDim.x) any |= b.overflow[m] != 0;
    if (!__syncthreads_or(any)) return;
    for (int map = 0; map < nmaps; ++map) {
        if (!b.overflow[map]) continue;
        raster_tile<TS, kVis, true>(mesh, mvps, cfg, b, out, map, blockIdx.x, tile, rec, nrec);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// G-buffer resolve: GBuffer.frag:35-37 (position, normal, gl_FrontFacing)
// ---------------------------------------------------------------------------
// Perspective-correct attributes of the visible piece at pixel k; false when
// the piece does not cover the pixel centre.
__device__ __forceinline__ bool resolve_attrs(const Tri<6>& t, int i, int j, float out[6]) {
    long long e[3];
    float z;
    if (!eval_pixel<6>(t, i, j, e, z)) return false;
    double b0 = (double)e[0] / (double)t.w[0];
    double b1 = (double)e[1] / (double)t.w[1];
    double b2 = (double)e[2] / (double)t.w[2];
    double den = (b0 + b1) + b2;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double num = (b0 * (double)t.a[0][q] + b1 * (double)t.a[1][q]) + b2 * (double)t.a[2][q];
        out[q] = (float)(num / den);
    }
    return true;
}

// Clipped pieces (rare) set up through the out-of-line clipper; kept apart so
// that the common path's triangle record stays in registers.
__device__ __noinline__ bool resolve_clipped(const MeshView& mesh, const float* mvp, int tri, int piece,
                                             const RasterCfg& cfg, int i, int j, float out[6], int& front) {
    Tri<6> t;
    bool ok;
    get_piece<6>(mesh, mvp, tri, piece, cfg, t, ok);
    front = t.front;
    return ok && resolve_attrs(t, i, j, out);
}

// facing_out (optional): gl_FrontFacing as 1 / 0 (GBuffer.frag:36), 0 for
// background.  The normal is stored negated on back faces.
__global__ __launch_bounds__(256) void k_gbuffer_resolve(MeshView mesh, Mat4 mvp_arg, RasterCfg cfg,
                                                         const unsigned long long* vis, float* pos_out,
                                                         float* nrm_out, float* facing_out) {
    size_t npx = (size_t)cfg.w * cfg.h;
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npx) return;
    unsigned long long key = vis[k];
    float out[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float facing = 0.0f;
    if (key != ~0ull) {
        uint32_t id = (uint32_t)(key & 0xffffffffull);
        const int tri = (int)(id >> 3), piece = (int)(id & 7u);
        const int i = (int)(k % cfg.w), j = (int)(k / cfg.w);
        Tri<6> t;
        bool ok;
        const int cls = classify<6>(mesh, mvp_arg.m, tri, cfg, t, ok);
        bool hit = false;
        int front = 0;
        if (cls == kInside) {
            hit = ok && piece == 0 && resolve_attrs(t, i, j, out);
            front = t.front;
        } else if (cls == kClip) {
            hit = resolve_clipped(mesh, mvp_arg.m, tri, piece, cfg, i, j, out, front);
        }
        if (hit) {
            facing = front ? 1.0f : 0.0f;
            if (!front) {
                out[3] = -out[3];
                out[4] = -out[4];
                out[5] = -out[5];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) out[q] = 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        pos_out[q * npx + k] = out[q];
        nrm_out[q * npx + k] = out[3 + q];
    }
    if (facing_out) facing_out[k] = facing;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side: workspace layout of the binned raster (shared by the single
// passes and the batched frames)
// ---------------------------------------------------------------------------
struct Workspace {
    BinBufs bins;
    void* vis;  // gbuffer only
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Tile-list entries per map: up to wide_span tiles per directly binned piece
// (pieces are rarely that wide, so 4 tiles per triangle at the 64-texel edge,
// scaled with the tile count for smaller edges) plus the wide pieces' tiles.
inline uint32_t entry_cap(int ntris, int ntiles, int ts = kTile) {
    const unsigned long long per = 4ull * (unsigned long long)(64 / ts) * (unsigned long long)(64 / ts);
    unsigned long long cap = per * (unsigned long long)ntris + 4ull * (unsigned long long)ntiles + 1024ull;
    if (cap > 0x7FFFFFFFull) cap = 0x7FFFFFFFull;
    return (uint32_t)cap;
}

inline uint32_t wide_cap(int ntris) { return 1024u + (uint32_t)ntris / 16u; }

inline size_t layout(int ntris, int nmaps, int w, int h, bool vis, Workspace* ws, char* base, int ts = kTile) {
    int tx = (w + ts - 1) / ts, ty = (h + ts - 1) / ts;
    size_t ntiles = (size_t)tx * ty;
    uint32_t cap = entry_cap(ntris, (int)ntiles, ts);
    uint32_t wcap = wide_cap(ntris);
    size_t off = 0;
    size_t o_counts = off; off = align256(off + sizeof(uint32_t) * ntiles * nmaps);
    size_t o_nwide = off; off = align256(off + sizeof(uint32_t) * nmaps);
    size_t o_wover = off; off = align256(off + sizeof(int) * nmaps);
    size_t o_nclip = off; off = align256(off + sizeof(uint32_t) * nmaps);
    size_t zero_end = off;  // [0, zero_end) is cleared before binning
    size_t o_offsets = off; off = align256(off + sizeof(uint32_t) * (ntiles + 1) * nmaps);
    size_t o_over = off; off = align256(off + sizeof(int) * nmaps);
    size_t o_wide = off; off = align256(off + sizeof(Setup) * (size_t)wcap * nmaps);
    size_t o_clip = off; off = align256(off + sizeof(uint32_t) * (size_t)wcap * nmaps);
    size_t o_entries = off; off = align256(off + sizeof(uint32_t) * (size_t)cap * nmaps);
    size_t o_tiles = off; off = align256(off + sizeof(uint32_t) * (size_t)(ntris > 0 ? ntris : 0) * nmaps);
    const bool records = 4 * (64 / ts) * (64 / ts) <= 4;  // the tile code's span (wide_span<ts>() <= 4)
    size_t o_setups = off;
    if (records) off = align256(off + sizeof(SetupRec) * (size_t)(ntris > 0 ? ntris : 0) * nmaps);
    size_t o_vis = off;
    if (vis) off = align256(off + sizeof(unsigned long long) * (size_t)w * h);
    if (ws) {
        BinBufs& b = ws->bins;
        b.counts = reinterpret_cast<uint32_t*>(base + o_counts);
        b.nwide = reinterpret_cast<uint32_t*>(base + o_nwide);
        b.wide_over = reinterpret_cast<int*>(base + o_wover);
        b.offsets = reinterpret_cast<uint32_t*>(base + o_offsets);
        b.overflow = reinterpret_cast<int*>(base + o_over);
        b.wide = reinterpret_cast<Setup*>(base + o_wide);
        b.nclip = reinterpret_cast<uint32_t*>(base + o_nclip);
        b.clip = reinterpret_cast<uint32_t*>(base + o_clip);
        b.entries = reinterpret_cast<uint32_t*>(base + o_entries);
        b.tiles = reinterpret_cast<uint32_t*>(base + o_tiles);
        b.setups = records ? reinterpret_cast<SetupRec*>(base + o_setups) : nullptr;
        b.tstride = (uint32_t)(ntris > 0 ? ntris : 0);
        b.cap = cap;
        b.wcap = wcap;
        ws->vis = vis ? reinterpret_cast<void*>(base + o_vis) : nullptr;
    }
    (void)zero_end;
    return off;
}

// Words to clear before binning: counts, nwide and wide_over are laid out
// contiguously at the start of the workspace.
inline size_t zero_words(int nmaps, const RasterCfg& cfg) {
    size_t ntiles = (size_t)cfg.tiles_x * cfg.tiles_y;
    size_t off = align256(sizeof(uint32_t) * ntiles * nmaps);
    off = align256(off + sizeof(uint32_t) * nmaps);
    off = align256(off + sizeof(int) * nmaps);
    off = align256(off + sizeof(uint32_t) * nmaps);
    return off / sizeof(uint32_t);
}

inline bool size_ok(int w, int h) { return w > 0 && h > 0 && w <= GI_MAX_EDGE && h <= GI_MAX_EDGE; }

}  // namespace
}  // namespace gi
