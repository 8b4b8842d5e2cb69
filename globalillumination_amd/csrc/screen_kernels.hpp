// screen_kernels.hpp — device side of the full-screen shadow passes (see
// screen.hip).  Included by screen.hip and batch.hip; internal linkage.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"

namespace gi {
namespace {


// A workgroup covers a 16x16-pixel tile made of four 8x8-pixel wave tiles.
// A wave's pixels then project to a compact patch of each shadow map (about
// 16x16 texels at the usual 2x minification), so one gather instruction
// touches a few cache lines; 64 pixels of one row would project along a
// slanted line and touch a line per lane.
constexpr int kBlockX = 16;
constexpr int kBlockY = 16;
constexpr int kBlockThreads = kBlockX * kBlockY;

// Pixel offset of thread `tid` inside its workgroup tile.
__device__ __forceinline__ void wave_tile_xy(int tid, int& px, int& py) {
    const int lane = tid & 63, wave = tid >> 6;
    px = (wave & 1) * 8 + (lane & 7);
    py = (wave >> 1) * 8 + (lane >> 3);
}

// Appends the threads with `want` of a workgroup to a list whose length is
// *count, one device-scope atomic per workgroup (those on one address
// serialize across the GPU, about 11 ns each, so one per wave costs tens of
// microseconds when thousands of waves append).  Returns each such thread's
// slot.  Every thread of the workgroup must call it; it holds two barriers.
// `lds` is two words of the workgroup's shared memory.
__device__ __forceinline__ unsigned block_append(bool want, unsigned* count, unsigned* lds) {
    if (threadIdx.x == 0) lds[0] = 0u;
    __syncthreads();
    const unsigned long long m = __ballot(want);
    const int lane = (int)(threadIdx.x & 63);
    const int leader = m ? __ffsll((long long)m) - 1 : 0;
    unsigned off = 0u;
    if (m && lane == leader) off = atomicAdd(&lds[0], (unsigned)__popcll(m));
    off = (unsigned)__shfl((int)off, leader, 64);
    __syncthreads();
    if (threadIdx.x == 0) lds[1] = lds[0] ? atomicAdd(count, lds[0]) : 0u;
    __syncthreads();
    return lds[1] + off + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
}

// Work distribution of the kernels that expand a list of pixels into lanes:
// wave gw of nw takes the groups of P entries starting at gw * P, (gw + nw) *
// P, ... until the list (length n) runs out.  A shared counter handing out
// groups would cost an atomic per group on one address, and those serialize
// across the GPU (about 11 ns each), slower than the groups themselves.
struct WaveGroups {
    unsigned n, step, base;
    __device__ WaveGroups(unsigned n_, unsigned P) : n(n_) {
        const unsigned wpb = blockDim.x / 64u;
        step = gridDim.x * wpb * P;
        base = (blockIdx.x * wpb + (threadIdx.x >> 6)) * P;
    }
    __device__ bool more() const { return base < n; }
    __device__ void advance() { base += step; }
};

// Sums `v` over the `cnt` lanes starting at lane g0 in lane order, one
// rounded addition at a time (`acc + v0 + v1 + ...` as a loop writes it).
__device__ __forceinline__ float ordered_sum(float acc, float v, int g0, int cnt) {
    for (int j = 0; j < cnt; ++j) acc = acc + __shfl(v, g0 + j, 64);
    return acc;
}

struct GBufferView {
    const float* __restrict__ pos;  // 3 planes
    const float* __restrict__ nrm;  // 3 planes (facing-flipped) or null
    int w, h;
};

struct HardUniforms {
    Mat4 light_mvp;
    Mat4 mv;
    Mat3 nm;
    float lpos[3];
    float si;
    int kernel_order;
    int penumbra;
    int max_search;
    float thr;
    float stepx, stepy;  // shadowMapStep (MyGLGeometryViewer.cpp:238)
    int pre_eval;
};

// Shadow.frag:222-238 (identical copies in the RBSM shaders :352-365).
__device__ __forceinline__ float pre_evaluation(const HardUniforms& u, float vx, float vy, float vz, float nx,
                                                float ny, float nz) {
    float ve[4];
    mat4_vec4(u.mv.m, vx, vy, vz, 1.0f, ve);
    float n3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float acc = u.nm.m[0 * 3 + i] * nx;
        acc = acc + u.nm.m[1 * 3 + i] * ny;
        acc = acc + u.nm.m[2 * 3 + i] * nz;
        n3[i] = acc;
    }
    float N[3], L[3];
    normalize3(n3, N);
    float ld[3] = {u.lpos[0] - ve[0], u.lpos[1] - ve[1], u.lpos[2] - ve[2]};
    normalize3(ld, L);
    float d = N[0] * L[0] + N[1] * L[1] + N[2] * L[2];
    return (gmax(d, 0.0f) == 0.0f) ? u.si : 1.0f;
}

// Shadow.frag:86-116 with bilinearPCF == 1.
__device__ float pcf(const DepthMap& D, const HardUniforms& u, float s, float t, float z) {
    float incrWidth = 1.0f / D.fw;
    float incrHeight = 1.0f / D.fh;
    float illum = 0.0f;
    float offset = (float)u.penumbra;
    float stepSize = 2.0f * offset / (float)u.kernel_order;
    int count = 0;
    for (float w = -offset; w < offset; w += stepSize) {
        float sw = s + w * incrWidth;
        for (float h = -offset; h < offset; h += stepSize) {
            float d = D.fetch(sw, t + h * incrHeight);
            illum = illum + ((z <= d) ? 1.0f : u.si);
            count++;
        }
    }
    return illum / (float)count;
}

// ---------------------------------------------------------------------------
// Revectorization (RBSM).  `kFiltered` selects FilteredRBSM.frag's
// estimateRelativePosition (:234-243) and revectorizeShadow (:266-272).
// ---------------------------------------------------------------------------
template <bool kFiltered>
struct Rbsm {
    DepthMap D;
    float sx, sy;  // shadowMapStep
    float thr;
    int max_search;

    __device__ __forceinline__ float tex(float s, float t) const { return D.fetch(s, t); }

    // NonConservativeSMSR.frag:56-94
    __device__ bool disc_f(float x, float y, float z, float dirx, float diry, float type) const {
        if (dirx == 0.0f) {
            x = x - sx;
            if (fabsf(le01(z, tex(x, y)) - type) == 0.0f) return true;
            x = x + 2.0f * sx;
            if (fabsf(le01(z, tex(x, y)) - type) == 0.0f) return true;
            x = x - sx;
        }
        if (diry == 0.0f) {
            y = y + sy;
            if (fabsf(le01(z, tex(x, y)) - type) == 0.0f) return true;
            y = y - 2.0f * sy;
            if (fabsf(le01(z, tex(x, y)) - type) == 0.0f) return true;
        }
        return false;
    }

    // NonConservativeSMSR.frag:96-178; writes the shader global `newDepth`.
    __device__ bool disc_v(float x, float y, float z, float dirx, float diry, float dr, float dg, float db,
                           float& newDepth) const {
        float rx = x, ry = y;
        newDepth = z;
        if (dirx == 0.0f) {
            if (dr == 0.5f || dr == 0.75f) {
                rx = x - sx;
                float d = tex(rx, ry);
                if (db == 1.0f && fabsf(z - d) < thr) { z = z - thr; newDepth = z; }
                if (fabsf(le01(z, d) - db) == 0.0f) return true;
            }
            if (dr == 0.75f || dr == 0.25f) {
                rx = x + sx;
                float d = tex(rx, ry);
                if (db == 1.0f && fabsf(z - d) < thr) { z = z - thr; newDepth = z; }
                if (fabsf(le01(z, d) - db) == 0.0f) return true;
            }
        }
        if (diry == 0.0f) {
            if (dg == 0.5f || dg == 0.75f) {
                ry = y + sy;
                float d = tex(rx, ry);
                if (db == 1.0f && fabsf(z - d) < thr) { z = z - thr; newDepth = z; }
                if (fabsf(le01(z, d) - db) == 0.0f) return true;
            }
            if (dg == 0.75f || dg == 0.25f) {
                ry = y - sy;
                float d = tex(rx, ry);
                if (db == 1.0f && fabsf(z - d) < thr) { z = z - thr; newDepth = z; }
                if (fabsf(le01(z, d) - db) == 0.0f) return true;
            }
        }
        return false;
    }

    // NonConservativeSMSR.frag:180-232.  A step's addresses do not depend on
    // the texels (only the depth nudges and the exit do), so each step
    // fetches its texel and the neighbours disc_v() may test together and
    // then runs disc_v's tests on them in disc_v's order: one wait per step
    // where disc_v's early exits made up to three dependent fetches.  Across
    // a vertical walk (dirx == 0) disc_v reads x -/+ sx, selected by dr;
    // across a horizontal one (diry == 0) y +/- sy, selected by dg.  The end
    // of edge test (disc_f) runs after the loop.  Written for K steps per
    // round; one step per round measured fastest (RPCF, 3 taps per axis, on a
    // 1920x1080 view of a 200k-triangle soup and a 2048^2 map: 0.150 ms
    // against 0.161 ms for two steps and 0.184 ms for four, which hold more
    // registers, and 0.167 ms for disc_v's own fetches).
    __device__ float length(float dr, float dg, float db, float x, float y, float z, float dirx, float diry,
                            float sub) const {
        return length_k<1>(dr, dg, db, x, y, z, dirx, diry, sub);
    }
    template <int K>
    __device__ float length_k(float dr, float dg, float db, float x, float y, float z, float dirx, float diry,
                              float sub) const {
        float cx = x, cy = y, cz = z;
        float found = 0.0f, dist = 0.0f;
        const float stx = dirx * sx, sty = diry * sy;
        cx = cx + stx;
        cy = cy + sty;
        const bool vert = dirx == 0.0f, horiz = diry == 0.0f;
        const bool na = vert ? (dr == 0.5f || dr == 0.75f) : (horiz && (dg == 0.5f || dg == 0.75f));
        const bool nb = vert ? (dr == 0.75f || dr == 0.25f) : (horiz && (dg == 0.75f || dg == 0.25f));
        bool go = true, edge_end = false;
        for (int it0 = 0; go && it0 < max_search; it0 += K) {
            float dv[K], av[K], bv[K];
            float px = cx, py = cy;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                dv[i] = tex(px, py);
                av[i] = 0.0f;
                bv[i] = 0.0f;
                if (vert) {
                    if (na) av[i] = tex(px - sx, py);
                    if (nb) bv[i] = tex(px + sx, py);
                } else {
                    if (na) av[i] = tex(px, py + sy);
                    if (nb) bv[i] = tex(px, py - sy);
                }
                px = px + stx;
                py = py + sty;
            }
#pragma unroll
            for (int i = 0; i < K; ++i) {
                if (!go || it0 + i >= max_search) {
                    go = false;
                    break;
                }
                const float d = dv[i];
                if (db == 0.0f && fabsf(cz - d) < thr) cz = cz - thr;
                const float center = le01(cz, d);
                if (fabsf(center - db) == 0.0f) {
                    edge_end = true;
                    go = false;
                    break;
                }
                // disc_v (:96-178) on the fetched neighbours
                float zz = cz, newDepth = cz;
                bool cont = false;
                if (na) {
                    const float dd = av[i];
                    if (db == 1.0f && fabsf(zz - dd) < thr) { zz = zz - thr; newDepth = zz; }
                    if (fabsf(le01(zz, dd) - db) == 0.0f) cont = true;
                }
                if (!cont && nb) {
                    const float dd = bv[i];
                    if (db == 1.0f && fabsf(zz - dd) < thr) { zz = zz - thr; newDepth = zz; }
                    if (fabsf(le01(zz, dd) - db) == 0.0f) cont = true;
                }
                if (!cont) {
                    go = false;
                    break;
                }
                dist = dist + 1.0f;
                cx = cx + stx;
                cy = cy + sty;
                if (db == 1.0f) cz = newDepth;
            }
        }
        if (edge_end) found = disc_f(cx, cy, cz, 0.0f, 0.0f, db) ? 1.0f : 0.0f;
        float a = -(dist + (1.0f - sub));
        float b = dist + (1.0f - sub);
        return gmix(a, b, found);
    }

    // estimateRelativePosition :234-243
    __device__ float rel(float dx, float dy, float shadow) const {
        float T = 1.0f;
        if (dx < 0.0f && dy < 0.0f) T = 0.0f;
        if (dx > 0.0f && dy > 0.0f) T = -2.0f;
        float edge = gmin(fabsf(dx) + fabsf(dy), (float)max_search);
        if (kFiltered) return (T * fabsf(gmax(T * dx, T * dy))) / edge;
        return (gmax(T, 2.0f * shadow - 1.0f) * fabsf(gmax(T * dx, T * dy))) / edge;
    }

    // revectorizeShadow :266-274
    __device__ float revectorize(float rx, float ry, float s) const {
        if (kFiltered) {
            if (rx * ry < 0.0f) return (1.0f - s) + (2.0f * s - 1.0f) * gmax(rx, ry);
            if (rx * ry == 0.0f) return s;
            return gclamp((1.0f - s) + (2.0f * s - 1.0f) * (rx + ry), 0.0f, 1.0f);
        }
        if ((rx * ry == 2.0f * s) ||
            ((fabsf(rx) * fabsf(ry) > 0.0f) && ((1.0f - s) + (2.0f * s - 1.0f) * (fabsf(rx) + fabsf(ry)) < 0.5f)))
            return 0.0f;
        return 1.0f;
    }

    // computeDiscontinuity :276-286 (getDisc :23-54 without early outs)
    __device__ void discontinuity(float x, float y, float z, float dC, float& r, float& g, float& b) const {
        float center = le01(z, dC);
        float xx = x - sx;
        float dl = le01(z, tex(xx, y));
        xx = xx + 2.0f * sx;
        float drt = le01(z, tex(xx, y));
        xx = xx - sx;
        float yy = y + sy;
        float db = le01(z, tex(xx, yy));
        yy = yy - 2.0f * sy;
        float dt = le01(z, tex(xx, yy));
        r = (2.0f * fabsf(dl - center) + fabsf(drt - center)) / 4.0f;
        g = (2.0f * fabsf(db - center) + fabsf(dt - center)) / 4.0f;
        b = 1.0f - center;
    }

    // orientateDS :245-254 -> estimateRelativePosition :256-264 -> revectorizeShadow
    __device__ float revectorized(float dr, float dg, float db, float x, float y, float z, float subx, float suby,
                                  float shadow) const {
        float left = length(dr, dg, db, x, y, z, -1.0f, 0.0f, 1.0f - subx);
        float right = length(dr, dg, db, x, y, z, 1.0f, 0.0f, subx);
        float down = length(dr, dg, db, x, y, z, 0.0f, -1.0f, 1.0f - suby);
        float up = length(dr, dg, db, x, y, z, 0.0f, 1.0f, suby);
        return revectorize(rel(left, right, shadow), rel(down, up, shadow), shadow);
    }

    // computeShadowFromSMSR :288-304
    __device__ float smsr(float x, float y, float z, float si) const {
        float dC = tex(x, y);
        float r, g, b;
        discontinuity(x, y, z, dC, r, g, b);
        float subx = gfract(x * D.fw);
        float suby = gfract(y * D.fh);
        float shadow = le01(z, dC);
        if (r > 0.0f || g > 0.0f) {
            if (r == 0.75f && g == 0.75f) return gmix(1.0f - shadow, 1.0f, si);
            return gmix(revectorized(r, g, b, x, y, z, subx, suby, shadow), 1.0f, si);
        }
        return gmix(shadow, 1.0f, si);
    }

    // An RPCF tap up to its edge walks: coordinates, depth test and
    // discontinuity.  Returns true where the tap sits on a discontinuity (the
    // walks of revectorized() then decide it, from the fields of `t`), else
    // leaves the tap's value in `v`.
    struct Tap {
        float px, py, r, g, b, shadow;
    };
    __device__ bool rpcf_front(float x, float y, float z, float si, float w, float h, Tap& t, float& v) const {
        const float incrWidth = 1.0f / D.fw;
        const float incrHeight = 1.0f / D.fh;
        t.px = x + w * incrWidth;
        t.py = y + h * incrHeight;
        float d = tex(t.px, t.py);
        t.shadow = le01(z, d);
        discontinuity(t.px, t.py, z, d, t.r, t.g, t.b);
        if (t.r > 0.0f || t.g > 0.0f) return true;
        v = gmix(t.shadow, 1.0f, si);
        return false;
    }
    // A tap's r and g take the values k/4 (k < 4) and b and shadow 0 or 1,
    // all exact in the 6 bits this code keeps; from_code() gives them back
    // bit for bit.
    __device__ static unsigned tests_code(const Tap& t) {
        return (unsigned)(t.r * 4.0f) | (unsigned)(t.g * 4.0f) << 2 | (unsigned)t.b << 4 | (unsigned)t.shadow << 5;
    }
    __device__ static Tap from_code(float px, float py, unsigned c) {
        return Tap{px, py, (float)(c & 3u) * 0.25f, (float)(c >> 2 & 3u) * 0.25f, (float)(c >> 4 & 1u),
                   (float)(c >> 5 & 1u)};
    }
    // The edge walks of a tap rpcf_front() found on a discontinuity.
    __device__ float rpcf_walks(const Tap& t, float z) const {
        return revectorized(t.r, t.g, t.b, t.px, t.py, z, gfract(t.px * D.fw), gfract(t.py * D.fh), t.shadow);
    }

    // One tap of computeShadowFromAccurateRPCF's loop (:313-345) at offsets
    // (w, h) in texels: revectorized where the tap sits on a discontinuity,
    // mixed with si.
    __device__ float rpcf_tap(float x, float y, float z, float si, float w, float h) const {
        Tap t;
        float v;
        if (rpcf_front(x, y, z, si, w, h, t, v)) return gmix(rpcf_walks(t, z), 1.0f, si);
        return v;
    }
    // computeShadowFromAccurateRPCF's loop (:306-349) runs over the pixel
    // list: k_hard_classify, then k_hard_taps / k_hard_taps_lanes.
};

// ---------------------------------------------------------------------------
// Conservative revectorization: ConservativeSMSR.frag, the shader
// computeHardShadows binds when ShadowParams.conservative is set
// (ShadowMapping/src/main.cpp:406-408).  Binary neighbour tests and walks that
// stop at the first umbra texel; the result is si or 1 per tap.
// ---------------------------------------------------------------------------
struct RbsmConservative {
    DepthMap D;
    float sx, sy;  // shadowMapStep
    float thr;
    int max_search;
    float si;

    __device__ __forceinline__ float tex(float s, float t) const { return D.fetch(s, t); }

    // computeDiscontinuity :23-48, abs(dir - 1): 1 where the neighbour is in umbra.
    // Returns left + right (in *lr) and the four-way sum.
    __device__ float disc(float x, float y, float z, float* lr) const {
        float xx = x - sx;
        const float l = fabsf(le01(z, tex(xx, y)) - 1.0f);
        xx = xx + 2.0f * sx;
        const float r = fabsf(le01(z, tex(xx, y)) - 1.0f);
        xx = xx - sx;
        float yy = y + sy;
        const float b = fabsf(le01(z, tex(xx, yy)) - 1.0f);
        yy = yy - 2.0f * sy;
        const float t = fabsf(le01(z, tex(xx, yy)) - 1.0f);
        if (lr) {
            lr[0] = l;
            lr[1] = r;
            lr[2] = b;
            lr[3] = t;
        }
        return ((l + r) + b) + t;
    }

    // computeRelativeDistance(vec4, vec2, float) :50-86
    __device__ float walk(float x, float y, float z, float dirx, float diry, float c) const {
        const float stx = dirx * sx, sty = diry * sy;
        float cx = x + stx, cy = y + sty, cz = z;
        float found = 0.0f, dist = 0.0f;
        for (int it = 0; it < max_search; ++it) {
            const float d = tex(cx, cy);
            if (fabsf(cz - d) < thr) cz = cz - thr;
            if (le01(cz, d) == 0.0f) {  // isCenterUmbra
                found = 1.0f;
                break;
            }
            if (disc(cx, cy, cz, nullptr) == 0.0f) break;
            dist = dist + 1.0f;
            cx = cx + stx;
            cy = cy + sty;
        }
        dist = dist + (1.0f - c);
        return gmix(-dist, dist, found);
    }

    // normalizeRelativeDistance(vec2) :99-108
    __device__ float norm(float a, float b) const {
        float T = 1.0f;
        if (a < 0.0f && b < 0.0f) T = 0.0f;
        if (a > 0.0f && b > 0.0f) T = -2.0f;
        const float len = gmin(fabsf(a) + fabsf(b), (float)max_search);
        return fabsf(gmax(T * a, T * b)) / len;
    }

    // computeRelativeDistance(vec4, vec2) :88-97 -> normalizeRelativeDistance
    // :110-118 -> revectorizeShadow :120-126
    __device__ float revectorized(float x, float y, float z, float cx, float cy) const {
        const float dl = walk(x, y, z, -1.0f, 0.0f, 1.0f - cx);
        const float dr = walk(x, y, z, 1.0f, 0.0f, cx);
        const float db = walk(x, y, z, 0.0f, -1.0f, 1.0f - cy);
        const float dt = walk(x, y, z, 0.0f, 1.0f, cy);
        const float rx = norm(dl, dr), ry = norm(db, dt);
        if ((rx * ry > 0.0f) && (1.0f - rx > ry)) return si;
        return 1.0f;
    }

    // computeShadowFromSMSR :128-144
    __device__ float smsr(float x, float y, float z) const {
        if (le01(z, tex(x, y)) == 0.0f) return si;
        float lr[4];
        if (disc(x, y, z, lr) == 0.0f) return 1.0f;
        if (lr[0] + lr[1] == 2.0f || lr[2] + lr[3] == 2.0f) return si;
        return revectorized(x, y, z, gfract(x * D.fw), gfract(y * D.fh));
    }

    // One tap of computeShadowFromRPCF's loop (:146-198) at offsets (w, h)
    // in texels.
    __device__ float rpcf_tap(float x, float y, float z, float w, float h) const {
        const float incrWidth = 1.0f / D.fw;
        const float incrHeight = 1.0f / D.fh;
        const float px = x + w * incrWidth;
        const float py = y + h * incrHeight;
        const float shadow = (z <= tex(px, py)) ? 1.0f : si;
        if (shadow == 1.0f) {
            float lr[4];
            disc(px, py, z, lr);
            if (lr[0] == 0.0f && lr[1] == 0.0f) return 1.0f;
            return revectorized(px, py, z, gfract(px * D.fw), gfract(py * D.fh));
        }
        return si;
    }
    // computeShadowFromRPCF's loop (:146-198) runs over the pixel list:
    // k_hard_classify, then k_hard_taps_lanes.
};

// Pixel index with an XCD-aware block order.  Hardware hands consecutive
// workgroups to the 8 XCDs round-robin; block b is remapped so that each XCD
// drains one contiguous range of a strip-major order: 8 * kSpx vertical
// strips of equal width (the last ones narrower or empty), dealt to the XCDs
// in turn (XCD k drains strips k, k + 8, ...), row-major inside each strip.
// Neighbouring pixels' shadow-map texels then stay in one XCD's L2.  One strip
// per XCD (kSpx = 1) suits the passes that stream the G-buffer: a strip edge
// splits the 128-byte lines of a pixel row between two XCDs, and wide strips
// have few edges.  Passes whose cost per pixel varies across the view (area-
// light and filtered revectorized passes) take more, narrower strips per XCD,
// so each XCD samples several parts of the view and the XCDs finish together.
// Block (bx, by) at position lin of that order on a bx_n x by_n grid:
template <int kSpx>
__device__ __forceinline__ void strip_block(int lin, int bx_n, int by_n, int& bx, int& by) {
    constexpr int kStrips = 8 * kSpx;
    const int sw = (bx_n + kStrips - 1) / kStrips;
    int acc = 0;
    bx = 0;
    by = 0;
    for (int o = 0; o < kStrips; ++o) {
        const int s = (o % kSpx) * 8 + o / kSpx;
        const int x0 = s * sw;
        const int wk = x0 >= bx_n ? 0 : min(sw, bx_n - x0);
        if (lin < acc + wk * by_n) {
            const int q = lin - acc;
            bx = x0 + q % wk;
            by = q / wk;
            return;
        }
        acc += wk * by_n;
    }
}

// Position of this workgroup in the strip-major order: XCD k (block b with
// b % 8 == k) drains the k-th of 8 contiguous ranges.
__device__ __forceinline__ int xcd_order(int nblocks) {
    const int b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    const int per = nblocks / 8;
    const int rem = nblocks - per * 8;
    const int xcd = b % 8, slot = b / 8;
    const int lin = (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
    return lin < nblocks ? lin : b;
}

template <int kSpx = 1>
__device__ __forceinline__ void pixel_of_block(int w, int h, int& x, int& y) {
    const int bx_n = (w + kBlockX - 1) / kBlockX;
    const int by_n = (int)gridDim.y;
    int bx, by, px, py;
    strip_block<kSpx>(xcd_order(bx_n * by_n), bx_n, by_n, bx, by);
    wave_tile_xy((int)threadIdx.x, px, py);
    x = bx * kBlockX + px;
    y = by * kBlockY + py;
}

// Strips per XCD of the passes whose per-pixel cost varies most: filtered
// revectorization and PCSS (2); the area-light pass has its own block shape
// and 4 strips (soft_block_pixel).
constexpr int kSpxFiltered = 2;

// The techniques without a tap loop of revectorized taps (naive, PCF, RBSM,
// conservative RBSM): a thread per pixel.  RPCF, FilteredRBSM and
// conservative RPCF run over a pixel list (k_hard_classify).
template <int kTech>
__global__ __launch_bounds__(kBlockThreads) void k_hard(GBufferView gb, DepthMap D, HardUniforms u,
                                                            float* __restrict__ out) {
    static_assert(kTech == GI_NAIVE || kTech == GI_PCF || kTech == GI_RBSM || kTech == GI_RBSM_CONSERVATIVE,
                  "the listed techniques run k_hard_classify");
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    float vx = gb.pos[k];
    if (vx == 0.0f) {  // Shadow.frag:244 discards; the cleared target holds 0
        out[k] = 0.0f;
        return;
    }
    float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    float sc[4];
    mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
    float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
    float shadow = 1.0f;
    if (u.pre_eval)
        shadow = pre_evaluation(u, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
    if constexpr (kTech == GI_NAIVE || kTech == GI_PCF) {
        if (sc[3] > 0.0f && shadow == 1.0f) {
            if constexpr (kTech == GI_NAIVE) {
                float d = D.fetch(n0, n1);
                shadow = (n2 <= d) ? 1.0f : u.si;
            } else {
                shadow = pcf(D, u, n0, n1, n2);
            }
        }
    } else if constexpr (kTech == GI_RBSM_CONSERVATIVE) {
        // ConservativeSMSR.frag:215-235 (no shadowCoord.w test)
        RbsmConservative rb{D, u.stepx, u.stepy, u.thr, u.max_search, u.si};
        if (shadow == 1.0f) shadow = rb.smsr(n0, n1, n2);
    } else {
        Rbsm<false> rb{D, u.stepx, u.stepy, u.thr, u.max_search};
        if (shadow == 1.0f) shadow = rb.smsr(n0, n1, n2, u.si);
    }
    out[k] = shadow;
}

// Offsets of the RPCF tap loop `for (float w = -offset; w <= offset; w +=
// stepSize)` (NonConservativeSMSR.frag:313-314, FilteredRBSM.frag:313-314):
// the count per axis, and the offset after i steps, accumulated one addition
// at a time as the loop does.
__host__ __device__ inline int rpcf_side(const HardUniforms& u) {
    const float offset = (float)u.penumbra, stepSize = 2.0f * offset / (float)u.kernel_order;
    int n = 0;
    for (float w = -offset; w <= offset; w += stepSize) ++n;
    return n;
}
__device__ __forceinline__ float rpcf_offset(const HardUniforms& u, int i) {
    const float offset = (float)u.penumbra, stepSize = 2.0f * offset / (float)u.kernel_order;
    float w = -offset;
    for (int k = 0; k < i; ++k) w += stepSize;
    return w;
}

// Workspace of the RPCF and FilteredRBSM passes: the list length (in a
// 256-byte header), then one pixel index per view pixel.
constexpr size_t kHardListHeader = 256;
inline size_t hard_list_bytes(int w, int h) { return kHardListHeader + (size_t)w * h * 4u; }

// The pixels the pre-evaluation lets through run RPCF's tap loop; the others
// get their value here (k_hard's branches), and the loop's pixels go to the
// list for k_hard_taps.
template <int kTech>
__global__ __launch_bounds__(kBlockThreads) void k_hard_classify(GBufferView gb, HardUniforms u,
                                                                  float* __restrict__ out, unsigned* __restrict__ count,
                                                                  unsigned* __restrict__ list) {
    int x, y;
    pixel_of_block<kSpxFiltered>(gb.w, gb.h, x, y);
    bool loop = false;
    size_t k = 0;
    if (x < gb.w && y < gb.h) {
        const size_t plane = (size_t)gb.w * gb.h;
        k = (size_t)y * gb.w + x;
        const float vx = gb.pos[k];
        float shadow = 0.0f;  // Shadow.frag:244 discards; the cleared target holds 0
        if (vx != 0.0f) {
            shadow = 1.0f;
            if (u.pre_eval) {
                const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
                shadow = pre_evaluation(u, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
            }
            loop = shadow == 1.0f;
        }
        if (!loop) out[k] = shadow;
    }
    __shared__ unsigned lds[2];
    const unsigned slot = block_append(loop, count, lds);
    if (loop) list[slot] = (unsigned)k;
}

// computeShadowFromAccurateRPCF (:306-349) for the listed pixels, a lane per
// tap.  The taps are summed in the loop's order (w, then h) and divided by
// their count, as the shader does.
//
// A lane per tap without packing: 64 / taps pixels per wave group when the
// loop has at most 64 taps, else one pixel in passes of 64 taps.  Serves
// RPCF and FilteredRBSM loops of more than 64 taps, and ConservativeSMSR's
// RPCF (RbsmConservative::rpcf_tap) at any size.
template <int kTech>
__device__ __forceinline__ float hard_tap(const DepthMap& D, const HardUniforms& u, float x, float y, float z,
                                          float w, float h) {
    if constexpr (kTech == GI_RPCF_CONSERVATIVE) {
        const RbsmConservative rb{D, u.stepx, u.stepy, u.thr, u.max_search, u.si};
        return rb.rpcf_tap(x, y, z, w, h);
    } else {
        const Rbsm<kTech == GI_RBSM_FILTERED> rb{D, u.stepx, u.stepy, u.thr, u.max_search};
        return rb.rpcf_tap(x, y, z, u.si, w, h);
    }
}
template <int kTech>
__global__ __launch_bounds__(kBlockThreads) void k_hard_taps_lanes(GBufferView gb, DepthMap D, HardUniforms u,
                                                                    float* __restrict__ out,
                                                                    const unsigned* __restrict__ count,
                                                                    const unsigned* __restrict__ list) {
    const unsigned n = *count;
    const int side = rpcf_side(u);
    const int ntaps = side * side;
    const int per = ntaps < 64 ? ntaps : 64;
    const int P = 64 / per;
    const int lane = (int)(threadIdx.x & 63);
    const int slot = lane / per, sl = lane - slot * per;
    const int g0 = (slot < P ? slot : P - 1) * per;
    const size_t plane = (size_t)gb.w * gb.h;
    for (WaveGroups wg(n, (unsigned)P); wg.more(); wg.advance()) {
        const unsigned qi = wg.base + (unsigned)slot;
        const bool has = slot < P && qi < n;
        const size_t k = list[has ? qi : wg.base];
        const float vx = gb.pos[k], vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
        float sc[4];
        mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
        const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
        float illum = 0.0f;
        for (int t0 = 0; t0 < ntaps; t0 += per) {
            const int t = t0 + sl;
            float v = 0.0f;
            if (has && t < ntaps) {
                const int tw = t / side;
                v = hard_tap<kTech>(D, u, n0, n1, n2, rpcf_offset(u, tw), rpcf_offset(u, t - tw * side));
            }
            illum = ordered_sum(illum, v, g0, ntaps - t0 < per ? ntaps - t0 : per);
        }
        if (has && sl == 0) out[k] = illum / (float)ntaps;
    }
}

// Loops of at most 64 taps: 64 / taps pixels per wave group, and a wave takes
// three groups at a time.  Every lane runs its three taps up to the edge
// walks, then the taps that sit on a discontinuity (about half of them on a
// 200k-triangle soup) are packed into consecutive lanes through the wave's
// LDS area and walked together: two passes of 64 walks for the three groups
// where each group alone would leave its lanes without a discontinuity idle
// through a pass of its own.  The walks get the same inputs, so every tap's
// value is the same; the sums run over the taps in loop order as before.
// The whole pass on that soup (RPCF, 3 taps per axis, 1920x1080): one group
// per pass 0.180 ms, two 0.175, three 0.168, four 0.172 (70 VGPRs, 7 waves
// per SIMD).  The tests travel as 6 bits (tests_code) so that three groups'
// records fit 8 waves per SIMD.
constexpr int kTapPack = 3;  // groups packed per wave pass
template <int kTech>
__global__ __launch_bounds__(kBlockThreads) void k_hard_taps(GBufferView gb, DepthMap D, HardUniforms u,
                                                              float* __restrict__ out,
                                                              const unsigned* __restrict__ count,
                                                              const unsigned* __restrict__ list) {
    using R = Rbsm<kTech == GI_RBSM_FILTERED>;
    // per wave: the packed walks' inputs (px, py, z and the tap's tests as
    // bits); the result replaces px
    __shared__ float pk[kBlockThreads / 64][4][64 * kTapPack];
    const unsigned n = *count;
    const int side = rpcf_side(u);
    const int per = side * side;  // at most 64: the host launches k_hard_taps_lanes beyond
    if (per > 64) return;
    const int P = 64 / per;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int slot = lane / per, sl = lane - slot * per;
    const int g0 = (slot < P ? slot : P - 1) * per;
    const size_t plane = (size_t)gb.w * gb.h;
    const R rb{D, u.stepx, u.stepy, u.thr, u.max_search};
    const int tw = sl / side;
    const float ow = rpcf_offset(u, tw), oh = rpcf_offset(u, sl - tw * side);
    float(*buf)[64 * kTapPack] = pk[wave];
    WaveGroups wg(n, (unsigned)P);
    while (wg.more()) {
        bool has[kTapPack], disc[kTapPack];
        unsigned kk[kTapPack];
        float v[kTapPack];
        int rank[kTapPack], total = 0;
#pragma unroll
        for (int r = 0; r < kTapPack; ++r) {
            const unsigned base = wg.base;
            const unsigned qi = base + (unsigned)slot;
            has[r] = base < n && slot < P && qi < n;
            disc[r] = false;
            v[r] = 0.0f;
            kk[r] = has[r] ? list[qi] : 0u;
            typename R::Tap tp;
            float z = 0.0f;
            if (has[r]) {
                const size_t k = kk[r];
                const float vx = gb.pos[k], vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
                float sc[4];
                mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
                const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3];
                z = sc[2] / sc[3];
                disc[r] = rb.rpcf_front(n0, n1, z, u.si, ow, oh, tp, v[r]);
            }
            // pack the taps on a discontinuity, ranked over the groups in order
            const unsigned long long m = __ballot(disc[r]);
            rank[r] = total + (int)__popcll(m & ((1ull << lane) - 1ull));
            total += (int)__popcll(m);
            if (disc[r]) {
                const int j = rank[r];
                buf[0][j] = tp.px; buf[1][j] = tp.py; buf[2][j] = z;
                buf[3][j] = __uint_as_float(R::tests_code(tp));
            }
            wg.advance();
        }
        // LDS within one wave: the fences keep the compiler from moving the
        // other lanes' accesses across these points
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int j0 = 0; j0 < total; j0 += 64) {
            const int j = j0 + lane;
            if (j < total) {
                const typename R::Tap t = R::from_code(buf[0][j], buf[1][j], __float_as_uint(buf[3][j]));
                buf[0][j] = gmix(rb.rpcf_walks(t, buf[2][j]), 1.0f, u.si);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < kTapPack; ++r) {
            if (disc[r]) v[r] = buf[0][rank[r]];
            const float illum = ordered_sum(0.0f, v[r], g0, per);
            if (has[r] && sl == 0) out[kk[r]] = illum / (float)per;
        }
        // the next pass writes only after every lane's reads above
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// Percentage-closer soft shadows: PlausibleSoftShadow.frag with PCSS == 1
// (renderSoftShadows, SoftShadowMapping/src/main.cpp:925-1016).  One depth
// map from the light centre; a blocker search sets the PCF kernel's width.
// ---------------------------------------------------------------------------
struct PcssUniforms {
    float bsw;     // blockerSearchWidth (:171-173), formed on the host in FP32
    float bfw;     // (blockerSearchSize - 1.0) * 0.5
    float kfw;     // (kernelSize - 1.0) * 0.5
    float radius;  // float(lightSourceRadius)
    float znear;   // float(zNear), an int uniform
    int ks;        // kernelSize
};

// The loops `for (int h = -filterWidth; h <= filterWidth; h++)` start at the
// float bound converted to int (truncation toward zero) and compare as float.
__device__ __forceinline__ int loop_start(float fw) { return (int)(-fw); }

// computeAverageBlockerDepthBasedOnPCF :166-192
__device__ __forceinline__ float pcss_blocker(const DepthMap& D, const PcssUniforms& q, float x, float y, float z) {
    float avg = 0.0f;
    int n = 0;
    const int h0 = loop_start(q.bfw);
    for (int h = h0; (float)h <= q.bfw; ++h) {
        const float ty = y + ((float)h * q.bsw) / q.bfw;
        for (int w = h0; (float)w <= q.bfw; ++w) {
            const float d = D.fetch(x + ((float)w * q.bsw) / q.bfw, ty);
            if (z > d) {
                avg = avg + d;
                ++n;
            }
        }
    }
    return n == 0 ? 1.0f : avg / (float)n;
}

// computePenumbraWidth :365-374
__device__ __forceinline__ float pcss_penumbra(const PcssUniforms& q, float avg, float z) {
    if (avg < 0.99f) return 0.0f;
    const float pw = ((z - avg) / avg) * q.radius;
    return (q.znear * pw) / z;
}

// PCF :376-398
__device__ __forceinline__ float pcss_pcf(const DepthMap& D, const PcssUniforms& q, float si, float pw, float x,
                                          float y, float z) {
    const float stepSize = 2.0f * pw / (float)q.ks;
    if (stepSize <= 0.0f || stepSize >= 1.0f) return 1.0f;
    float illum = 0.0f;
    const int h0 = loop_start(q.kfw);
    for (int h = h0; (float)h <= q.kfw; ++h) {
        const float ty = y + ((float)h * pw) / q.kfw;
        for (int w = h0; (float)w <= q.kfw; ++w) {
            const float d = D.fetch(x + ((float)w * pw) / q.kfw, ty);
            illum = illum + ((z <= d) ? 1.0f : si);
        }
    }
    return illum / (float)(q.ks * q.ks);
}

// main :605-632, percentageCloserSoftShadows :556-563
__global__ __launch_bounds__(kBlockThreads) void k_pcss(GBufferView gb, DepthMap D, HardUniforms u, PcssUniforms q,
                                                         float* __restrict__ out) {
    int x, y;
    pixel_of_block<kSpxFiltered>(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    const float vx = gb.pos[k];
    if (vx == 0.0f) {  // discarded: the cleared target holds 0
        out[k] = 0.0f;
        return;
    }
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    float sc[4];
    mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
    const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
    float shadow = 1.0f;
    if (u.pre_eval)
        shadow = pre_evaluation(u, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
    if (sc[3] > 0.0f && shadow == 1.0f) {
        const float avg = pcss_blocker(D, q, n0, n1, n2);
        const float pw = pcss_penumbra(q, avg, n2);
        shadow = pcss_pcf(D, q, u.si, pw, n0, n1, n2);
    }
    out[k] = shadow;
}

constexpr int kMaxTaps = 16;

struct SoftUniforms {
    float m[12];  // rows 0-3 of columns 0-2 of B*lightMVP (AccurateSoftShadow.frag:70-73)
    float si;
    int nsamples;
    int pcf;
    // si lies in [0, 1] on a 2^-16 grid: every partial sum of up to 16 tap
    // values (1 or si) is then a multiple of 2^-16 below 32, exact in FP32, so
    // a sample's tap sum equals lit + (taps - lit) * si in any order
    int exact_sum;
    int ntaps;             // taps per axis of the PCF loop
    float taps[kMaxTaps];  // loop values w of `for (w = -offset; w </<= offset; w += step)`
    float tapw[kMaxTaps];  // taps[a] * (1/W) and taps[a] * (1/H): the shader's products, in FP32
    float taph[kMaxTaps];
    // A sample whose clip coordinates satisfy cw > 0 and cx < lo_x * cw,
    // cx > hi_x * cw, cy < lo_y * cw or cy > hi_y * cw has every PCF tap
    // outside the map: lo = -(max |tap offset| + 0.02), hi = 1 - lo, so after
    // the division and each tap's addition the coordinate stays below 0 (or
    // at least 1) under any rounding.
    float lo_x, hi_x, lo_y, hi_y;
};

// Rows of NT consecutive texels; selecting element k of a row.
template <int NT>
struct Row { float v[NT]; };

template <int NT>
__device__ __forceinline__ float pick(const Row<NT>& r, int k) {
    float x = r.v[0];
#pragma unroll
    for (int q = 1; q < NT; ++q) x = (k == q) ? r.v[q] : x;
    return x;
}

// One sample's term of AccurateSoftShadow.frag:75-121 for a pixel whose
// columns-0-2 product with B*lightMVP is `cm` (monteCarlo: accFactor = 1).
// NT > 0: the PCF loop has NT taps per axis, and when every tap lies inside
// the map the NT x NT taps are served by NT row loads of NT consecutive
// texels (the taps of one sample span fewer than NT texels per axis, which
// the guard below checks).  NT == 0: generic per-tap loop.
// `quot` (optional, exact-sum uniforms only): quot[lit] = (lit + (NT*NT - lit)
// * si) / (NT*NT) for lit = 0..NT*NT, the same expression evaluated once per
// workgroup, which replaces the per-sample division where the tap sum is
// formed from the lit count.
template <int NT>
__device__ __forceinline__ float soft_term(const DepthMap& D, const SoftUniforms& u, const float cm[4], float4 T,
                                           float incrWidth, float incrHeight,
                                           const float* __restrict__ quot = nullptr) {
    float cx = cm[0] + T.x, cy = cm[1] + T.y, cz = cm[2] + T.z, cw = cm[3] + T.w;
    if (NT > 0 && u.pcf && cw > 0.0f &&
        (cx < u.lo_x * cw || cx > u.hi_x * cw || cy < u.lo_y * cw || cy > u.hi_y * cw)) {
        // Every tap lies outside the map and reads the border depth 0 (a
        // pixel outside this sample's light frustum): the taps' tests are all
        // sz <= 0, added in the loop's order as the per-tap path below adds them.
        const float sz = cz / cw;
        if (quot && u.exact_sum) return quot[(sz <= 0.0f) ? NT * NT : 0];
        const float v = (sz <= 0.0f) ? 1.0f : u.si;
        float illum = 0.0f;
#pragma unroll
        for (int q = 0; q < (NT > 0 ? NT * NT : 1); ++q) illum = illum + v;
        return illum / (float)(NT * NT);
    }
    float sx = cx / cw, sy = cy / cw, sz = cz / cw;
    if (!u.pcf) {
        float d = D.fetch(sx, sy);
        return (sz <= d) ? 1.0f : u.si;
    }
    if constexpr (NT > 0) {
        float tx[NT], ty[NT];
        int iu[NT], iv[NT];
#pragma unroll
        for (int a = 0; a < NT; ++a) {
            tx[a] = sx + u.tapw[a];
            ty[a] = sy + u.taph[a];
            iu[a] = (int)floorf(tx[a] * D.fw);
            iv[a] = (int)floorf(ty[a] * D.fh);
        }
        // every tap inside [0, W) x [0, H): the taps increase with a and rounding
        // is monotone, so the first and last bound the others (NaN fails)
        bool inside = tx[0] * D.fw >= 0.0f && tx[NT - 1] * D.fw < D.fw && ty[0] * D.fh >= 0.0f &&
                      ty[NT - 1] * D.fh < D.fh;
        // one row of NT texels from iu[0] must cover every tap column
        inside = inside && (iu[NT - 1] - iu[0] <= NT - 1) && (iu[0] + NT - 1 < D.w);
        float illum = 0.0f;
        if (inside) {
            // each row is copied into its own small struct and then into
            // scalars, so the rows stay in VGPRs (an array of rows filled by
            // one memcpy each was placed in LDS by the compiler)
            // inside the map the texel index is below W * H <= 2^28: 32-bit
            // unsigned offsets from the map base (no 64-bit address math)
            float rv[NT][NT];
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                Row<NT> r;
                // byte offset in 32 bits (a map holds at most 2^28 texels):
                // the load adds it to the map base as an unsigned offset;
                // row and width are below 2^24, so a 24-bit multiply is exact
                const uint32_t boff = (__umul24((uint32_t)iv[b], (uint32_t)D.w) + (uint32_t)iu[0]) * 4u;
                __builtin_memcpy(&r, reinterpret_cast<const char*>(D.texels) + boff, sizeof(Row<NT>));
#pragma unroll
                for (int q = 0; q < NT; ++q) rv[b][q] = r.v[q];
            }
            if (u.exact_sum) {
                // each texel is tested once; tap column a reads row element
                // iu[a] - iu[0], so each element's count is weighted by the
                // number of tap columns that read it
                int weight[NT];
#pragma unroll
                for (int q = 0; q < NT; ++q) weight[q] = 0;
#pragma unroll
                for (int a = 0; a < NT; ++a) {
#pragma unroll
                    for (int q = 0; q < NT; ++q) weight[q] += (iu[a] - iu[0] == q) ? 1 : 0;
                }
                int lit = 0;
#pragma unroll
                for (int q = 0; q < NT; ++q) {
                    int col = 0;
#pragma unroll
                    for (int b = 0; b < NT; ++b) col += (sz <= rv[b][q]) ? 1 : 0;
                    lit += weight[q] * col;
                }
                if (quot) return quot[lit];
                illum = (float)lit + (float)(NT * NT - lit) * u.si;
            } else {
#pragma unroll
                for (int a = 0; a < NT; ++a) {
                    const int k = iu[a] - iu[0];
#pragma unroll
                    for (int b = 0; b < NT; ++b) {
                        float d = rv[b][0];
#pragma unroll
                        for (int q = 1; q < NT; ++q) d = (k == q) ? rv[b][q] : d;
                        illum = illum + ((sz <= d) ? 1.0f : u.si);
                    }
                }
            }
        } else {
#pragma unroll
            for (int a = 0; a < NT; ++a) {
#pragma unroll
                for (int b = 0; b < NT; ++b) {
                    float d = D.fetch(tx[a], ty[b]);
                    illum = illum + ((sz <= d) ? 1.0f : u.si);
                }
            }
        }
        return illum / (float)(NT * NT);
    } else {
        float illum = 0.0f;
        for (int a = 0; a < u.ntaps; ++a) {
            float txa = sx + u.taps[a] * incrWidth;
            for (int b = 0; b < u.ntaps; ++b) {
                float d = D.fetch(txa, sy + u.taps[b] * incrHeight);
                illum = illum + ((sz <= d) ? 1.0f : u.si);
            }
        }
        return illum / (float)(u.ntaps * u.ntaps);
    }
}

// adaptiveSampling == 1 (AccurateSoftShadow.frag:79-87): trans.w carries the
// sample's quad-tree level as level * 10000; the sample weighs
// 1 / (2^level + 1)^2 (level clamped to 4) and projects with the remainder.
__device__ __forceinline__ float adaptive_weight(float4& T) {
    int level = (int)(T.w / 10000.0f);
    if (level > 4) level = 4;
    T.w = T.w - (float)(level * 10000);
    float p = 1.0f;  // myPow(2.0, level)
    for (int e = 1; e <= level; ++e) p = p * 2.0f;
    const float f = p + 1.0f;
    return 1.0f / (f * f);
}

// A run of consecutive samples of one Monte-Carlo frame (accFactor 1): the
// running sum `acc` of the samples before the run is read (unless first) and
// the sum after it stored, or, on the last run, divided by the frame's sample
// count.  The additions are the single pass's, in the same order.
struct SoftChunk {
    float* acc;         // width*height running sums
    int first;          // no samples before this run
    float count_total;  // > 0 on the last run: the frame's samples (exact in FP32)
};

// AccurateSoftShadow.frag:52-133 for pixel (x, y): samples in order 0..N-1.
template <int NT, bool kAdaptive = false, bool kChunk = false>
__device__ __forceinline__ void soft_pixel(const GBufferView& gb, const float* __restrict__ maps, int smw, int smh,
                                           const float4* __restrict__ trans, const SoftUniforms& u,
                                           float* __restrict__ out, int x, int y, const SoftChunk& ch = {}) {
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    const float vx = gb.pos[k];
    if (vx == 0.0f) {  // Shadow.frag:244 background test; the cleared target holds 0
        if (!kChunk || ch.count_total > 0.0f) out[k] = 0.0f;
        return;
    }
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    float cm[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // :70-73, v.w ignored
        float a = u.m[0 * 4 + i] * vx;
        a = a + u.m[1 * 4 + i] * vy;
        a = a + u.m[2 * 4 + i] * vz;
        cm[i] = a;
    }
    const size_t mapsz = (size_t)smw * smh;
    DepthMap D{maps, smw, smh, (float)smw, (float)smh};
    const float incrWidth = 1.0f / D.fw, incrHeight = 1.0f / D.fh;
    float acc = 0.0f, count = 0.0f;
    if constexpr (kChunk) {
        if (!ch.first) acc = ch.acc[k];
    }
    // unrolled so several samples' gathers are in flight; the sum stays in order
#pragma unroll 4
    for (int i = 0; i < u.nsamples; ++i) {
        D.texels = maps + (size_t)i * mapsz;
        if constexpr (kAdaptive) {
            float4 T = trans[i];
            const float af = adaptive_weight(T);
            acc = acc + soft_term<NT>(D, u, cm, T, incrWidth, incrHeight) * af;
            count = count + af;
        } else {
            acc = acc + soft_term<NT>(D, u, cm, trans[i], incrWidth, incrHeight);
            count = count + 1.0f;
        }
    }
    if constexpr (kChunk) {
        if (ch.count_total > 0.0f)
            out[k] = acc / ch.count_total;
        else
            ch.acc[k] = acc;
    } else {
        out[k] = acc / count;
    }
}

// Block shape of the soft pass.  Each (block, sample) pair fetches every
// cache line its taps touch, and a line that straddles two blocks is fetched
// by both, since neighbouring blocks reach a given sample at different times
// (a model of this over the 64-sample frame reproduces the measured HBM
// traffic of 16x16 blocks to 3%).  A 128-byte line holds 32 texels of a map
// row, so blocks wider than tall straddle fewer lines per pixel: the model
// gives 1.83 GB for 16x16 blocks, 1.52 GB for 32x8 and 1.45 GB for 64x4.
// Each wave covers a 16 x 4 patch.  On the 64-sample 1080p frame (3x3 taps)
// the pass takes 0.355 ms with 64x4 blocks of 16x4 waves, against 0.395 ms
// with 16x16 blocks of 8x8 waves, 0.358 ms with 32x8 blocks of 16x4 waves,
// 0.362 ms with 64x4 blocks of 32x2 waves, 0.402 ms with 64x1 waves and
// 0.417 ms with 128x2 blocks; 2 or 8 strips per XCD instead of 4 took
// 0.394 and 0.360 ms.  Larger workgroups straddle fewer lines still (the
// model: 1.33 GB for 64x8, 1.23 GB for 128x8) but ran slower: 0.386 ms for
// 64x8, 0.413 for 128x4, 0.483 for 128x8, 0.492 for 256x4.
constexpr int kSoftBW = 64, kSoftBH = 4;
constexpr int kSoftThreads = kSoftBW * kSoftBH;
constexpr int kSoftWX = 16, kSoftWY = 64 / kSoftWX;
constexpr int kSpxSoftWide = 4;
static_assert(kSoftBW % kSoftWX == 0 && kSoftBH % kSoftWY == 0, "wave tiles must tile the block");
inline int soft_blocks(int w, int h) { return ((w + kSoftBW - 1) / kSoftBW) * ((h + kSoftBH - 1) / kSoftBH); }
__device__ __forceinline__ void soft_block_pixel(int w, int h, int& x, int& y) {
    const int bx_n = (w + kSoftBW - 1) / kSoftBW;
    const int by_n = (h + kSoftBH - 1) / kSoftBH;
    int bx, by;
    strip_block<kSpxSoftWide>(xcd_order(bx_n * by_n), bx_n, by_n, bx, by);
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    constexpr int kWavesX = kSoftBW / kSoftWX;
    x = bx * kSoftBW + (wave % kWavesX) * kSoftWX + lane % kSoftWX;
    y = by * kSoftBH + (wave / kWavesX) * kSoftWY + lane / kSoftWX;
}

// One thread per pixel, any tap count.  Blocks are ordered XCD-aware
// (soft_block_pixel) so each XCD's L2 serves one contiguous band of the view.
template <int NT, bool kAdaptive = false>
__global__ __launch_bounds__(kSoftThreads) void k_soft(GBufferView gb, const float* __restrict__ maps, int smw,
                                                            int smh, const float4* __restrict__ trans,
                                                            SoftUniforms u, float* __restrict__ out) {
    int x, y;
    soft_block_pixel(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    soft_pixel<NT, kAdaptive>(gb, maps, smw, smh, trans, u, out, x, y);
}

// Every pixel adds its samples in order, as soft_pixel does (same bits).
template <int NT, bool kAdaptive = false>
__global__ __launch_bounds__(kSoftThreads) void k_soft_wide(GBufferView gb, const float* __restrict__ maps, int smw,
                                                                 int smh, const float4* __restrict__ trans,
                                                                 SoftUniforms u, float* __restrict__ out) {
    int x, y;
    soft_block_pixel(gb.w, gb.h, x, y);
    const size_t plane = (size_t)gb.w * gb.h;
    float cm[4];
    float acc = 0.0f;
    bool fg = false;
    if (x < gb.w && y < gb.h) {
        const size_t k = (size_t)y * gb.w + x;
        const float vx = gb.pos[k];
        if (vx == 0.0f) {  // Shadow.frag:244 background test; the cleared target holds 0
            out[k] = 0.0f;
        } else {
            fg = true;
            const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
#pragma unroll
            for (int c = 0; c < 4; ++c) {  // AccurateSoftShadow.frag:70-73, v.w ignored
                float a = u.m[0 * 4 + c] * vx;
                a = a + u.m[1 * 4 + c] * vy;
                a = a + u.m[2 * 4 + c] * vz;
                cm[c] = a;
            }
        }
    }
    // the tap sum's quotient for each lit count (exact-sum uniforms)
    __shared__ float quot[NT * NT + 1];
    if ((int)threadIdx.x <= NT * NT) {
        const int lit = (int)threadIdx.x;
        quot[lit] = ((float)lit + (float)(NT * NT - lit) * u.si) / (float)(NT * NT);
    }
    __syncthreads();
    const float* qt = u.exact_sum ? quot : nullptr;
    const size_t mapsz = (size_t)smw * smh;
    DepthMap D{maps, smw, smh, (float)smw, (float)smh};
    const float incrWidth = 1.0f / D.fw, incrHeight = 1.0f / D.fh;
    float count = 0.0f;
    for (int s = 0; s < u.nsamples; ++s) {
        D.texels = maps + (size_t)s * mapsz;
        float4 T = trans[s];
        float af = 1.0f;
        if constexpr (kAdaptive) af = adaptive_weight(T);
        if (fg) {
            const float term = soft_term<NT>(D, u, cm, T, incrWidth, incrHeight, qt);
            if constexpr (kAdaptive)
                acc = acc + term * af;
            else
                acc = acc + term;
        }
        count = count + af;
    }
    if (fg) out[(size_t)y * gb.w + x] = acc / count;
}

// One run of samples of a Monte-Carlo frame (SoftChunk).
template <int NT>
__global__ __launch_bounds__(kSoftThreads) void k_soft_chunk(GBufferView gb, const float* __restrict__ maps, int smw,
                                                                  int smh, const float4* __restrict__ trans,
                                                                  SoftUniforms u, SoftChunk ch,
                                                                  float* __restrict__ out) {
    int x, y;
    soft_block_pixel(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    soft_pixel<NT, false, true>(gb, maps, smw, smh, trans, u, out, x, y, ch);
}

// ---------------------------------------------------------------------------
// Adaptive light sampling, one quad-tree node's test (evaluateSubAreaLightSource,
// SoftShadowMapping/src/main.cpp:440-571).
// ---------------------------------------------------------------------------
struct QtUniforms {
    Mat4 m[4];   // B * lightMVP of the node's four corner samples
    int map[4];  // their slots in the map array (shadowMapIndices)
};

constexpr float kQtBackground = 0.63f;  // the reprojection target's clear colour R (main.cpp:511)

// QuadTree/Reprojection.frag:8-24: how many of the four corner maps light
// the pixel (0..4); background pixels keep the clear colour.  Block (0, 0)
// also clears the node's pass counter for k_qt_evaluate.
__global__ __launch_bounds__(kBlockThreads) void k_qt_reproject(GBufferView gb, const float* __restrict__ maps,
                                                                 int smw, int smh, QtUniforms u,
                                                                 float* __restrict__ img,
                                                                 unsigned* __restrict__ count) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *count = 0u;
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    const float vx = gb.pos[k];
    if (vx == 0.0f) {
        img[k] = kQtBackground;
        return;
    }
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    const size_t mapsz = (size_t)smw * smh;
    float sh[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float sc[4];
        mat4_vec4(u.m[i].m, vx, vy, vz, 1.0f, sc);
        const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
        DepthMap D{maps + (size_t)u.map[i] * mapsz, smw, smh, (float)smw, (float)smh};
        sh[i] = (n2 <= D.fetch(n0, n1)) ? 1.0f : 0.0f;
    }
    img[k] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Texel of the reprojection image that GL_NEAREST samples for reduced-grid
// coordinate c (texture coordinate (c + 0.5) / nd over an n-texel axis),
// or -1 outside [0, n) (CLAMP_TO_BORDER reads 0 there).  Exact arithmetic:
// floor((c + 0.5) * n / nd).
__device__ __forceinline__ int qt_texel(int c, int n, int nd) {
    if (c < 0) return -1;
    const long long a = ((2ll * c + 1) * n) / (2ll * nd);
    return a < n ? (int)a : -1;
}

// QuadTree/Evaluation.frag:6-22 over the (W/div) x (H/div) viewport: a
// penumbra pixel (1..3 lit corners) whose 3x3 neighbourhood all agree passes;
// passes are counted on the device (the reference's GL_SAMPLES_PASSED query).
__global__ __launch_bounds__(256) void k_qt_evaluate(const float* __restrict__ img, int w, int h, int wd, int hd,
                                                     unsigned* __restrict__ count) {
    __shared__ unsigned block_count;
    if (threadIdx.x == 0) block_count = 0u;
    __syncthreads();
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    bool pass = false;
    if (id < wd * hd) {
        const int i = id % wd, j = id / wd;
        auto tex = [&](int ci, int cj) -> float {
            const int a = qt_texel(ci, w, wd), b = qt_texel(cj, h, hd);
            return (a < 0 || b < 0) ? 0.0f : img[(size_t)b * w + a];
        };
        const float sum = tex(i, j);
        pass = !(sum == kQtBackground || sum == 0.0f || sum == 4.0f || sum == 5.0f);
        for (int dy = -1; dy <= 1 && pass; ++dy)
            for (int dx = -1; dx <= 1 && pass; ++dx)
                if (sum != tex(i + dx, j + dy)) pass = false;
    }
    const unsigned long long b = __ballot(pass);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&block_count, (unsigned)__popcll(b));
    __syncthreads();
    if (threadIdx.x == 0 && block_count) atomicAdd(count, block_count);
}

// ---------------------------------------------------------------------------
// Deferred Phong shading: PhongShading.frag:11-46 (ShadowMapping and
// SoftShadowMapping share it; the shadow R channel is the hard or soft pass's
// output).
// ---------------------------------------------------------------------------
struct PhongUniforms {
    Mat4 mv;
    Mat3 nm;
    float lpos[3];
    float si;
    float color[4];  // G-buffer colour of every pixel when no colour plane is given
    float clear[4];  // background: shadeScene's glClearColor (main.cpp:452)
};

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// `pow(x, 0.3 * shininess)` with shininess 10: the FP32 exponent is exactly 3.
// GLSL leaves pow's precision to the implementation; this build evaluates it
// as (x * x) * x.
__device__ __forceinline__ float pow3(float x) { return (x * x) * x; }

__device__ __forceinline__ float4 phong_pixel(const GBufferView& gb, const float* __restrict__ facing,
                                              const float* __restrict__ shadow, const float4* __restrict__ color,
                                              const PhongUniforms& u, size_t k) {
    const size_t plane = (size_t)gb.w * gb.h;
    const float vx = gb.pos[k];
    if (vx == 0.0f) return make_float4(u.clear[0], u.clear[1], u.clear[2], u.clear[3]);  // :40 discard
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    // the stored normal is negated on back faces; PhongShading reads it raw (:42)
    const bool front = facing[k] != 0.0f;
    float nr[3] = {gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]};
    if (!front) {
        nr[0] = -nr[0];
        nr[1] = -nr[1];
        nr[2] = -nr[2];
    }
    const float sh = shadow[k];
    const float specShadow = sh - u.si;  // :18
    float v[4];
    mat4_vec4(u.mv.m, vx, vy, vz, 1.0f, v);  // :20
    float n3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {  // :21 normalMatrix * normal.xyz
        float acc = u.nm.m[0 * 3 + i] * nr[0];
        acc = acc + u.nm.m[1 * 3 + i] * nr[1];
        acc = acc + u.nm.m[2 * 3 + i] * nr[2];
        n3[i] = acc;
    }
    float N[3], L[3];
    normalize3(n3, N);
    const float ld[3] = {u.lpos[0] - v[0], u.lpos[1] - v[1], u.lpos[2] - v[2]};
    normalize3(ld, L);  // :23
    // :24 E = normalize(-vertex) normalizes all four components, then keeps xyz
    const float e4[4] = {-v[0], -v[1], -v[2], -v[3]};
    const float e_sqr = ((e4[0] * e4[0] + e4[1] * e4[1]) + e4[2] * e4[2]) + e4[3] * e4[3];
    const float e_inv = 1.0f / sqrtf(e_sqr);
    const float E[3] = {e4[0] * e_inv, e4[1] * e_inv, e4[2] * e_inv};
    // :25 R = normalize(-reflect(L, N)), reflect(I, N) = I - 2.0 * dot(N, I) * N
    const float dNL = dot3(N, L);
    const float two_d = 2.0f * dNL;
    const float mr[3] = {-(L[0] - two_d * N[0]), -(L[1] - two_d * N[1]), -(L[2] - two_d * N[2])};
    float R[3];
    normalize3(mr, R);
    const float diff = gmax(dNL, 0.0f);                 // :28 max(dot(normal.xyz, L), 0.0)
    const float spec = pow3(gmax(dot3(R, E), 0.0f));     // :29
    const float amb[4] = {0.4f, 0.4f, 0.4f, 1.0f}, ldif[4] = {0.5f, 0.5f, 0.5f, 1.0f},
                lspe[4] = {0.25f, 0.25f, 0.25f, 1.0f};
    float col[4];
    if (color) {
        const float4 c = color[k];
        col[0] = c.x; col[1] = c.y; col[2] = c.z; col[3] = c.w;
    } else {
        col[0] = u.color[0]; col[1] = u.color[1]; col[2] = u.color[2]; col[3] = u.color[3];
    }
    float o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float Idiff = ldif[c] * diff;
        const float Ispec = (specShadow * lspe[c]) * spec;
        o[c] = (sh * col[c]) * ((Idiff + Ispec) + amb[c]);  // :32
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// Rows of 64 pixels per wave (a workgroup covers 64 x 4): each load of a
// G-buffer plane reads two whole 128-byte lines.
inline int row_blocks(int w, int h) { return ((w + 63) / 64) * ((h + 3) / 4); }
__device__ __forceinline__ void row_block_pixel(int w, int h, int& x, int& y) {
    const int bx_n = (w + 63) / 64, by_n = (h + 3) / 4;
    int bx, by;
    strip_block<1>(xcd_order(bx_n * by_n), bx_n, by_n, bx, by);
    x = bx * 64 + (int)(threadIdx.x & 63);
    y = by * 4 + (int)(threadIdx.x >> 6);
}

__global__ __launch_bounds__(kBlockThreads) void k_phong(GBufferView gb, const float* __restrict__ facing,
                                                          const float* __restrict__ shadow,
                                                          const float4* __restrict__ color, PhongUniforms u,
                                                          float4* __restrict__ out) {
    int x, y;
    row_block_pixel(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t k = (size_t)y * gb.w + x;
    out[k] = phong_pixel(gb, facing, shadow, color, u, k);
}

// Tap values of the PCF loop `for (w = -offset; w < offset (or <=); w += step)`
// with step = 2*offset/kernelOrder, accumulated in FP32 exactly as the shader
// loop does (AccurateSoftShadow.frag:101-109, Shadow.frag:93-99).
int pcf_taps(float offset, int kernel_order, bool inclusive, float* taps, int max_taps) {
    float stepSize = 2.0f * offset / (float)kernel_order;
    int n = 0;
    for (float w = -offset; inclusive ? (w <= offset) : (w < offset); w += stepSize) {
        if (n == max_taps) return -1;
        taps[n++] = w;
    }
    return n;
}

// Host side: argument checks of the soft-shadow pass and its uniforms
// (shared by the single pass and the batched frames).  Returns a GI_* status.
inline bool soft_finite(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

inline int soft_setup(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                      const float light_mvp[16], const float* trans_dev, const gi_soft_params* p,
                      const float* out_dev, SoftUniforms& u) {
    if (!gbuffer || !gbuffer->position_dev || gbuffer->width <= 0 || gbuffer->height <= 0 ||
        gbuffer->width > GI_MAX_EDGE || gbuffer->height > GI_MAX_EDGE)
        return GI_EINVAL;
    if (!maps_dev || !light_mvp || !trans_dev || !p || !out_dev) return GI_EINVAL;
    if (nsamples <= 0 || nsamples > GI_MAX_SAMPLES) return GI_EINVAL;
    if (smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE) return GI_EINVAL;
    if (!soft_finite(light_mvp, 16) || !std::isfinite(p->shadow_intensity) || !std::isfinite(p->penumbra))
        return GI_ENONFINITE;
    if (p->pcf && (p->kernel_order <= 0 || p->kernel_order > 14 || !(p->penumbra > 0.0f) || p->penumbra > 64.0f))
        return GI_EINVAL;
    if ((reinterpret_cast<uintptr_t>(trans_dev) & 15u) != 0) return GI_EINVAL;
    for (int i = 0; i < 12; ++i) u.m[i] = light_mvp[i];
    u.si = p->shadow_intensity;
    u.nsamples = nsamples;
    u.pcf = p->pcf ? 1 : 0;
    {
        const float g = u.si * 65536.0f;  // exact: a power-of-two scale of a value in [0, 1]
        u.exact_sum = (u.si >= 0.0f && u.si <= 1.0f && g == std::floor(g)) ? 1 : 0;
    }
    u.ntaps = 1;
    for (int i = 0; i < kMaxTaps; ++i) u.taps[i] = u.tapw[i] = u.taph[i] = 0.0f;
    u.lo_x = u.lo_y = -INFINITY;  // no shortcut unless the PCF taps are set up below
    u.hi_x = u.hi_y = INFINITY;
    if (p->pcf) {
        u.ntaps = pcf_taps(p->penumbra, p->kernel_order, p->inclusive != 0, u.taps, kMaxTaps);
        if (u.ntaps <= 0) return GI_ENOTSUP;
        // the kernel's incrWidth = 1.0f / (float)W; the taps must increase for
        // the extremes-only bounds test
        const float incr_w = 1.0f / (float)smw, incr_h = 1.0f / (float)smh;
        float mx = 0.0f, my = 0.0f;
        for (int i = 0; i < u.ntaps; ++i) {
            u.tapw[i] = u.taps[i] * incr_w;
            u.taph[i] = u.taps[i] * incr_h;
            if (i > 0 && !(u.taps[i] > u.taps[i - 1])) return GI_ENOTSUP;
            mx = std::max(mx, std::fabs(u.tapw[i]));
            my = std::max(my, std::fabs(u.taph[i]));
        }
        u.lo_x = -(mx + 0.02f);
        u.hi_x = 1.0f + mx + 0.02f;
        u.lo_y = -(my + 0.02f);
        u.hi_y = 1.0f + my + 0.02f;
    }
    return GI_OK;
}

// Template variant of k_soft / soft_pixel for uniforms u: 3 or 4 taps per
// axis take the row-load path, everything else the generic loop.
inline int soft_variant(const SoftUniforms& u) { return (u.pcf && (u.ntaps == 3 || u.ntaps == 4)) ? u.ntaps : 0; }

}  // namespace
}  // namespace gi
