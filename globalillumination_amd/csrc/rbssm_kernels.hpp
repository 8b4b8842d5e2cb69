// rbssm_kernels.hpp — revectorization-based soft shadows (RBSSM) for gfx950.
//
// SoftShadowMapping/Shaders/SoftShadow/RBSSM.frag: a PCSS-sized kernel
// (blocker search and penumbra width as PlausibleSoftShadow.frag) whose taps
// that sit on a shadow-map discontinuity are revectorized: the discontinuity
// is oriented by four edge walks (computeDiscontinuityLength :193-253),
// normalized (:255-309) and filled by smoothONDS (:311-1117).  Every branch,
// coordinate step and `mix`/`step` is evaluated as the shader writes it, in
// FP32; the neighbour tests are the RBSM ones (Rbsm<false>::disc_f / disc_v /
// discontinuity, identical in RBSSM.frag :37-186, :1137-1147).
#pragma once

namespace gi {
namespace {

__device__ __forceinline__ float gstep(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
__device__ __forceinline__ float fbool(bool b) { return b ? 1.0f : 0.0f; }

struct Rbssm {
    Rbsm<false> rb;  // shadow map, shadowMapStep, depthThreshold, maxSearch
    mutable float newDepth;
    int cap;         // scan bound of the accurate smoothONDS: 4 (W + H) + 16

    __device__ __forceinline__ float sx() const { return rb.sx; }
    __device__ __forceinline__ float sy() const { return rb.sy; }
    __device__ __forceinline__ float tex(float s, float t) const { return rb.tex(s, t); }
    // getDisc(coord, vec2 dir, float discType) :63-101
    __device__ __forceinline__ bool discf(float x, float y, float z, float dx, float dy, float type) const {
        return rb.disc_f(x, y, z, dx, dy, type);
    }
    // getDisc(coord, vec2 dir, vec4 discType) :103-186 (writes newDepth)
    __device__ __forceinline__ bool discv(float x, float y, float z, float dx, float dy, float r, float g,
                                          float b) const {
        float nd;
        const bool h = rb.disc_v(x, y, z, dx, dy, r, g, b, nd);
        newDepth = nd;
        return h;
    }

    static __device__ __forceinline__ float compress(float nd) { return -2.0f - ((0.5f - nd) * 2.0f); }
    static __device__ __forceinline__ float decompress(float nd) { return 0.5f - ((nd + 2.0f) * -1.0f) / 2.0f; }

    // computeDiscontinuityLength :193-253
    __device__ float length(float ir, float ig, float ib, float x, float y, float z, float dx, float dy) const {
        float foundEdgeEnd = 0.0f;
        if (dx == 0.0f && ir == 0.0f && ig != 0.0f) return -1.0f;
        if (dy == 0.0f && ir != 0.0f && ig == 0.0f) return -1.0f;
        if (((0.5f - ir) * 8.0f - 1.0f) == dx) return 1.0f;
        if (((ig - 0.5f) * 8.0f + 1.0f) == dy) return 1.0f;
        float dist = 1.0f;
        const float stx = dx * sx(), sty = dy * sy();
        float cx = x + stx, cy = y + sty, cz = z;
        for (int it = 0; it < rb.max_search; ++it) {
            const float d = tex(cx, cy);
            if (ib == 0.0f)
                if (fabsf(cz - d) < rb.thr) cz = cz - rb.thr;
            const float center = le01(cz, d);
            if (fabsf(center - ib) == 0.0f) {
                foundEdgeEnd = discf(cx, cy, cz, 0.0f, 0.0f, ib) ? 1.0f : 0.0f;
                break;
            }
            if (!discv(cx, cy, cz, dx, dy, ir, ig, ib)) break;
            dist = dist + 1.0f;
            cx = cx + stx;
            cy = cy + sty;
            if (ib == 1.0f) cz = newDepth;
        }
        return gmix(-dist, dist, foundEdgeEnd);
    }

    // normalizeDiscontinuitySpace :255-280
    __device__ float normalize(float a, float b, float sub) const {
        if (a < 0.0f && b < 0.0f) return -1.0f;
        const float edgeLength = gmin(fabsf(a) + fabsf(b) - 1.0f, (float)rb.max_search);
        float nd = 1.0f - gmax(a, b) / edgeLength;
        if (a == b)
            nd = nd + gmix(sub / edgeLength, (1.0f - sub) / edgeLength, gstep(0.5f, sub));
        else if (a == gmax(a, b))
            nd = nd + (1.0f - sub) / edgeLength;
        else
            nd = nd + sub / edgeLength;
        if (a > 0.0f && b > 0.0f) return compress(nd);
        return nd;
    }

    // smoothONDS, RBSSM.frag:311-1117 (kAcc = false) and
    // RevectorizationBasedAccurateSoftShadow.frag:320-1138 (kAcc = true).  The
    // two differ in three early exits and in eight scans that RBSSM.frag runs
    // once (their `while` lines are commented out there) and the accurate
    // shader repeats until the scan condition fails.  Each scan moves at least
    // one texel per step and ends at the map border (depth 0) for any finite
    // depth; `cap` steps (4 (W + H) + 16) bound it for non-finite depths.
    // lightCoord, normalizedDiscontinuity and discontinuity are by-value
    // copies, as in the shader.
    template <bool kAcc>
    __device__ float smooth(float lx, float ly, float lz, float nx, float ny, float dr, float dg, float db, float ux,
                            float uy) const {
        const float sxx = sx(), syy = sy();
        if (db == 0.0f) {
            if (kAcc) {
                if (nx == -1.0f && ny == -1.0f) return 1.0f;
                if (dr == 0.75f && dg == 0.75f) return 0.0f;
            }
            // positive entering discontinuity on both axes, on both sides of one axis
            if (nx <= -2.0f && ny <= -2.0f && (dr == 0.75f || dg == 0.75f)) {
                bool left = true, right = true, bottom = true, top = true;
                if (dr == 0.75f) {
                    if (dg == 0.0f) {
                        ly = ly + syy;
                        top = discf(lx, ly, lz, 0.0f, 1.0f, db);
                        ly = ly - 2.0f * syy;
                        bottom = discf(lx, ly, lz, 0.0f, 1.0f, db);
                        ly = ly + syy;
                        if (top && bottom) return 0.0f;
                        ly = ly + gmix(-syy, syy, gstep(1.0f, fbool(top)));
                        left = discv(lx, ly, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                        right = !left;
                    } else {
                        if (dg == 0.5f)
                            top = false;
                        else if (dg == 0.25f)
                            bottom = false;
                        ly = ly + ((0.5f - dg) * 8.0f - 1.0f) * syy;
                        left = discv(lx, ly, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                        right = discv(lx, ly, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                        if (left && right) {
                            if (ny <= -2.0f) ny = decompress(ny) + 0.5f;
                            const float a = gclamp(ux - (ny - 1.0f), 0.0f, 1.0f);
                            const float b = gclamp((1.0f - ux) - (ny - 1.0f), 0.0f, 1.0f);
                            const float c = gmix(1.0f - uy, uy, gstep(dg, 1.0f));
                            return gmin(gmin(a, b), c);
                        }
                    }
                }
                if (dg == 0.75f) {
                    if (dr == 0.0f) {
                        lx = lx - sxx;
                        left = discf(lx, ly, lz, 1.0f, 0.0f, db);
                        lx = lx + 2.0f * sxx;
                        right = discf(lx, ly, lz, 1.0f, 0.0f, db);
                        lx = lx - sxx;
                        if (left && right) return 0.0f;
                        lx = lx + gmix(sxx, -sxx, gstep(1.0f, fbool(left)));
                        bottom = discv(lx, ly, lz, 0.0f, 0.25f, 0.5f, 0.0f, 0.0f);
                        top = !bottom;
                    } else {
                        if (dr == 0.5f)
                            right = false;
                        else if (dr == 0.25f)
                            left = false;
                        lx = lx - ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                        bottom = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.5f, 0.0f);
                        top = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.25f, 0.0f);
                        if (bottom && top) {
                            if (nx <= -2.0f) nx = decompress(nx) + 0.5f;
                            const float a = gclamp(uy - (nx - 1.0f), 0.0f, 1.0f);
                            const float b = gclamp((1.0f - uy) - (nx - 1.0f), 0.0f, 1.0f);
                            const float c = gmix(1.0f - ux, ux, gstep(dr, 1.0f));
                            return gmin(gmin(a, b), c);
                        }
                    }
                }
                if (!left && !bottom)
                    return gclamp((1.0f - ux) - (1.0f - uy), 0.0f, 1.0f);
                else if (!right && !bottom)
                    return gclamp(ux - (1.0f - uy), 0.0f, 1.0f);
                else if (!left && !top)
                    return gclamp((1.0f - uy) - ux, 0.0f, 1.0f);
                else if (!right && !top)
                    return gclamp((1.0f - uy) - (1.0f - ux), 0.0f, 1.0f);
            }

            if (dr == 0.75f || dg == 0.75f) {
                // entering left and right discontinuity: scan while umbra on both sides
                if (dr == 0.75f && dg != 0.0f) {
                    bool left, right;
                    int guard = 0;
                    do {
                        ly = ly + ((dg - 0.75f) * 4.0f) * syy;
                        left = discv(lx, ly, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                        right = discv(lx, ly, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                        if (!kAcc || lz > tex(lx, ly)) break;
                    } while (left && right && ++guard < cap);
                    if (!left && !right) return gclamp(1.0f - ny, 0.0f, 1.0f);
                    const float sub = gmix(1.0f - uy, uy, gstep(1.0f, dg));
                    const float a =
                        gmix(sub, gclamp((1.0f - ux) - (ny - 1.0f), 0.0f, 1.0f), gstep(1.0f, fbool(right)));
                    const float b = gmix(sub, gclamp(ux - (ny - 1.0f), 0.0f, 1.0f), gstep(1.0f, fbool(left)));
                    return gmin(a, b);
                }
                // entering left and right discontinuity only: scan up and down for the ends
                if (dr == 0.75f && dg == 0.0f) {
                    bool topLeft = true, topRight = true, bottomLeft = true, bottomRight = true;
                    bool topCenter = false, bottomCenter = false;
                    float ty = ly, by = ly;
                    int guard = 0;
                    do {
                        if (!topCenter) {
                            ty = ty + syy;
                            topCenter = !(le01(lz, tex(lx, ty)) != 0.0f);
                            topLeft = discv(lx, ty, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                            topRight = discv(lx, ty, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                        }
                        if (!bottomCenter) {
                            by = by - syy;
                            bottomCenter = !(le01(lz, tex(lx, by)) != 0.0f);
                            bottomLeft = discv(lx, by, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                            bottomRight = discv(lx, by, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                        }
                        if (!kAcc) break;
                    } while (topLeft && topRight && bottomLeft && bottomRight && (!topCenter || !bottomCenter) &&
                             ++guard < cap);
                    if (topCenter) topLeft = topRight = true;
                    if (bottomCenter) bottomLeft = bottomRight = true;
                    if (ny <= -2.0f) {
                        ny = decompress(ny) + 0.5f;
                        const float a = gclamp(ux - (ny - 1.0f), 0.0f, 1.0f);
                        const float b = gclamp((1.0f - ux) - (ny - 1.0f), 0.0f, 1.0f);
                        if (!bottomRight || !topRight) return a;
                        else if (!bottomLeft || !topLeft) return b;
                        else return gmin(a, b);
                    }
                    if ((!bottomRight && !bottomLeft) || (!topRight && !topLeft))
                        return gclamp(1.0f - ny, 0.0f, 1.0f);
                    else if (!bottomRight || !topRight)
                        return gclamp(ux - (ny - 1.0f), 0.0f, 1.0f);
                    else
                        return gclamp((1.0f - ux) - (ny - 1.0f), 0.0f, 1.0f);
                }
                // entering top and bottom discontinuity: scan while umbra above and below
                if (dr != 0.0f && dg == 0.75f) {
                    bool bottom, top;
                    int guard = 0;
                    do {
                        lx = lx - ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                        bottom = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.5f, 0.0f);
                        top = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.25f, 0.0f);
                        if (!kAcc || lz > tex(lx, ly)) break;
                    } while (bottom && top && ++guard < cap);
                    if (!bottom && !top) return gclamp(1.0f - nx, 0.0f, 1.0f);
                    const float sub = gmix(ux, 1.0f - ux, gstep(dr, 0.25f));
                    const float a = gmix(sub, gclamp(uy - (nx - 1.0f), 0.0f, 1.0f), gstep(1.0f, fbool(top)));
                    const float b =
                        gmix(sub, gclamp((1.0f - uy) - (nx - 1.0f), 0.0f, 1.0f), gstep(1.0f, fbool(bottom)));
                    return gmin(a, b);
                }
                // entering top and bottom discontinuity only: scan right and left for the ends
                if (dr == 0.0f && dg == 0.75f) {
                    bool topLeft = true, topRight = true, bottomLeft = true, bottomRight = true;
                    bool leftCenter = false, rightCenter = false;
                    float rx = lx, lxx = lx;
                    int guard = 0;
                    do {
                        if (!rightCenter) {
                            rx = rx + sxx;
                            rightCenter = !(le01(lz, tex(rx, ly)) != 0.0f);
                            bottomRight = discv(rx, ly, lz, 1.0f, 0.0f, 0.0f, 0.5f, 0.0f);
                            topRight = discv(rx, ly, lz, 1.0f, 0.0f, 0.0f, 0.25f, 0.0f);
                        }
                        if (!leftCenter) {
                            lxx = lxx - sxx;
                            leftCenter = !(le01(lz, tex(lxx, ly)) != 0.0f);
                            bottomLeft = discv(lxx, ly, lz, 1.0f, 0.0f, 0.0f, 0.5f, 0.0f);
                            topLeft = discv(lxx, ly, lz, 1.0f, 0.0f, 0.0f, 0.25f, 0.0f);
                        }
                        if (!kAcc) break;
                    } while (topLeft && topRight && bottomLeft && bottomRight && (!rightCenter || !leftCenter) &&
                             ++guard < cap);
                    if (rightCenter) bottomRight = topRight = true;
                    if (leftCenter) topLeft = bottomLeft = true;
                    if (nx <= -2.0f) {
                        nx = decompress(nx) + 0.5f;
                        const float a = gclamp(uy - (nx - 1.0f), 0.0f, 1.0f);
                        const float b = gclamp((1.0f - uy) - (nx - 1.0f), 0.0f, 1.0f);
                        if (!bottomRight || !bottomLeft) return a;
                        else if (!topRight || !topLeft) return b;
                        else return gmin(a, b);
                    }
                    if ((!bottomRight && !topRight) || (!bottomLeft && !topLeft))
                        return gclamp(1.0f - nx, 0.0f, 1.0f);
                    else if (!bottomRight || !bottomLeft)
                        return gclamp(uy - (nx - 1.0f), 0.0f, 1.0f);
                    else
                        return gclamp((1.0f - uy) - (nx - 1.0f), 0.0f, 1.0f);
                }
            }

            // discontinuity in both axes (corner)
            if (dr > 0.0f && dg > 0.0f) {
                lx = lx - ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                const bool horizontal = discf(lx, ly, lz, 1.0f, 0.0f, 0.0f);
                lx = lx + ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                ly = ly + ((0.5f - dg) * 8.0f - 1.0f) * syy;
                const bool vertical = discf(lx, ly, lz, 0.0f, 1.0f, 0.0f);
                if (horizontal && !vertical) {
                    dr = 0.0f;
                } else if (!horizontal && vertical) {
                    dg = 0.0f;
                } else if (!horizontal && !vertical) {
                    if (dg == 0.5f) return gclamp((1.0f - nx) - (uy - 1.0f), 0.0f, 1.0f);
                    else if (dg == 0.25f) return gclamp((1.0f - nx) + uy, 0.0f, 1.0f);
                } else {
                    float a, b;
                    if (dr == 0.5f && dg == 0.5f) {
                        a = gmix(1.0f - uy, gclamp(1.0f - (nx - (1.0f - uy)), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(ux, gclamp(1.0f - (ny - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmin(a, b);
                    } else if (dr == 0.5f && dg == 0.25f) {
                        a = gmix(uy, gclamp(1.0f - (nx - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(ux, gclamp(1.0f - (ny - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmin(a, b);
                    } else if (dr == 0.25f && dg == 0.5f) {
                        a = gmix(1.0f - uy, gclamp(1.0f - (nx - (1.0f - uy)), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(1.0f - ux, gclamp(1.0f - (ny - (1.0f - ux)), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmin(a, b);
                    } else if (dr == 0.25f && dg == 0.25f) {
                        a = gmix(uy, gclamp(1.0f - (nx - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(1.0f - ux, gclamp(1.0f - (ny - (1.0f - ux)), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmin(a, b);
                    }
                }
            }

            if (nx <= -2.0f) return gmix(1.0f - uy, uy, gstep(dg, 0.25f));
            if (ny <= -2.0f) return gmix(ux, 1.0f - ux, gstep(dr, 0.25f));
            if (dg > 0.0f) {
                if (dg == 0.5f) return gclamp((1.0f - uy) - (nx - 1.0f), 0.0f, 1.0f);
                else return gclamp(uy - (nx - 1.0f), 0.0f, 1.0f);
            }
            if (dr > 0.0f) {
                if (dr == 0.5f) return gclamp(ux - (ny - 1.0f), 0.0f, 1.0f);
                else return gclamp((1.0f - ux) - (ny - 1.0f), 0.0f, 1.0f);
            }
        } else {  // exiting discontinuity
            if (kAcc && dr == 0.75f && dg == 0.75f) return 1.0f;
            if (dr == 0.75f || dg == 0.75f) {
                // exiting left and right discontinuity only
                if (dr == 0.75f && dg == 0.0f) {
                    bool left = discv(lx - sxx, ly, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                    bool right = discv(lx + sxx, ly, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                    if (left && right) {
                        return gclamp(ny, 0.0f, 1.0f);
                    } else if (left || right) {
                        if (!left) return gmix(1.0f - ux, gclamp(ny - ux, 0.0f, 1.0f), gstep(-2.0f, ny));
                        else return gmix(ux, gclamp(ny - (1.0f - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                    } else {
                        // rows +1, -1, +2, -2, ... until umbra beside the column
                        int count = 0;
                        do {
                            const int mult = count / 2 + 1;
                            const float ry = (count % 2 == 0) ? ly + (float)mult * syy : ly - (float)mult * syy;
                            left = discf(lx - sxx, ry, lz, 1.0f, 0.0f, 0.0f);
                            right = discf(lx + sxx, ry, lz, 1.0f, 0.0f, 0.0f);
                            if (!kAcc || lz <= tex(lx, ly)) break;
                            ++count;
                        } while (!left && !right && count < cap);
                        if (left && right) {
                            return gclamp(ny, 0.0f, 1.0f);
                        } else {
                            if (!left) return gmix(1.0f - ux, gclamp(ny - ux, 0.0f, 1.0f), gstep(-2.0f, ny));
                            else return gmix(ux, gclamp(ny - (1.0f - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        }
                    }
                }
                // exiting left and right discontinuity
                if (dr == 0.75f && dg != 0.0f) {
                    ly = ly + ((0.5f - dg) * 8.0f - 1.0f) * syy;
                    bool left = discv(lx, ly, lz, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f);
                    bool right = discv(lx, ly, lz, 0.0f, 1.0f, 0.25f, 0.0f, 0.0f);
                    float a, b;
                    if (left && right) {
                        return gclamp(ny, 0.0f, 1.0f);
                    } else if (left || right) {
                        if (left) a = gmix(uy, 1.0f - uy, gstep(dg, 0.25f));
                        else a = gmix(1.0f - ux, gclamp(ny - ux, 0.0f, 1.0f), gstep(-2.0f, ny));
                        if (right) b = gmix(uy, 1.0f - uy, gstep(dg, 0.25f));
                        else b = gmix(ux, gclamp(ny - (1.0f - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmax(a, b);
                    } else {
                        if (dg == 0.75f) return 0.0f;
                        int guard = 0;
                        do {
                            left = discf(lx - sxx, ly, lz, 1.0f, 0.0f, 0.0f);
                            right = discf(lx + sxx, ly, lz, 1.0f, 0.0f, 0.0f);
                            if (!kAcc || lz <= tex(lx, ly)) break;
                            ly = ly + ((0.5f - dg) * 8.0f - 1.0f) * syy;
                        } while (!left && !right && ++guard < cap);
                        if (left && right) {
                            return gclamp(ny, 0.0f, 1.0f);
                        } else {
                            a = gmix(1.0f - ux, ux, gstep(1.0f, fbool(left)));
                            b = gmix(uy, 1.0f - uy, gstep(dg, 0.25f));
                            return gmax(a, b);
                        }
                    }
                }
                // exiting top and bottom discontinuity only
                if (dr == 0.0f && dg == 0.75f) {
                    bool top = discf(lx, ly - syy, lz, 0.0f, 1.0f, 0.0f);
                    bool bottom = discf(lx, ly + syy, lz, 0.0f, 1.0f, 0.0f);
                    if (bottom && top) {
                        return gclamp(nx, 0.0f, 1.0f);
                    } else if (bottom || top) {
                        if (!top) return gmix(1.0f - uy, gclamp(nx - uy, 0.0f, 1.0f), gstep(-2.0f, nx));
                        else return gmix(uy, gclamp(nx - (1.0f - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                    } else {
                        // columns +1, -1, +2, -2, ... until umbra above or below
                        int count = 0;
                        do {
                            const int mult = count / 2 + 1;
                            const float rx = (count % 2 == 0) ? lx + (float)mult * sxx : lx - (float)mult * sxx;
                            top = discf(rx, ly - syy, lz, 0.0f, 1.0f, 0.0f);
                            bottom = discf(rx, ly + syy, lz, 0.0f, 1.0f, 0.0f);
                            if (!kAcc || lz <= tex(lx, ly)) break;
                            ++count;
                        } while (!top && !bottom && count < cap);
                        if (top && bottom) {
                            return gclamp(nx, 0.0f, 1.0f);
                        } else {
                            if (!top) return gmix(1.0f - uy, gclamp(nx - uy, 0.0f, 1.0f), gstep(-2.0f, nx));
                            else return gmix(uy, gclamp(nx - (1.0f - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        }
                    }
                }
                // exiting top and bottom discontinuity
                if (dr != 0.0f && dg == 0.75f) {
                    lx = lx - ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                    bool bottom = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.5f, 0.0f);
                    bool top = discv(lx, ly, lz, 1.0f, 0.0f, 0.0f, 0.25f, 0.0f);
                    float a, b;
                    if (bottom && top) {
                        return gclamp(nx, 0.0f, 1.0f);
                    } else if (bottom || top) {
                        if (top) a = gmix(1.0f - ux, ux, gstep(dr, 0.25f));
                        else a = gmix(1.0f - uy, gclamp(nx - uy, 0.0f, 1.0f), gstep(-2.0f, nx));
                        if (bottom) b = gmix(1.0f - ux, ux, gstep(dr, 0.25f));
                        else b = gmix(uy, gclamp(nx - (1.0f - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        return gmax(a, b);
                    } else {
                        if (dr == 0.75f) return 0.0f;
                        int guard = 0;
                        do {
                            top = discf(lx, ly - syy, lz, 0.0f, 1.0f, 0.0f);
                            bottom = discf(lx, ly + syy, lz, 0.0f, 1.0f, 0.0f);
                            if (!kAcc || lz <= tex(lx, ly)) break;
                            lx = lx - ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                        } while (!top && !bottom && ++guard < cap);
                        if (top && bottom) {
                            return gclamp(nx, 0.0f, 1.0f);
                        } else {
                            a = gmix(1.0f - uy, uy, gstep(1.0f, fbool(top)));
                            b = gmix(1.0f - ux, ux, gstep(dr, 0.25f));
                            return gmax(a, b);
                        }
                    }
                }
            }
            // discontinuity in both axes (corner)
            if (dr > 0.0f && dg > 0.0f) {
                lx = lx + ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                const bool horizontal = discf(lx, ly, lz, 1.0f, 0.0f, db);
                lx = lx + ((0.5f - dr) * 8.0f - 1.0f) * sxx;
                ly = ly + ((0.5f - dg) * 8.0f - 1.0f) * syy;
                const bool vertical = discf(lx, ly, lz, 0.0f, 1.0f, db);
                if (horizontal && !vertical) {
                    dr = 0.0f;
                } else if (!horizontal && vertical) {
                    dg = 0.0f;
                } else if (!horizontal && !vertical) {
                    if (dg == 0.5f) return gclamp(uy - (1.0f - nx), 0.0f, 1.0f);
                    else if (dg == 0.25f) return gclamp((1.0f - uy) - (1.0f - nx), 0.0f, 1.0f);
                } else {
                    float a, b;
                    if (dr == 0.5f && dg == 0.5f) {
                        a = gmix(uy, gclamp(nx - (1.0f - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(1.0f - ux, gclamp(ny - ux, 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmax(a, b);
                    } else if (dr == 0.5f && dg == 0.25f) {
                        a = gmix(1.0f - uy, gclamp(nx - uy, 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(1.0f - ux, gclamp(ny - ux, 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmax(a, b);
                    } else if (dr == 0.25f && dg == 0.5f) {
                        a = gmix(uy, gclamp(nx - (1.0f - uy), 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(ux, gclamp(ny - (1.0f - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmax(a, b);
                    } else if (dr == 0.25f && dg == 0.25f) {
                        a = gmix(1.0f - uy, gclamp(nx - uy, 0.0f, 1.0f), gstep(-2.0f, nx));
                        b = gmix(ux, gclamp(ny - (1.0f - ux), 0.0f, 1.0f), gstep(-2.0f, ny));
                        return gmax(a, b);
                    }
                }
            }
            if (nx <= -2.0f) return gmix(uy, 1.0f - uy, gstep(dg, 0.25f));
            if (ny <= -2.0f) return gmix(1.0f - ux, ux, gstep(dr, 0.25f));
            if (dg > 0.0f) {
                if (dg == 0.5f) return gclamp(nx - (1.0f - uy), 0.0f, 1.0f);
                else return gclamp(nx - uy, 0.0f, 1.0f);
            }
            if (dr > 0.0f) {
                if (dr == 0.5f) return gclamp(ny - ux, 0.0f, 1.0f);
                else return gclamp(ny - (1.0f - ux), 0.0f, 1.0f);
            }
        }
        return 1.0f - db;
    }

    // RSMSS, RevectorizationBasedAccurateSoftShadow.frag:1175-1186: one tap on
    // a discontinuity, oriented, normalized and filled by the accurate
    // smoothONDS, then mixed with the shadow intensity.
    __device__ float rsmss(float x, float y, float z, float dr, float dg, float db, float si) const {
        const float ux = gfract(x * rb.D.fw), uy = gfract(y * rb.D.fh);
        const float left = length(dr, dg, db, x, y, z, -1.0f, 0.0f);
        const float right = length(dr, dg, db, x, y, z, 1.0f, 0.0f);
        const float down = length(dr, dg, db, x, y, z, 0.0f, -1.0f);
        const float up = length(dr, dg, db, x, y, z, 0.0f, 1.0f);
        const float nx = normalize(left, right, ux), ny = normalize(down, up, uy);
        return gmix(smooth<true>(x, y, z, nx, ny, dr, dg, db, ux, uy), 1.0f, si);
    }

    // One tap of RBSSM's kernel loop (:1202-1339) at loop indices (h, w):
    // revectorized (oriented, normalized, filled by smoothONDS) where the tap
    // sits on a discontinuity, else the plain depth test, mixed with si.
    __device__ float tap(const PcssUniforms& q, float si, float pw, float x, float y, float z, int h, int w) const {
        const float ty = y + ((float)h * pw) / q.kfw;
        const float tx = x + ((float)w * pw) / q.kfw;
        const float ux = gfract(tx * rb.D.fw), uy = gfract(ty * rb.D.fh);
        const float d = tex(tx, ty);
        float dr, dg, db;
        rb.discontinuity(tx, ty, z, d, dr, dg, db);
        if (dr > 0.0f || dg > 0.0f) {
            const float left = length(dr, dg, db, tx, ty, z, -1.0f, 0.0f);
            const float right = length(dr, dg, db, tx, ty, z, 1.0f, 0.0f);
            const float down = length(dr, dg, db, tx, ty, z, 0.0f, -1.0f);
            const float up = length(dr, dg, db, tx, ty, z, 0.0f, 1.0f);
            const float nx = normalize(left, right, ux), ny = normalize(down, up, uy);
            const float fill = smooth<false>(tx, ty, z, nx, ny, dr, dg, db, ux, uy);
            return gmix(fill, 1.0f, si);
        }
        return (z <= d) ? 1.0f : si;
    }
};

// Loop indices of RBSSM's kernel: `for (int h = int(-kfw); float(h) <= kfw;
// h++)`, the same for w.  Returns the count per axis (kernelSize for odd
// sizes, one less for even ones) and the first index in h0.
__device__ __forceinline__ int rbssm_side(const PcssUniforms& q, int& h0) {
    h0 = loop_start(q.kfw);
    int n = 0;
    for (int h = h0; (float)h <= q.kfw; ++h) ++n;
    return n;
}

// Workspace of the RBSSM pass: the list length (in a 256-byte header), then
// one (pixel, penumbra width) pair per view pixel.
constexpr size_t kListHeader = 256;
inline size_t rbssm_workspace_bytes(int w, int h) { return kListHeader + (size_t)w * h * 8u; }

// Clears a list's length before the classifying kernel appends to it (a
// kernel rather than a memset, so the pass can be captured in a graph).
__global__ void k_list_reset(unsigned* count) {
    if (threadIdx.x == 0) *count = 0u;
}

// RBSSM main :1362-1378 up to the kernel loop: pre-evaluation, blocker
// search, penumbra width.  Pixels that skip the loop get their value here;
// the others go to the list with their penumbra width for k_rbssm_taps.
__global__ __launch_bounds__(kBlockThreads) void k_rbssm_classify(GBufferView gb, DepthMap D, HardUniforms u,
                                                                   PcssUniforms q, float* __restrict__ out,
                                                                   unsigned* __restrict__ counters,
                                                                   uint2* __restrict__ list) {
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    bool loop = false;
    size_t k = 0;
    float pw = 0.0f;
    if (x < gb.w && y < gb.h) {
        const size_t plane = (size_t)gb.w * gb.h;
        k = (size_t)y * gb.w + x;
        const float vx = gb.pos[k];
        float shadow = 0.0f;  // background: the cleared target holds 0
        if (vx != 0.0f) {
            const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
            float sc[4];
            mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
            const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
            shadow = 1.0f;
            if (u.pre_eval)
                shadow = pre_evaluation(u, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
            if (sc[3] > 0.0f && shadow == 1.0f) {
                const float avg = pcss_blocker(D, q, n0, n1, n2);
                pw = pcss_penumbra(q, avg, n2);
                const float stepSize = 2.0f * pw / (float)q.ks;
                loop = !(stepSize <= 0.0f || stepSize >= 1.0f);
            }
        }
        if (!loop) out[k] = shadow;
    }
    __shared__ unsigned lds[2];
    const unsigned slot = block_append(loop, &counters[0], lds);
    if (loop) list[slot] = make_uint2((unsigned)k, __float_as_uint(pw));
}

// RBSSM's kernel loop for the listed pixels, a lane per tap: 64 / taps
// pixels per wave group when the kernel has at most 64 taps, else one pixel
// in passes of 64 taps.  The taps are summed in the loop's order (h, then w)
// and divided by kernelSize^2, as the shader does (:1337).
__global__ __launch_bounds__(kBlockThreads) void k_rbssm_taps(GBufferView gb, DepthMap D, HardUniforms u,
                                                               PcssUniforms q, float* __restrict__ out,
                                                               unsigned* __restrict__ counters,
                                                               const uint2* __restrict__ list) {
    const unsigned n = counters[0];
    int h0;
    const int side = rbssm_side(q, h0);
    const int ntaps = side * side;
    const int per = ntaps < 64 ? ntaps : 64;
    const int P = 64 / per;
    const int lane = (int)(threadIdx.x & 63);
    const int slot = lane / per, sl = lane - slot * per;
    const int g0 = (slot < P ? slot : P - 1) * per;
    const size_t plane = (size_t)gb.w * gb.h;
    const Rbssm r{Rbsm<false>{D, u.stepx, u.stepy, u.thr, u.max_search}, 0.0f, 0};
    for (WaveGroups wg(n, (unsigned)P); wg.more(); wg.advance()) {
        const unsigned qi = wg.base + (unsigned)slot;
        const bool has = slot < P && qi < n;
        const uint2 e = list[has ? qi : wg.base];
        const size_t k = e.x;
        const float pw = __uint_as_float(e.y);
        const float vx = gb.pos[k], vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
        float sc[4];
        mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
        const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
        float illum = 0.0f;
        for (int t0 = 0; t0 < ntaps; t0 += per) {
            const int t = t0 + sl;
            float v = 0.0f;
            if (has && t < ntaps) {
                const int th = t / side;
                v = r.tap(q, u.si, pw, n0, n1, n2, h0 + th, h0 + (t - th * side));
            }
            illum = ordered_sum(illum, v, g0, ntaps - t0 < per ? ntaps - t0 : per);
        }
        if (has && sl == 0) out[k] = illum / (float)(q.ks * q.ks);
    }
}

// ---------------------------------------------------------------------------
// Revectorization-based adaptive sampling (ShadowParams.
// revectorizationBasedAdaptiveSampling, SoftShadowMapping/src/main.cpp:440-571
// and :813-892).
// ---------------------------------------------------------------------------

// Discontinuity.frag writes (r, g, b, 1) into a GL_RGBA8 layer; r and g are
// multiples of 0.25, b is 0 or 1.  One byte holds the same information:
// bits 0-1 r * 4, bits 2-3 g * 4, bit 4 b, bit 5 the alpha of an evaluated
// pixel.  Background pixels keep the layer's clear colour (1, 1, 1, 1).
constexpr unsigned char kDiscBackground = 0xFF;

__device__ __forceinline__ unsigned char disc_encode(float r, float g, float b) {
    return (unsigned char)((int)(r * 4.0f) | ((int)(g * 4.0f) << 2) | (b != 0.0f ? 16 : 0) | 32);
}

// correctDiscontinuity (RevectorizationBasedAccurateSoftShadow.frag:1188-1205)
// restores the exact quarters from the 8-bit values; from the byte code they
// are exact already.
__device__ __forceinline__ void disc_decode(unsigned char c, float& r, float& g, float& b) {
    r = (float)(c & 3) * 0.25f;
    g = (float)((c >> 2) & 3) * 0.25f;
    b = (float)((c >> 4) & 1);
}

// Discontinuity.frag:66-88: the discontinuity of the G-buffer pixel in one
// corner map (computeDiscontinuity when the pre-evaluation lets it through and
// w > 0, else 0).
__global__ __launch_bounds__(kBlockThreads) void k_rev_discontinuity(GBufferView gb, DepthMap D, HardUniforms u,
                                                                      unsigned char* __restrict__ layer) {
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    const float vx = gb.pos[k];
    if (vx == 0.0f) {
        layer[k] = kDiscBackground;
        return;
    }
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    float sc[4];
    mat4_vec4(u.light_mvp.m, vx, vy, vz, 1.0f, sc);
    const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
    const float shadow = pre_evaluation(u, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
    unsigned char code = 0;
    if (sc[3] > 0.0f && shadow == 1.0f) {
        const Rbsm<false> rb{D, u.stepx, u.stepy, u.thr, u.max_search};
        float dr, dg, db;
        rb.discontinuity(n0, n1, n2, D.fetch(n0, n1), dr, dg, db);
        code = disc_encode(dr, dg, db);
    }
    layer[k] = code;
}

struct RevQtUniforms {
    QtUniforms q;
    int level;        // quadTreeLevel of the node under test
    int disc_layers;  // layers of the discontinuity array; reads clamp to the last
};

// QuadTree/RevectorizationBasedReprojection.frag:14-51: lit-corner count as
// Reprojection.frag, forced to 5 (never a penumbra pixel for Evaluation.frag)
// where two or more corner maps see a discontinuity at the pixel, and the
// visibility classification (0 umbra, 1 lit, 0.5 penumbra) started at level
// 0 and demoted to 0.5 where a deeper node disagrees.  Each pixel reads and
// writes only its own classification, so the update is in place (the
// reference renders a copy and copies it back).  Block (0, 0) clears the
// node's pass counter for k_qt_evaluate.
__global__ __launch_bounds__(kBlockThreads) void k_rev_qt_reproject(GBufferView gb, const float* __restrict__ maps,
                                                                     int smw, int smh, RevQtUniforms u,
                                                                     const unsigned char* __restrict__ disc,
                                                                     float* __restrict__ img,
                                                                     float* __restrict__ vis,
                                                                     unsigned* __restrict__ count) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *count = 0u;
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    if (x >= gb.w || y >= gb.h) return;
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t k = (size_t)y * gb.w + x;
    const float vx = gb.pos[k];
    if (vx == 0.0f) {  // both targets keep the clear colour
        img[k] = kQtBackground;
        vis[k] = kQtBackground;
        return;
    }
    const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
    const size_t mapsz = (size_t)smw * smh;
    float sh[4];
    int ndisc = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float sc[4];
        mat4_vec4(u.q.m[i].m, vx, vy, vz, 1.0f, sc);
        const float n0 = sc[0] / sc[3], n1 = sc[1] / sc[3], n2 = sc[2] / sc[3];
        DepthMap D{maps + (size_t)u.q.map[i] * mapsz, smw, smh, (float)smw, (float)smh};
        sh[i] = (n2 <= D.fetch(n0, n1)) ? 1.0f : 0.0f;
        const int layer = u.q.map[i] < u.disc_layers ? u.q.map[i] : u.disc_layers - 1;
        if ((disc[(size_t)layer * plane + k] & 15) != 0) ++ndisc;
    }
    float sum = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    if (ndisc >= 2) sum = 5.0f;
    float v = vis[k];
    if (u.level == 0) {
        v = (sum == 0.0f || sum == 4.0f) ? sum / 4.0f : 0.5f;
    } else if ((v == 0.0f && sum != 0.0f) || (v == 1.0f && sum != 4.0f)) {
        v = 0.5f;
    }
    img[k] = sum;
    vis[k] = v;
}

struct RevSoftUniforms {
    float m[12];      // rows 0-3 of columns 0-2 of B*lightMVP
    HardUniforms h;   // pre-evaluation (mv, nm, lpos, si), maxSearch, depthThreshold, shadowMapStep
    int nsamples;
    int disc_layers;
    int cap;          // scan bound of the accurate smoothONDS
};

// One light sample of RevectorizationBasedAccurateSoftShadow.frag:1275-1325
// for the pixel at world position v: the sample's shadow value times its
// quad-tree weight, and the weight.  kRpcf == false: one tap, revectorized
// where the stored discontinuity of that sample is set; kRpcf == true: a
// 3 x 3 tap kernel, each tap's discontinuity computed on the fly and
// revectorized.
template <bool kRpcf>
__device__ __forceinline__ void rev_sample(const float* __restrict__ maps, int smw, int smh,
                                           const float4* __restrict__ trans, const unsigned char* __restrict__ disc,
                                           size_t plane, size_t k, const RevSoftUniforms& u, const float cm[4], int i,
                                           float& term, float& af) {
    const float si = u.h.si;
    float4 T = trans[i];
    af = adaptive_weight(T);
    const float cx = cm[0] + T.x, cy = cm[1] + T.y, cz = cm[2] + T.z, cw = cm[3] + T.w;
    const float sx = cx / cw, sy = cy / cw, sz = cz / cw;
    const DepthMap D{maps + (size_t)i * ((size_t)smw * smh), smw, smh, (float)smw, (float)smh};
    const Rbssm r{Rbsm<false>{D, u.h.stepx, u.h.stepy, u.h.thr, u.h.max_search}, 0.0f, u.cap};
    float value;
    if constexpr (!kRpcf) {
        const float d = D.fetch(sx, sy);
        const int layer = i < u.disc_layers ? i : u.disc_layers - 1;
        float dr, dg, db;
        disc_decode(disc[(size_t)layer * plane + k], dr, dg, db);
        if (dr > 0.0f || dg > 0.0f)
            value = r.rsmss(sx, sy, sz, dr, dg, db, si);
        else
            value = (sz <= d) ? 1.0f : si;
    } else {
        const float incrWidth = 1.0f / D.fw, incrHeight = 1.0f / D.fh;
        float illum = 0.0f;
        int n = 0;
        for (float w = -1.0f; w <= 1.0f; w += 1.0f) {
            for (float hh = -1.0f; hh <= 1.0f; hh += 1.0f) {
                const float tx = sx + w * incrWidth, ty = sy + hh * incrHeight;
                const float d = D.fetch(tx, ty);
                float dr, dg, db;
                r.rb.discontinuity(tx, ty, sz, d, dr, dg, db);
                if (dr > 0.0f || dg > 0.0f)
                    illum = illum + r.rsmss(tx, ty, sz, dr, dg, db, si);
                else
                    illum = illum + ((sz <= d) ? 1.0f : si);
                ++n;
            }
        }
        value = illum / (float)n;
    }
    term = value * af;
}

// Workspace of the revectorization-based final pass: the list length (in a
// 256-byte header), then one pixel index per view pixel.
inline size_t rev_soft_workspace_bytes(int w, int h) { return kListHeader + (size_t)w * h * 4u; }

// RevectorizationBasedAccurateSoftShadow.frag:1252-1331 up to the sample
// loop.  Pixels the pre-evaluation lets through take their classification
// (clamped to [si, 1]); penumbra pixels (0.5) go to the list for
// k_rev_soft_samples.
__global__ __launch_bounds__(kBlockThreads) void k_rev_soft_classify(GBufferView gb, const float* __restrict__ vis,
                                                                      RevSoftUniforms u, float* __restrict__ out,
                                                                      unsigned* __restrict__ counters,
                                                                      unsigned* __restrict__ list) {
    int x, y;
    pixel_of_block(gb.w, gb.h, x, y);
    bool pen = false;
    size_t k = 0;
    if (x < gb.w && y < gb.h) {
        const size_t plane = (size_t)gb.w * gb.h;
        k = (size_t)y * gb.w + x;
        const float vx = gb.pos[k];
        float shadow = 0.0f;  // discard: the cleared target holds 0
        if (vx != 0.0f) {
            const float vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
            shadow = pre_evaluation(u.h, vx, vy, vz, gb.nrm[k], gb.nrm[plane + k], gb.nrm[2 * plane + k]);
            if (shadow == 1.0f) shadow = gclamp(vis[k], u.h.si, 1.0f);
            pen = shadow == 0.5f;
        }
        if (!pen) out[k] = shadow;
    }
    __shared__ unsigned lds[2];
    const unsigned slot = block_append(pen, &counters[0], lds);
    if (pen) list[slot] = (unsigned)k;
}

// The sample loop of the listed penumbra pixels, a lane per light sample:
// 64 / nsamples pixels per wave group when nsamples <= 64, else one pixel in
// passes of 64 samples.  Penumbra pixels are a small fraction of the view
// (those along shadow edges), so a lane per pixel would leave most of the
// GPU idle while a few waves run every sample.  The weighted values and the
// weights are summed in sample order, the shader's `acc += v * w; count += w`
// operation for operation.
template <bool kRpcf>
__global__ __launch_bounds__(kBlockThreads) void k_rev_soft_samples(GBufferView gb, const float* __restrict__ maps,
                                                                     int smw, int smh, const float4* __restrict__ trans,
                                                                     const unsigned char* __restrict__ disc,
                                                                     RevSoftUniforms u, float* __restrict__ out,
                                                                     unsigned* __restrict__ counters,
                                                                     const unsigned* __restrict__ list) {
    const unsigned n = counters[0];
    const int S = u.nsamples;
    const int per = S < 64 ? S : 64;
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
        float cm[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float a = u.m[0 * 4 + i] * vx;
            a = a + u.m[1 * 4 + i] * vy;
            a = a + u.m[2 * 4 + i] * vz;
            cm[i] = a;
        }
        float acc = 0.0f, count = 0.0f;
        for (int i0 = 0; i0 < S; i0 += per) {
            const int i = i0 + sl;
            float term = 0.0f, af = 0.0f;
            if (has && i < S) rev_sample<kRpcf>(maps, smw, smh, trans, disc, plane, k, u, cm, i, term, af);
            const int cnt = S - i0 < per ? S - i0 : per;
            acc = ordered_sum(acc, term, g0, cnt);
            count = ordered_sum(count, af, g0, cnt);
        }
        if (has && sl == 0) out[k] = acc / count;
    }
}

// The RPCF sample loop of the listed penumbra pixels, a lane per tap of each
// sample's 3 x 3 kernel.  Lane per sample ran each sample's nine taps (and
// their edge walks) one after another on a few thousand waves; here a wave
// holds 7 (pixel, sample) pairs of 9 lanes and a workgroup 28 pairs, so the
// same work spreads over nine times as many waves with one tap each.
//
// A workgroup takes units of Q = 28 / nsamples whole pixels (one pixel when
// nsamples > 28, in passes of 28 samples).  Each pair's nine taps are summed
// in the loop's order (w outer, h inner) by shuffles and divided by 9; the
// pair's weighted value and weight go to LDS, and one lane per pixel adds
// them in sample order, `acc += v * w; count += w` as the shader does.
constexpr int kRevPairsPerWave = 7;
constexpr int kRevPairs = kRevPairsPerWave * (kBlockThreads / 64);
__global__ __launch_bounds__(kBlockThreads) void k_rev_soft_taps(GBufferView gb, const float* __restrict__ maps,
                                                                  int smw, int smh,
                                                                  const float4* __restrict__ trans,
                                                                  RevSoftUniforms u, float* __restrict__ out,
                                                                  const unsigned* __restrict__ counters,
                                                                  const unsigned* __restrict__ list) {
    __shared__ float lterm[kRevPairs], laf[kRevPairs];
    const unsigned n = counters[0];
    const int S = u.nsamples;
    const int Q = S <= kRevPairs ? kRevPairs / S : 1;  // pixels per unit
    const int pairs = Q * S;                              // pairs per unit
    const unsigned units = (n + (unsigned)Q - 1u) / (unsigned)Q;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int pw = lane / 9, tap = lane - pw * 9;  // pair within the wave, tap within the pair
    const int slot = wave * kRevPairsPerWave + pw;  // pair within the pass (lane 63: none)
    const int g0 = (pw < kRevPairsPerWave ? pw : kRevPairsPerWave - 1) * 9;
    const float tw = (float)(tap / 3) - 1.0f, th = (float)(tap % 3) - 1.0f;  // the loop's w and h
    const size_t plane = (size_t)gb.w * gb.h;
    const size_t mapsz = (size_t)smw * smh;
    const float fw = (float)smw, fh = (float)smh;
    const float incrWidth = 1.0f / fw, incrHeight = 1.0f / fh;
    const int owner = (int)threadIdx.x;  // thread q < Q sums pixel q of the unit
    float acc = 0.0f, count = 0.0f;
    for (unsigned unit = blockIdx.x; unit < units; unit += gridDim.x) {
        for (int p0 = 0; p0 < pairs; p0 += kRevPairs) {
            const int pl = p0 + slot;
            const unsigned q = unit * (unsigned)Q + (unsigned)(pl / S);
            const bool has = pw < kRevPairsPerWave && pl < pairs && q < n;
            float v = 0.0f, af = 0.0f;
            if (has) {
                const size_t k = list[q];
                const int i = pl % S;
                const float vx = gb.pos[k], vy = gb.pos[plane + k], vz = gb.pos[2 * plane + k];
                float cm[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float a = u.m[0 * 4 + c] * vx;
                    a = a + u.m[1 * 4 + c] * vy;
                    a = a + u.m[2 * 4 + c] * vz;
                    cm[c] = a;
                }
                float4 T = trans[i];
                af = adaptive_weight(T);
                const float cx = cm[0] + T.x, cy = cm[1] + T.y, cz = cm[2] + T.z, cw = cm[3] + T.w;
                const float sx = cx / cw, sy = cy / cw, sz = cz / cw;
                const DepthMap D{maps + (size_t)i * mapsz, smw, smh, fw, fh};
                const Rbssm r{Rbsm<false>{D, u.h.stepx, u.h.stepy, u.h.thr, u.h.max_search}, 0.0f, u.cap};
                const float tx = sx + tw * incrWidth, ty = sy + th * incrHeight;
                const float d = D.fetch(tx, ty);
                float dr, dg, db;
                r.rb.discontinuity(tx, ty, sz, d, dr, dg, db);
                if (dr > 0.0f || dg > 0.0f)
                    v = r.rsmss(tx, ty, sz, dr, dg, db, u.h.si);
                else
                    v = (sz <= d) ? 1.0f : u.h.si;
            }
            const float illum = ordered_sum(0.0f, v, g0, 9);
            if (has && tap == 0) {
                lterm[slot] = (illum / 9.0f) * af;
                laf[slot] = af;
            }
            __syncthreads();
            if (owner < Q) {
                // this pixel's pairs in this pass, in sample order
                const int a0 = owner * S, a1 = a0 + S;
                const int lo = a0 > p0 ? a0 : p0, hi = a1 < p0 + kRevPairs ? a1 : p0 + kRevPairs;
                const unsigned qo = unit * (unsigned)Q + (unsigned)owner;
                if (qo < n) {
                    if (lo == a0) acc = 0.0f, count = 0.0f;
                    for (int e = lo; e < hi; ++e) {
                        acc = acc + lterm[e - p0];
                        count = count + laf[e - p0];
                    }
                    if (hi == a1) out[list[qo]] = acc / count;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace
}  // namespace gi
