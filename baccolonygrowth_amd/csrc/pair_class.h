// pair_class.h - FP32 decisions about a pair of cells: the list build's
// end-point prefilter (PreCol), the neighbour-list
// predicate (classify_pair) and "surely not in contact" (surely_apart).
//
// reduceNeighbours (ABM/Neighbours.cpp:72-88) keeps neighbour B of cell A when
// an end point of B lies within re = L_A/2 + Rc R_A of A's centre and the
// segment distance is below rd = Rc R_A, both compared in FP64. The list build
// decides most pairs from the FP32 candidate records instead, and hands the
// rest to the FP64 predicate. The function is host-callable so that a CPU
// harness can check it against the oracle on many pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace bcg {

// The list build's FP32 prefilter: the first test of reduceNeighbours
// (ABM/Neighbours.cpp:69-73), an end point e of the candidate within
// r = L_A/2 + Rc R_A of the centre c of A, made conservative. End points are
// stored relative to the origin o of their bin column (prefilter_point: x, y,
// z and w = x^2 + y^2 + z^2 of the rounded coordinates), so that against one
// column
//   |e - c'|^2 - r'^2 = w - 2 x c'_x - 2 y c'_y - 2 z c'_z + (|c'|^2 - r'^2),
// c' = c - o: three FMAs and an add per end point, the column's constants in
// PreCol, and the verdict is the sign bit of the smaller of the two sums.
// Conservative: r carries 1e-3 for its own rounding (prefilter_reach),
// r'^2 = 1.0001 r^2 + 4e-6 E^2 with E = |c'_x| + |c'_y| + |c'_z| + 2r. For an
// end point within reach every partial sum is below 5 E^2 in magnitude, so
// the five roundings of the sum and the roundings of c0 and of the stored
// coordinates stay below 30 units of 2^-24 E^2 (1.8e-6 E^2); the exact FP64
// test decides every pair this keeps.
__host__ __device__ inline float4 prefilter_point(double x, double y, double z) {
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    return make_float4(fx, fy, fz, (float)((double)fx * fx + (double)fy * fy + (double)fz * fz));
}
__host__ __device__ inline float prefilter_reach(double re) { return (float)re + 1e-3f; }
struct PreCol {
    float m2x, m2y, m2z, c0;
    // (in the list build each FMA takes one operand from the record, which
    // its walk holds in scalar registers)
    __host__ __device__ float key(const float4 p, const float4 q) const {
        const float sp = fmaf(p.z, m2z, fmaf(p.y, m2y, fmaf(p.x, m2x, c0))) + p.w;
        const float sq = fmaf(q.z, m2z, fmaf(q.y, m2y, fmaf(q.x, m2x, c0))) + q.w;
        return fminf(sp, sq);
    }
};
// The constants of A's test against one column: cx, cy its centre relative to
// the column origin, cz its z, r its reach (prefilter_reach).
__host__ __device__ inline PreCol prefilter_column(float cx, float cy, float cz, float r) {
    const float e = fabsf(cx) + fabsf(cy) + fabsf(cz) + 2.0f * r;
    PreCol c;
    c.m2x = -2.0f * cx;
    c.m2y = -2.0f * cy;
    c.m2z = -2.0f * cz;
    c.c0 = (cx * cx + cy * cy + cz * cz) - r * r * 1.0001f - 4e-6f * (e * e);
    return c;
}

// 1 keep, 0 reject, -1 undecided (the FP64 predicate decides).
// a_p, a_q: A's end points in the frame of A's bin column (x, y relative to the
// column origin, z as stored); b_p, b_q: B's end points in the frame of B's
// column, moved into A's by (sx, sy); re, rd as above; rcp: a reciprocal,
// exact or approximate (the bounds below hold for any parameters it yields).
//
// End points: |e - c|^2 from the differences. Segment distance: any (s, t) in
// [0,1]^2 gives the upper bound f(s, t) = |A(s) - B(t)|^2, and since f is a
// convex quadratic, f(s, t) + the minimum over the box of its gradient's
// linear term,
//   f + min(-g_s s, g_s (1 - s)) + min(-g_t t, g_t (1 - t)),
// is a lower bound of its minimum over the box. Near the closest points both
// bounds are tight (the gradient vanishes or points out of the box there),
// also for near-parallel segments whose FP32 parameters are poorly
// determined.
//
// Tolerances: 1e-4 of the squared thresholds (the FP64 reference's rounding
// and its root comparisons are far below it) plus 1e-5 E^2, E the largest
// coordinate magnitude of the four end points; the FP32 rounding of the stored
// coordinates and of these sums is a few tens of 2^-24 E^2.
// FMA helpers: the bounds below carry their own margins, so contraction is
// free here (fmaf rounds once on the host and on the device alike).
__host__ __device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return fmaf(az, bz, fmaf(ay, by, ax * bx));
}

__host__ __device__ inline float coord_max(float4 a_p, float4 a_q, float4 b_p, float4 b_q) {
    return fmaxf(fmaxf(fmaxf(fmaxf(fabsf(a_p.x), fabsf(a_p.y)), fmaxf(fabsf(a_p.z), fabsf(a_q.x))),
                       fmaxf(fmaxf(fabsf(a_q.y), fabsf(a_q.z)), fmaxf(fabsf(b_p.x), fabsf(b_p.y)))),
                 fmaxf(fmaxf(fabsf(b_p.z), fabsf(b_q.x)), fmaxf(fabsf(b_q.y), fabsf(b_q.z))));
}

// Upper and lower bounds (ub, lb) of the squared distance between segments
// A = [a_p, a_q] and B = [b_p, b_q] (see above): (s, t) is the line-line
// parameter s0, t optimal for it and clamped, s optimal for that t and
// clamped (the closest points when the segments' closest pair is interior or
// has t at an end; near them otherwise, which only loosens the bounds).
// Returns false for (near-)degenerate segments (no bounds).
template <class Rcp>
__host__ __device__ inline bool seg_dist2_bounds(float4 a_p, float4 a_q, float4 b_p, float4 b_q, float tol, Rcp rcp,
                                                 float& ub, float& lb) {
    const float v1x = a_q.x - a_p.x, v1y = a_q.y - a_p.y, v1z = a_q.z - a_p.z;
    const float v2x = b_q.x - b_p.x, v2y = b_q.y - b_p.y, v2z = b_q.z - b_p.z;
    const float rx = a_p.x - b_p.x, ry = a_p.y - b_p.y, rz = a_p.z - b_p.z;
    const float l1 = dot3(v1x, v1y, v1z, v1x, v1y, v1z), l2 = dot3(v2x, v2y, v2z, v2x, v2y, v2z);
    const bool ok = (l1 > 1e-4f * tol) && (l2 > 1e-4f * tol);
    const float f = dot3(v2x, v2y, v2z, rx, ry, rz), c = dot3(v1x, v1y, v1z, rx, ry, rz);
    const float b = dot3(v1x, v1y, v1z, v2x, v2y, v2z);
    const float den = fmaf(l1, l2, -b * b);
    auto clamp01 = [](float x) { return fmaxf(0.0f, fminf(1.0f, x)); };  // NaN -> 1
    const float s0 = den > 0.0f ? clamp01(fmaf(b, f, -c * l2) * rcp(den)) : 0.0f;
    const float t = clamp01(fmaf(b, s0, f) * rcp(l2));
    const float s = clamp01(fmaf(b, t, -c) * rcp(l1));
    const float wx = fmaf(-t, v2x, fmaf(s, v1x, rx));
    const float wy = fmaf(-t, v2y, fmaf(s, v1y, ry));
    const float wz = fmaf(-t, v2z, fmaf(s, v1z, rz));
    ub = dot3(wx, wy, wz, wx, wy, wz);
    const float gs = 2.0f * dot3(wx, wy, wz, v1x, v1y, v1z);
    const float gt = -2.0f * dot3(wx, wy, wz, v2x, v2y, v2z);
    lb = ub + fminf(-gs * s, fmaf(-gs, s, gs)) + fminf(-gt * t, fmaf(-gt, t, gt));
    return ok;
}

template <class Rcp>
__host__ __device__ inline int classify_pair(float4 a_p, float4 a_q, float re, float rd, float4 b_p, float4 b_q,
                                             float sx, float sy, Rcp rcp) {
    // straight-line code (no early exits): a wave runs it for 64 pairs at once
    b_p.x += sx;
    b_p.y += sy;
    b_q.x += sx;
    b_q.y += sy;
    const float E = coord_max(a_p, a_q, b_p, b_q);
    const float tol = 1e-5f * E * E;
    // first test: an end point of B within re of A's centre
    const float cx = 0.5f * (a_p.x + a_q.x), cy = 0.5f * (a_p.y + a_q.y), cz = 0.5f * (a_p.z + a_q.z);
    const float px = b_p.x - cx, py = b_p.y - cy, pz = b_p.z - cz;
    const float qx = b_q.x - cx, qy = b_q.y - cy, qz = b_q.z - cz;
    const float e2 = fminf(dot3(px, py, pz, px, py, pz), dot3(qx, qy, qz, qx, qy, qz));
    const float re2 = re * re;
    const bool end_no = e2 > fmaf(re2, 1.0001f, tol);
    const bool end_yes = e2 < fmaf(re2, 0.9999f, -tol);
    // second test: the segment distance below rd
    float ub, lb;
    const bool ok = seg_dist2_bounds(a_p, a_q, b_p, b_q, tol, rcp, ub, lb);
    const float rd2 = rd * rd;
    const bool dist_no = lb > fmaf(rd2, 1.0001f, tol);
    const bool dist_yes = ub < fmaf(rd2, 0.9999f, -tol);
    if (end_no) return 0;
    if (!ok) return -1;  // (near-)degenerate segments: the FP64 predicate decides, whatever it does
    if (dist_no) return 0;
    if (end_yes && dist_yes) return 1;
    return -1;
}

// True when the segments' FP64 distance (ABM/Neighbours.cpp:107-160) is surely
// at least sigma = R_A + R_B, i.e. the cells do not touch and F_cc (ABM/Forces.cpp:
// 13-84) adds only signed zeros. End points relative to a common origin near the
// pair (FP32 of the FP64 differences); margins as in classify_pair.
template <class Rcp>
__host__ __device__ inline bool surely_apart(float4 a_p, float4 a_q, float4 b_p, float4 b_q, float sigma, Rcp rcp) {
    const float E = coord_max(a_p, a_q, b_p, b_q);
    const float tol = 1e-5f * E * E;
    float ub, lb;
    const bool ok = seg_dist2_bounds(a_p, a_q, b_p, b_q, tol, rcp, ub, lb);
    return ok && lb > fmaf(sigma * sigma, 1.0001f, tol);
}

}  // namespace bcg
