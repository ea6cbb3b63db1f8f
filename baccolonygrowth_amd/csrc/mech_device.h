// mech_device.h - device physics of the mechanics step (gfx950, FP64).
//
// Each function restates one reference function (file:line relative to
// AgentBasedModel-BacColonyGrowth/). Expressions keep the reference's
// evaluation order and the library is built with -ffp-contract=off, so +,-,*,
// /, sqrt round exactly as in the reference. The reference's pow(delta, 3/2)
// is evaluated as delta * sqrt(delta) (pow15), within an ulp or so of glibc's
// pow.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "../../include/bcg.h"

namespace bcg {

constexpr double kPI = 3.141592653589793238462643383279502884197;  // ABM/tools.h:12

// Hot record: the fields the force gather reads, one 128-byte line per cell.
//   [0..2] p   [3..5] q   [6..8] Velocity   [9..11] AngularVelocity
//   [12] Length  [13] Radius  [14] Ldiv  [15] GrowthRate
constexpr int HOT = 16;

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 scl(d3 a, double c) { return mk(a.x * c, a.y * c, a.z * c); }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// ABM/tools.h:121-124
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double clampd(double n, double lo, double hi) {  // ABM/Neighbours.cpp:95
    return n < lo ? lo : (n > hi ? hi : n);
}

// Cell geometry + kinematics held in registers.
struct Seg {
    d3 p, q;
    double L, R;
};
struct Kin {
    d3 v, w;
};

__device__ __forceinline__ d3 cm_of(const Seg& s) { return scl(add(s.p, s.q), 0.5); }  // ABM/tools.h:149

__device__ __forceinline__ void load_seg(const double* __restrict__ hot, int id, Seg& s) {
    const double2* r = reinterpret_cast<const double2*>(hot + (size_t)id * HOT);
    double2 a = r[0], b = r[1], c = r[2], g = r[6];
    s.p = mk(a.x, a.y, b.x);
    s.q = mk(b.y, c.x, c.y);
    s.L = g.x;
    s.R = g.y;
}
__device__ __forceinline__ void load_kin(const double* __restrict__ hot, int id, Kin& k) {
    const double2* r = reinterpret_cast<const double2*>(hot + (size_t)id * HOT);
    double2 a = r[3], b = r[4], c = r[5];
    k.v = mk(a.x, a.y, b.x);
    k.w = mk(b.y, c.x, c.y);
}

// sqrt(x) < t (x >= 0, t > 0 normal) exactly as the FP64 expression with a
// correctly rounded square root decides it: when x is below t^2 (1 - 2^-48)
// the rounded root is below t(1 - 2^-50) < t, and above t^2 (1 + 2^-48) it
// exceeds t; only in between is the root taken.
__device__ __forceinline__ bool sqrt_lt(double x, double t) {
    const double t2 = t * t;
    bool r = x < t2 * (1.0 - 0x1p-48);
    const bool amb = !r & (x <= t2 * (1.0 + 0x1p-48));
    if (amb) r = sqrt(x) < t;
    return r;
}

// Segment-segment closest points: ABM/Neighbours.cpp:107-160 (the squared
// distance; min_distance takes the root as the reference does).
__device__ __forceinline__ double min_distance_sq(const Seg& A, const Seg& B, d3& c1, d3& c2) {
    d3 v1 = sub(A.q, A.p), v2 = sub(B.q, B.p), r = sub(A.p, B.p);
    double l1 = dot(v1, v1), l2 = dot(v2, v2), f = dot(v2, r), c = dot(v1, r), b = dot(v1, v2);
    double denom = l1 * l2 - b * b;
    double s, t;
    if (denom != 0) s = clampd((b * f - c * l2) / denom, 0, 1);
    else s = 0;
    double tnom = b * s + f;
    if (tnom < 0) {
        t = 0;
        s = clampd(-c / l1, 0, 1);
    } else if (tnom > l2) {
        t = 1;
        s = clampd((b - c) / l1, 0, 1);
    } else {
        t = tnom / l2;
    }
    c1 = add(A.p, scl(v1, s));
    c2 = add(B.p, scl(v2, t));
    d3 cv = sub(c1, c2);
    return dot(cv, cv);
}
__device__ __forceinline__ double min_distance(const Seg& A, const Seg& B, d3& c1, d3& c2) {
    return sqrt(min_distance_sq(A, B, c1, c2));
}

// reduceNeighbours predicate for neighbour B of cell A: ABM/Neighbours.cpp:72-88.
// The square roots are compared through sqrt_lt, which decides each
// comparison as the reference's sqrt(...) < threshold does.
__device__ __forceinline__ bool keep_neighbour(const Seg& A, d3 cmA, const Seg& B, double Rc) {
    const double thr = A.L / 2.0 + Rc * A.R;
    d3 v = sub(B.p, cmA);
    const double dp2 = dot(v, v);
    v = sub(B.q, cmA);
    const double dq2 = dot(v, v);
    if (sqrt_lt(dp2, thr) || sqrt_lt(dq2, thr)) {
        d3 c1, c2;
        return sqrt_lt(min_distance_sq(A, B, c1, c2), Rc * A.R);
    }
    return false;
}

// delta^(3/2) for delta >= 0 (the reference's pow(delta, 3.0 / 2.0)): a
// correctly rounded root and one rounded product, within about one ulp of
// the exact power, as glibc's pow is.
__device__ __forceinline__ double pow15(double delta) { return delta * sqrt(delta); }

struct PhysParams {
    double k_cc, k_wc, gamma_t, cell_mu, wall_mu, viscosity, tension, DH, cellRadius, BoxLength, Rc;
    double L_divide, varL, varAngle;
    int BoxX, BoxY, BoxZ, ref, nhx, nhy, planar, crossFeeding;
};

// UniformGrid::GetAddress (ABM/UniformGrid.h:87-100); -1 when outside. A
// NaN coordinate is outside too: x86-64 converts it to INT_MIN (outside),
// the GPU's conversion to 0, which would pass it as box 0.
__device__ __forceinline__ int bin_of(const PhysParams& P, d3 pos, int& x, int& y, int& z) {
    x = (int)(pos.x / P.BoxLength + P.BoxX / 2);
    y = (int)(pos.y / P.BoxLength + P.BoxY / 2 + 0.5);
    z = (int)(pos.z / P.BoxLength);
    const bool nan = (pos.x != pos.x) | (pos.y != pos.y) | (pos.z != pos.z);
    if (nan | !((x < P.BoxX) & (y < P.BoxY) & (z < P.BoxZ) & (x >= 0) & (y >= 0) & (z >= 0))) return -1;
    return x * (P.BoxY * P.BoxZ) + y * P.BoxZ + z;
}

// F_cc: ABM/Forces.cpp:13-84. The force F of neighbour B on A and its point
// of action relative to A's centre, r - cmA. The kinematics come from kinA()
// and kinB(), called only if the cells touch.
template <class KA_, class KB_>
__device__ __forceinline__ void F_cc_r(const PhysParams& P, const Seg& A, d3 cmA, const Seg& B, KA_ kinA, KB_ kinB,
                                       d3& F, d3& r) {
    const double gamma_n = 100.0;
    d3 c1, c2;
    double sigma = A.R + B.R;
    double L1 = A.L + 4.0 / 3 * A.R;
    double M1 = L1 * A.R * A.R * kPI;
    double L2 = B.L + 4.0 / 3 * B.R;
    double M2 = L2 * B.R * B.R * kPI;
    double Meff = M1 * M2 / (M1 + M2);
    double d = min_distance(A, B, c1, c2);
    double delta = dmax(sigma - d, 0.0);
    d3 vn = scl(sub(c2, c1), 1.0 / d);
    double FE_mag = sqrt(A.R * B.R / (A.R + B.R)) * P.k_cc * pow15(delta);
    F = scl(vn, -FE_mag);
    r = scl(add(scl(c1, B.R), scl(c2, A.R)), 1 / sigma);
    if (FE_mag > DBL_EPSILON) {
        const Kin KA = kinA();
        const Kin KB = kinB();
        d3 rcm1 = sub(r, cmA), rcm2 = sub(r, cm_of(B));
        d3 va1 = cross(KA.w, rcm1), va2 = cross(KB.w, rcm2);
        d3 v1 = add(KA.v, va1), v2 = add(KB.v, va2);
        d3 dv = sub(v2, v1);
        d3 dvn = scl(vn, dot(dv, vn));
        d3 FDn = scl(dvn, gamma_n * Meff * delta);
        d3 dvt = sub(dv, dvn);
        double v_mag = sqrt(dot(dvt, dvt));
        d3 FDt = mk(0.0, 0.0, 0.0);
        if (v_mag > 3 * DBL_EPSILON) FDt = scl(dvt, dmin(P.gamma_t * Meff * sqrt(delta), P.cell_mu * FE_mag / v_mag));
        F = add(add(F, FDt), FDn);
    }
    r = sub(r, cmA);
}

// ... and its torque Tc = (r - cmA) x F about A's centre.
template <class KA_, class KB_>
__device__ __forceinline__ void F_cc(const PhysParams& P, const Seg& A, d3 cmA, const Seg& B, KA_ kinA, KB_ kinB, d3& F,
                                     d3& Tc) {
    d3 r;
    F_cc_r(P, A, cmA, B, kinA, kinB, F, r);
    Tc = cross(r, F);
}

// F_cc accumulated into A's running sums in the reference's order.
__device__ __forceinline__ void F_cc_acc(const PhysParams& P, const Seg& A, const Kin& KA, d3 cmA, const Seg& B,
                                         const double* __restrict__ hot, int idB, d3& Fnet, d3& Tnet) {
    d3 F, Tc;
    F_cc(
        P, A, cmA, B, [&]() { return KA; },
        [&]() {
            Kin KB;
            load_kin(hot, idB, KB);
            return KB;
        },
        F, Tc);
    Fnet = add(Fnet, F);
    Tnet = add(Tnet, Tc);
}

// One locus of F_cw: ABM/Forces.cpp:98-152 / :157-201.
__device__ __forceinline__ d3 F_cw_locus(const PhysParams& P, const Seg& S, const Kin& K, d3 cm, d3 vert,
                                         double wall_z, double M, d3& rr, d3& sf, d3& df) {
    const double gamma_n = 100.0;
    double sigma = S.R;
    rr = mk(vert.x, P.planar ? 0.0 : vert.y, wall_z);  // var_pos == 0: no roughness draw term
    double d = vert.z - rr.z;
    double delta = dmax(sigma - d, 0.0);
    double FE_mag = sqrt(S.R) * P.k_wc * pow15(delta);
    if (FE_mag > DBL_EPSILON) {
        d3 FE = mk(0.0, 0.0, FE_mag);
        d3 rcm = sub(rr, cm);
        d3 va = cross(K.w, rcm);
        d3 v = add(K.v, va);
        d3 dv = scl(v, -1.0);
        d3 dvn = mk(0.0, 0.0, dv.z);
        d3 dvt = sub(dv, dvn);
        double v_mag = sqrt(dot(dvt, dvt));
        d3 FDt;
        if (v_mag > DBL_EPSILON) FDt = scl(dvt, dmin(P.gamma_t * M * sqrt(delta), P.wall_mu * FE_mag / v_mag));
        else FDt = mk(0, 0, 0);
        d3 FDn = scl(dvn, gamma_n * M * delta);
        sf = add(sf, scl(dvt, P.gamma_t * M * sqrt(delta)));
        df = add(df, scl(dvt, P.wall_mu * FE_mag / v_mag));
        return add(FE, add(FDt, FDn));
    }
    return mk(0.0, 0.0, 0.0);
}

// Array2D<double>::linear_interp (ABM/Array.h:281-288), float weights.
// Reads past the array (the reference reads there with weight 0 when Ny==1)
// contribute nothing.
__device__ __forceinline__ double height_interp(const double* __restrict__ H, int nx, int ny, int x0, int y0, double dx,
                                                double dy) {
    float x = dx < 0 ? 1 + dx : dx;
    float y = dy < 0 ? 1 + dy : dy;
    auto at = [&](int i, int j) -> double {
        if (ny == 1) {
            long f = (long)i * ny + j;
            return (f >= 0 && f < (long)nx * ny) ? H[f] : 0.0;
        }
        return (i >= 0 && i < nx && j >= 0 && j < ny) ? H[(long)i * ny + j] : 0.0;
    };
    return at(x0, y0) * (1 - x) * (1 - y) + at(x0 + 1, y0) * x * (1 - y) + at(x0, y0 + 1) * (1 - x) * y +
           at(x0 + 1, y0 + 1) * x * y;
}

// Height lookup at a point: ABM/Forces.cpp:222-233 (3-D model: same rule in y).
__device__ __forceinline__ double surf_height(const PhysParams& P, const double* __restrict__ H, d3 pt) {
    const double hx = P.BoxLength / P.ref;
    int i = (int)(pt.x / hx + P.BoxX / 2 * P.ref);
    double dx = pt.x / hx - (i - P.BoxX / 2 * P.ref + 0.5);
    i = (int)ceil(dx) + i - 1;
    int j = 0;
    double dy = 0;
    if (!P.planar) {
        j = (int)(pt.y / hx + P.BoxY / 2 * P.ref);
        dy = pt.y / hx - (j - P.BoxY / 2 * P.ref + 0.5);
        j = (int)ceil(dy) + j - 1;
    }
    return height_interp(H, P.nhx, P.nhy, i, j, dx, dy);
}

// The operands of F_surf_tension for a cell: the surface height at its end
// points and centre, and the surface normal of its height column (read
// whether or not the force applies: the column is clamped to the grid).
struct SurfOps {
    double hp, hq, hcm;
    d3 Nv;
};
__device__ __forceinline__ SurfOps surf_ops(const PhysParams& P, const Seg& S, d3 cm, const double* __restrict__ H,
                                            const double* __restrict__ N) {
    int xa = (int)((cm.x + P.BoxX / 2 * P.BoxLength) / (P.BoxLength / P.ref));
    int ya = 0;
    if (!P.planar) ya = (int)((cm.y + P.BoxY / 2 * P.BoxLength) / (P.BoxLength / P.ref));
    xa = xa < 0 ? 0 : (xa > P.nhx - 1 ? P.nhx - 1 : xa);
    ya = ya < 0 ? 0 : (ya > P.nhy - 1 ? P.nhy - 1 : ya);
    SurfOps o;
    o.hp = surf_height(P, H, S.p);
    o.hq = surf_height(P, H, S.q);
    o.hcm = surf_height(P, H, cm);
    const double* n = N + 3 * ((long)xa * P.nhy + ya);
    o.Nv = mk(n[0], n[1], n[2]);
    return o;
}

// F_surf_tension: ABM/Forces.cpp:209-317, from its operands.
__device__ __forceinline__ void F_surf_tension_ops(const PhysParams& P, const Seg& S, d3 cm, const SurfOps& o, d3& F,
                                                   d3& T) {
    const double height_p = o.hp, height_q = o.hq, height_cm = o.hcm;
    if (height_cm > (cm.z + P.BoxLength)) {
        F = mk(0, 0, 0);
        T = mk(0, 0, 0);
        return;
    }
    const d3 Nv = o.Nv;
    d3 F1 = mk(0, 0, 0), F2 = mk(0, 0, 0);
    double dh = S.p.z - (height_p - P.DH / Nv.z);
    if (!(dh < 0.0)) {
        double distance = -dh * Nv.z;
        double F0 = 2 * kPI * P.tension * dmin(distance / (P.cellRadius / 5.0), 1.0);
        F1 = scl(Nv, F0);
    }
    dh = S.q.z - (height_q - P.DH / Nv.z);
    if (!(dh < 0.0)) {
        double distance = -dh * Nv.z;
        double F0 = 2 * kPI * P.tension * dmin(distance / (P.cellRadius / 5.0), 1.0);
        F2 = scl(Nv, F0);
    }
    F = add(F1, F2);
    T = cross(sub(S.p, cm), F1);
    T = add(T, cross(sub(S.q, cm), F2));
}
__device__ __forceinline__ void F_surf_tension(const PhysParams& P, const Seg& S, d3 cm, const double* __restrict__ H,
                                               const double* __restrict__ N, d3& F, d3& T) {
    F_surf_tension_ops(P, S, cm, surf_ops(P, S, cm, H, N), F, T);
}

// Half kick shared by UpdatePositions / UpdateVelocities:
// ABM/Integrate.cpp:47-65 (identical to :106-124).
__device__ __forceinline__ void half_kick(double dt, d3 F, d3 T, const Seg& S, const Kin& K, Kin& out) {
    double R2 = S.R * S.R;
    double L2 = S.L * S.L;
    double M = kPI * R2 * (S.L + 4.0 / 3.0 * S.R);
    double Ixy = kPI * R2 * (S.L * L2 / 12.0 + S.R * (8.0 * R2 / 15.0 + L2 / 3.0));
    double Iz = kPI * R2 * R2 * (S.L / 2.0 + 8.0 * S.R / 15.0);
    d3 z = sub(S.q, S.p);
    z = scl(z, 1.0 / S.L);
    d3 vaz = scl(z, dot(z, K.w));
    d3 vaxy = sub(K.w, vaz);
    d3 Tz = scl(z, dot(z, T));
    d3 Txy = sub(T, Tz);
    out.v = add(K.v, scl(F, dt / (2.0 * M)));
    out.w = add(add(vaz, scl(Tz, dt / (2.0 * Iz))), add(vaxy, scl(Txy, dt / (2.0 * Ixy))));
}

}  // namespace bcg
