// zmarch.h - 7-point-stencil passes over a level that march up the z planes.
//
// A workgroup owns a ZM_TX x ZM_TY column tile and the planes [k0, k1). The
// stencil operand of plane k (tile plus a 1-point ring, mirrored at the
// Neumann faces) sits in LDS; the operand below and above the thread's own
// point stays in registers, so every plane of the operand is read from
// memory once (plus the ring) instead of three times. At step k a thread
//   1. issues the loads of plane k+1 (its own point, and one ring point for
//      the first threads) and of its point's other inputs on plane k+1 (two
//      register sets, the loop unrolled by two, so that they stay in flight
//      for a whole step);
//   2. applies the pass at (i, j, k) from the LDS plane and the registers;
//   3. stores plane k+1 into the other LDS slot; one barrier per step.
// The passes (an Op) decide what the operand is (e.g. x + e/scale, formed
// while loading), what else a point reads and what it writes; per-workgroup
// sums of squares are reduced in a fixed order.
//
// The functions are host-and-device so that the same code can be stepped on
// the host, one phase for all threads at a time.
#pragma once

#include "rbgs_fused.h"  // z_ranges, RBGS_HD

#ifndef RBGS_HD
#define RBGS_HD __host__ __device__
#endif

namespace bcg {
namespace mg {

constexpr int ZM_TX = 64, ZM_TY = 8;
constexpr int ZM_THREADS = ZM_TX * ZM_TY;
constexpr int ZM_PX = ZM_TX + 2, ZM_PY = ZM_TY + 2;  // tile + 1-point ring
constexpr int ZM_PLANE = ZM_PX * ZM_PY;
constexpr int ZM_RING = ZM_PLANE - ZM_THREADS;
static_assert(ZM_RING <= ZM_THREADS, "one ring point per thread at most");

template <class T>
struct ZmShared {
    T U[2][ZM_PLANE];
};

template <class T>
struct ZmThread {
    int x0, y0, k0, k1;
    bool act;          // own point inside the grid
    int li;            // LDS index of the own point
    int lr;            // LDS index of the ring point this thread loads (-1: none)
    unsigned pb;       // in-plane byte offset of the own point (mirrored into the grid when !act)
    unsigned rb;       // in-plane byte offset of the ring point (mirrored)
    T below, center;   // operand at the own column, planes k-1 and k
    T n0, n1;          // operand on plane k+1 (own point, ring point), in flight
};

RBGS_HD inline int zm_mirror(int i, int n) {
    i = i < 0 ? -i : i;
    i = i >= n ? 2 * (n - 1) - i : i;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// LDS index of ring point q (top row, bottom row, left column, right column).
RBGS_HD inline int zm_ring_index(int q) {
    if (q < ZM_PX) return q;
    if (q < 2 * ZM_PX) return (ZM_PY - 1) * ZM_PX + (q - ZM_PX);
    if (q < 2 * ZM_PX + ZM_TY) return (1 + q - 2 * ZM_PX) * ZM_PX;
    return (1 + q - 2 * ZM_PX - ZM_TY) * ZM_PX + ZM_PX - 1;
}

template <class T>
RBGS_HD inline void zm_init(int nx, int ny, int nz, int bx, int by, int bz, int kchunk, int tid, ZmThread<T>& t) {
    t.x0 = bx * ZM_TX;
    t.y0 = by * ZM_TY;
    t.k0 = bz * kchunk;
    t.k1 = t.k0 + kchunk < nz ? t.k0 + kchunk : nz;
    const int tx = tid % ZM_TX, ty = tid / ZM_TX;
    const int i = t.x0 + tx, j = t.y0 + ty;
    t.act = i < nx && j < ny;
    t.li = (ty + 1) * ZM_PX + tx + 1;
    t.pb = (unsigned)((zm_mirror(j, ny) * nx + zm_mirror(i, nx)) * (int)sizeof(T));
    if (tid < ZM_RING) {
        t.lr = zm_ring_index(tid);
        const int r = t.lr / ZM_PX, c = t.lr - r * ZM_PX;
        t.rb = (unsigned)((zm_mirror(t.y0 - 1 + r, ny) * nx + zm_mirror(t.x0 - 1 + c, nx)) * (int)sizeof(T));
    } else {
        t.lr = -1;
        t.rb = 0u;
    }
    t.below = t.center = t.n0 = t.n1 = T(0);
}

// Before the first step (then a barrier): plane k0 into LDS slot k0 & 1, the
// own column's operand on planes k0-1 and k0, the other inputs on plane k0.
template <class T, class Op>
RBGS_HD inline void zm_prologue(const Op& op, ZmShared<T>& sh, ZmThread<T>& t, typename Op::Pre& pre) {
    T* U = sh.U[t.k0 & 1];
    t.center = op.u(t.k0, t.pb);
    U[t.li] = t.center;
    if (t.lr >= 0) U[t.lr] = op.u(t.k0, t.rb);
    t.below = t.k0 >= 1 ? op.u(t.k0 - 1, t.pb) : T(0);
    op.fetch(t.k0, t.pb, pre);
}

// Step k, phase 1: loads of plane k+1 (when there is one) and of the point's
// other inputs on plane k+1 (when there is a step k+1).
template <class T, class Op>
RBGS_HD inline void zm_fetch(const Op& op, ZmThread<T>& t, typename Op::Pre& next, int nz, int k) {
    if (k + 1 <= nz - 1) {
        t.n0 = op.u(k + 1, t.pb);
        if (t.lr >= 0) t.n1 = op.u(k + 1, t.rb);
    }
    if (k + 1 < t.k1) op.fetch(k + 1, t.pb, next);
}

// Step k, phase 2: the pass at the own point; returns its sum-of-squares term.
template <class T, class Op>
RBGS_HD inline double zm_point(const Op& op, const ZmShared<T>& sh, const ZmThread<T>& t, const typename Op::Pre& pre,
                               int nz, int k) {
    if (!t.act) return 0.0;
    const T* U = sh.U[k & 1];
    const int i = t.li;
    const T above = k + 1 <= nz - 1 ? t.n0 : t.below;  // mirror across the top face
    const T s = op.lx * (U[i - 1] + U[i + 1]) + op.ly * (U[i - ZM_PX] + U[i + ZM_PX]) + op.lz * (t.below + above);
    return op.point(k, t.pb, t.center, s, pre);
}

// Step k, phase 3 (then a barrier): plane k+1 into the other LDS slot.
template <class T>
RBGS_HD inline void zm_finish(ZmShared<T>& sh, ZmThread<T>& t, int nz, int k) {
    if (k + 1 <= nz - 1) {
        T* U = sh.U[(k + 1) & 1];
        U[t.li] = t.n0;
        if (t.lr >= 0) U[t.lr] = t.n1;
    }
    t.below = t.center;
    t.center = t.n0;
}

// Workgroups: tiles_x * tiles_y tiles times nchunk z ranges of kchunk planes.
struct ZmGrid {
    int tiles_x, tiles_y, nchunk, kchunk;
    long blocks() const { return (long)tiles_x * tiles_y * nchunk; }
};

inline ZmGrid zm_grid(int nx, int ny, int nz, long resident) {
    ZmGrid g;
    g.tiles_x = (nx + ZM_TX - 1) / ZM_TX;
    g.tiles_y = (ny + ZM_TY - 1) / ZM_TY;
    const int nc = z_ranges((long)g.tiles_x * g.tiles_y, nz, resident, 1, 4);
    g.kchunk = (nz + nc - 1) / nc;
    g.nchunk = (nz + g.kchunk - 1) / g.kchunk;
    return g;
}

template <class T>
RBGS_HD inline const T* zm_at(const T* base, unsigned byte_off) {
    return (const T*)((const char*)base + byte_off);
}

// ----------------------------------------------------------------- passes
// Plane p of an array at in-plane byte offset b.
template <class T>
RBGS_HD inline T zm_ld(const T* a, long sxy, int p, unsigned b) {
    return *zm_at(a + p * sxy, b);
}

// r = f - A u  (Dirichlet rows: r = f - u). Same arithmetic as k_residual.
template <class T>
struct ZmResidual {
    const T* uin;
    const T* f;
    const T* a0;
    T* r;
    long sxy;
    T lx, ly, lz;
    struct Pre {
        T f, a0;
    };
    RBGS_HD T u(int p, unsigned b) const { return zm_ld(uin, sxy, p, b); }
    RBGS_HD void fetch(int k, unsigned b, Pre& q) const {
        q.f = zm_ld(f, sxy, k, b);
        q.a0 = k >= 1 ? zm_ld(a0, sxy, k, b) : T(0);
    }
    RBGS_HD double point(int k, unsigned b, T c, T s, const Pre& q) const {
        T v;
        if (k == 0) {
            v = q.f - c;
        } else {
            const T d = q.a0 + T(2) * (lx + ly + lz);
            v = q.f - (d * c - s);
        }
        *(T*)zm_at(r + k * sxy, b) = v;
        return 0.0;
    }
};

// ||b - A u||^2 terms, FP64, nothing written (the FP64 solver's convergence
// check), with the arithmetic of ZmResidual.
struct ZmResNorm {
    const double* uin;
    const double* bv;
    const double* a0;
    long sxy;
    double lx, ly, lz;
    struct Pre {
        double b, a0;
    };
    RBGS_HD double u(int p, unsigned b) const { return zm_ld(uin, sxy, p, b); }
    RBGS_HD void fetch(int k, unsigned b, Pre& q) const {
        q.b = zm_ld(bv, sxy, k, b);
        q.a0 = k >= 1 ? zm_ld(a0, sxy, k, b) : 0.0;
    }
    RBGS_HD double point(int k, unsigned, double c, double s, const Pre& q) const {
        const double r = k == 0 ? q.b - c : q.b - ((q.a0 + 2.0 * (lx + ly + lz)) * c - s);
        return r * r;
    }
};

// Mixed-precision outer step: operand x = x_old + e / scale (e == null: x_old),
// written to x_new (x_new == null: not written); r = b - A x in FP64,
// r32 = r * rscale, sum of r^2. Forming
// x while loading the operand fuses the update into the residual pass; x_old
// and x_new are different buffers, so no workgroup reads a value another
// one writes.
struct ZmIrUpdate {
    const double* xo;
    double* xn;
    const float* e;
    double inv_scale;
    const double* bv;
    const double* a0;
    float* r32;
    double rscale;
    long sxy;
    double lx, ly, lz;
    struct Pre {
        double b, a0;
    };
    RBGS_HD double u(int p, unsigned b) const {
        const double x = zm_ld(xo, sxy, p, b);
        return e ? x + (double)*zm_at(e + p * sxy, b / 2) * inv_scale : x;
    }
    RBGS_HD void fetch(int k, unsigned b, Pre& q) const {
        q.b = zm_ld(bv, sxy, k, b);
        q.a0 = k >= 1 ? zm_ld(a0, sxy, k, b) : 0.0;
    }
    RBGS_HD double point(int k, unsigned b, double c, double s, const Pre& q) const {
        if (xn) *(double*)zm_at(xn + k * sxy, b) = c;
        const double r = k == 0 ? q.b - c : q.b - ((q.a0 + 2.0 * (lx + ly + lz)) * c - s);
        *(float*)zm_at(r32 + k * sxy, b / 2) = (float)(r * rscale);
        return r * r;
    }
};

}  // namespace mg
}  // namespace bcg
