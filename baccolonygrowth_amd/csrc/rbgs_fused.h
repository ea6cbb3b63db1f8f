// rbgs_fused.h - one red-black Gauss-Seidel sweep (both colours) of the 3-D
// reaction-diffusion operator in a single launch, marching up the z planes.
//
// The two-launch sweep updates the red points from the old black values,
// then the black points from the new red values; each launch streams u, f
// and a0 through full cache lines although it updates half of the points.
// Here a workgroup owns a TX x TY column tile and the planes [k0, k1) and
// walks upwards. At step k it
//   1. issues the loads of the old iterate of plane k+3 (AHEAD = 1) or k+4
//      (AHEAD = 2) on the tile plus a 2-point ring, and of f, a0 on plane
//      k+2, into registers; with AHEAD = 2 the plane loads stay in
//      flight for two steps in two register sets, the loop unrolled by two so
//      that the sets never move;
//   2. computes the new red values of plane k+1 on the tile plus a 1-point
//      ring (2 rotating LDS slots) from the old planes k+1, k+2 in LDS (3
//      rotating slots); the ring repeats work of the neighbouring tiles so
//      that no workgroup waits for another;
//   3. computes the new black values of plane k on the tile and writes plane
//      k of the new iterate, both colours, to a second buffer;
//   4. stores plane k+3 into the LDS slot of plane k; one barrier per step.
// A workgroup starts two steps below k0 to rebuild the red planes k0-1 and
// k0. A thread owns an x-pair of columns of the red region: on every step
// exactly one column of the pair has a red point on plane k+1 and a black
// point on plane k, so every thread works on every step. The old value at
// the thread's own column and the new red value below it stay in registers.
//
// The point updates are the expressions of the two-launch sweep (k_rbgs),
// with the same boundary handling: ring positions outside the grid hold
// mirror images (Neumann faces x = 0, nx-1, y = 0, ny-1), plane nz mirrors
// plane nz-2 across the top face, and plane 0 (Dirichlet) takes f.
//
// The functions are host-and-device so that the same code can be stepped on
// the host, one phase for all threads at a time.
#pragma once

#ifndef RBGS_HD
#define RBGS_HD __host__ __device__
#endif

namespace bcg {
namespace mg {

// Output tiles: 30 x 30 (FP64: 44 KB of LDS; FP32 on the smaller levels) and
// 62 x 30 (FP32 on the finest levels: 43 KB, 1024 threads), where the wider
// tile spreads the fixed cost of a step over twice the points and trims the
// ring re-reads.
template <int TX_, int TY_>
struct RbTile {
    static constexpr int TX = TX_, TY = TY_;        // output tile
    static constexpr int RX = TX + 2, RY = TY + 2;  // red region: tile + 1-point ring
    static constexpr int UX = TX + 4, UY = TY + 4;  // old-iterate region: tile + 2-point ring
    static constexpr int THREADS = (RX / 2) * RY;   // one thread per x-pair of the red region
    static constexpr int PLANE = UY * UX;           // old-iterate values per plane
    static_assert(PLANE > 2 * THREADS && PLANE <= 3 * THREADS, "three plane loads per thread");
    static_assert(THREADS <= 1024, "one workgroup");
};
using RbNarrow = RbTile<30, 30>;
using RbWide = RbTile<62, 30>;

template <class T, class D>
struct RbShared {
    T U[3][D::UY * D::UX];  // old iterate, planes k, k+1, k+2 (slot = plane mod 3)
    T R[2][D::RY * D::RX];  // new red values, planes k, k+1 (slot = plane mod 2)
};

template <class T>
struct RbArgs {
    const T* uin;
    T* uout;
    const T* f;
    const T* a0;
    int nx, ny, nz;
    T lx, ly, lz;
};

// Per-thread state. The two columns of the pair are separate scalars chosen
// with selects (a runtime-indexed array would live in scratch memory).
// Global accesses use 32-bit byte offsets from a per-plane base pointer that
// is the same for the whole workgroup (one scalar base, one vector offset).
template <class T, class D>
struct RbThread {
    int x0, y0, k0, k1, kfirst;  // workgroup-uniform
    int par;                     // (x + y) & 1 of the pair's first column
    int ui, ri;                  // LDS index of the first column in U and R planes
    unsigned cb0, cb1;           // in-plane byte offsets of the two columns (mirrored)
    unsigned ob;                 // in-plane byte offset of the second column (output)
    unsigned ob0, ob1, ob2;      // in-plane byte offsets of this thread's share of a plane load
    bool in0, in1;               // column in the output tile (and in the grid)
    bool has2;                   // a third plane value to load
    T ucol0, ucol1;   // old value on the plane below the column's next red point
    T rcol0, rcol1;   // newest red value computed in the column
    T cf0, cf1, ca0, ca1;    // f, a0 on plane k+1 (red points now, black points next step)
    T bf, ba;                // f, a0 of this step's black point (plane k)
    T nf0, nf1, na0, na1;    // f, a0 on plane k+2, in flight
};

// This thread's share of one plane of the old iterate, in flight.
template <class T>
struct RbPlane {
    T v0, v1, v2;
};

template <class V>
RBGS_HD inline V rb_pick(int s, V v0, V v1) {
    return s ? v1 : v0;
}

template <class V>
RBGS_HD inline void rb_put(int s, V& v0, V& v1, V v) {
    if (s) v1 = v;
    else v0 = v;
}

// Index of grid point i on an axis of n points, mirrored at both faces;
// positions further out (never read by a grid point) map into the grid.
RBGS_HD inline int rb_mirror(int i, int n) {
    i = i < 0 ? -i : i;
    i = i >= n ? 2 * (n - 1) - i : i;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

RBGS_HD inline int rb_slot3(int p) { return (p + 3) % 3; }

template <class T>
RBGS_HD inline const T* rb_at(const T* base, unsigned byte_off) {
    return (const T*)((const char*)base + byte_off);
}

// column of the pair that has a red point on plane k+1 (0 or 1)
template <class T, class D>
RBGS_HD inline int rb_active(const RbThread<T, D>& t, int k) {
    return (t.par + k + 1) & 1;
}

template <class T, class D>
RBGS_HD inline void rb_init(const RbArgs<T>& a, int bx, int by, int bz, int kchunk, int tid, RbThread<T, D>& t) {
    t.x0 = bx * D::TX;
    t.y0 = by * D::TY;
    t.k0 = bz * kchunk;
    t.k1 = t.k0 + kchunk < a.nz ? t.k0 + kchunk : a.nz;
    t.kfirst = t.k0 - 2 > -1 ? t.k0 - 2 : -1;
    const int px = tid % (D::RX / 2), ry = tid / (D::RX / 2);
    const int y = t.y0 - 1 + ry;
    const int xa = t.x0 - 1 + 2 * px;
    t.par = (xa + y) & 1;
    t.ui = (ry + 1) * D::UX + 2 * px + 1;
    t.ri = ry * D::RX + 2 * px;
    const int gy = rb_mirror(y, a.ny);
    t.cb0 = (unsigned)((gy * a.nx + rb_mirror(xa, a.nx)) * (int)sizeof(T));
    t.cb1 = (unsigned)((gy * a.nx + rb_mirror(xa + 1, a.nx)) * (int)sizeof(T));
    const bool yin = y >= t.y0 && y < t.y0 + D::TY && y < a.ny;
    t.in0 = yin && xa >= t.x0 && xa < t.x0 + D::TX && xa < a.nx;
    t.in1 = yin && xa + 1 >= t.x0 && xa + 1 < t.x0 + D::TX && xa + 1 < a.nx;
    t.ob = (t.in0 || t.in1) ? (unsigned)((y * a.nx + xa + 1) * (int)sizeof(T)) : 0u;
    auto src = [&](int e) {
        const int r = e / D::UX;
        const int cx = e - r * D::UX;
        return (unsigned)((rb_mirror(t.y0 - 2 + r, a.ny) * a.nx + rb_mirror(t.x0 - 2 + cx, a.nx)) * (int)sizeof(T));
    };
    t.ob0 = src(tid);
    t.ob1 = src(tid + D::THREADS);
    t.has2 = tid + 2 * D::THREADS < D::PLANE;
    t.ob2 = t.has2 ? src(tid + 2 * D::THREADS) : 0u;
    t.ucol0 = t.ucol1 = t.rcol0 = t.rcol1 = T(0);
    t.cf0 = t.cf1 = t.ca0 = t.ca1 = t.bf = t.ba = T(0);
    t.nf0 = t.nf1 = t.na0 = t.na1 = T(0);
}

// This thread's share of the old iterate of plane p (0 <= p <= nz; plane
// nz is the mirror of nz-2) on the tile + 2-point ring, into registers. A
// null uin is the zero iterate (a sweep from a zero initial guess).
template <class T, class D>
RBGS_HD inline void rb_fetch_plane(const RbArgs<T>& a, const RbThread<T, D>& t, int p, RbPlane<T>& r) {
    if (!a.uin) {
        r.v0 = r.v1 = r.v2 = T(0);
        return;
    }
    const int pp = p == a.nz ? a.nz - 2 : p;
    const T* src = a.uin + (long)pp * a.nx * a.ny;
    r.v0 = *rb_at(src, t.ob0);
    r.v1 = *rb_at(src, t.ob1);
    r.v2 = t.has2 ? *rb_at(src, t.ob2) : T(0);
}

// ... and from the registers into the LDS slot of plane p.
template <class T, class D>
RBGS_HD inline void rb_store_plane(RbShared<T, D>& sh, const RbThread<T, D>& t, int p, int tid, const RbPlane<T>& r) {
    T* dst = sh.U[rb_slot3(p)];
    dst[tid] = r.v0;
    dst[tid + D::THREADS] = r.v1;
    if (t.has2) dst[tid + 2 * D::THREADS] = r.v2;
}

// f, a0 on plane p (0 <= p <= nz-1) at both columns of the pair, into the
// in-flight registers. Each plane's values are loaded once: the active
// column's on one step (red point), the other column's on the next (black).
template <class T, class D>
RBGS_HD inline void rb_fetch_coef(const RbArgs<T>& a, RbThread<T, D>& t, int p) {
    const T* fp = a.f + (long)p * a.nx * a.ny;
    const T* ap = a.a0 + (long)p * a.nx * a.ny;
    t.nf0 = *rb_at(fp, t.cb0);
    t.nf1 = *rb_at(fp, t.cb1);
    t.na0 = *rb_at(ap, t.cb0);
    t.na1 = *rb_at(ap, t.cb1);
}

// End of step k: the black point of step k+1 is the other column on plane
// k+1; the in-flight plane k+2 becomes the current one.
template <class T, class D>
RBGS_HD inline void rb_take_coef(RbThread<T, D>& t, int k) {
    const int s = rb_active(t, k);
    t.bf = rb_pick(s, t.cf1, t.cf0);
    t.ba = rb_pick(s, t.ca1, t.ca0);
    t.cf0 = t.nf0;
    t.cf1 = t.nf1;
    t.ca0 = t.na0;
    t.ca1 = t.na1;
}

// Plane p feeds the red points of plane p-1, computed on step p-2.
template <class T, class D>
RBGS_HD inline bool rb_plane_needed(const RbArgs<T>& a, const RbThread<T, D>& t, int p) {
    return p >= 0 && p <= a.nz && p - 2 < t.k1;
}

// Before the first step (then a barrier): old planes kfirst .. kfirst+2 (those
// that exist) into LDS, f and a0 of the first step.
template <class T, class D>
RBGS_HD inline void rb_prologue(const RbArgs<T>& a, RbShared<T, D>& sh, RbThread<T, D>& t, int tid) {
    for (int p = t.kfirst; p <= t.kfirst + 2; p++)
        if (rb_plane_needed(a, t, p)) {
            RbPlane<T> r;
            rb_fetch_plane(a, t, p, r);
            rb_store_plane(sh, t, p, tid, r);
        }
    rb_fetch_coef(a, t, t.kfirst + 1);  // kfirst + 1 <= k0 - 1 < nz
    rb_take_coef(t, t.kfirst - 1);      // (no black point on the first step)
}

// After the prologue is visible (then a barrier, as its slots are reused):
// the old values below each column's first red point.
template <class T, class D>
RBGS_HD inline void rb_prologue_init(const RbShared<T, D>& sh, RbThread<T, D>& t) {
    const int s0 = rb_active(t, t.kfirst);  // active at kfirst; the other column at kfirst + 1
    if (t.kfirst >= 0) rb_put(s0, t.ucol0, t.ucol1, sh.U[rb_slot3(t.kfirst)][t.ui + s0]);
    if (t.kfirst + 1 >= 0) rb_put(1 - s0, t.ucol0, t.ucol1, sh.U[rb_slot3(t.kfirst + 1)][t.ui + 1 - s0]);
}

// Step k, phase 1: issue the loads of f, a0 on plane k+2 (red points of
// step k+1) and of the old iterate on plane k+2+ahead (into `later`; stored
// on step k+ahead-1).
template <class T, class D>
RBGS_HD inline void rb_fetch_next(const RbArgs<T>& a, RbThread<T, D>& t, int k, int ahead, RbPlane<T>& later) {
    if (k + 1 < t.k1 && k + 2 <= a.nz - 1) rb_fetch_coef(a, t, k + 2);
    if (rb_plane_needed(a, t, k + 2 + ahead)) rb_fetch_plane(a, t, k + 2 + ahead, later);
}

// Step k, phase 3 (then a barrier): plane k+3 (`now`, in flight since step
// k or k-1) into the slot of plane k.
template <class T, class D>
RBGS_HD inline void rb_finish(const RbArgs<T>& a, RbShared<T, D>& sh, RbThread<T, D>& t, int k, int tid,
                              const RbPlane<T>& now) {
    if (rb_plane_needed(a, t, k + 3)) rb_store_plane(sh, t, k + 3, tid, now);
    rb_take_coef(t, k);
}

// Step k, phase 2: red on plane k+1, black on plane k, write plane k of the
// new iterate.
template <class T, class D>
RBGS_HD inline void rb_compute(const RbArgs<T>& a, RbShared<T, D>& sh, RbThread<T, D>& t, int k) {
    const int s = rb_active(t, k);
    const T d2 = T(2) * (a.lx + a.ly + a.lz);
    const bool do_red = k + 1 <= a.nz - 1;
    T red = T(0);
    if (do_red) {
        const int i = t.ui + s;
        const T above = sh.U[rb_slot3(k + 2)][i];
        const T fr = rb_pick(s, t.cf0, t.cf1);
        if (k + 1 == 0) {
            red = fr;
        } else {
            const T* Up = sh.U[rb_slot3(k + 1)];
            const T sum = a.lx * (Up[i - 1] + Up[i + 1]) + a.ly * (Up[i - D::UX] + Up[i + D::UX]) +
                          a.lz * (rb_pick(s, t.ucol0, t.ucol1) + above);
            red = (fr + sum) / (rb_pick(s, t.ca0, t.ca1) + d2);
        }
        rb_put(s, t.ucol0, t.ucol1, above);
        sh.R[(k + 1) & 1][t.ri + s] = red;
    }
    if (k >= t.k0) {
        T* out = a.uout + (long)k * a.nx * a.ny;
        const T rown = rb_pick(s, t.rcol0, t.rcol1);  // new red value below (plane k-1)
        if (rb_pick(s, t.in0, t.in1)) {
            T blk;
            if (k == 0) {
                blk = t.bf;
            } else {
                const T* Rk = sh.R[k & 1];
                const int i = t.ri + s;
                const T above = k == a.nz - 1 ? rown : red;
                const T sum = a.lx * (Rk[i - 1] + Rk[i + 1]) + a.ly * (Rk[i - D::RX] + Rk[i + D::RX]) +
                              a.lz * (rown + above);
                blk = (t.bf + sum) / (t.ba + d2);
            }
            *(T*)rb_at(out, t.ob - (unsigned)((1 - s) * (int)sizeof(T))) = blk;
        }
        // the other column's red point on plane k, computed on the previous step
        if (rb_pick(s, t.in1, t.in0)) *(T*)rb_at(out, t.ob - (unsigned)(s * (int)sizeof(T))) = rb_pick(s, t.rcol1, t.rcol0);
    }
    if (do_red) rb_put(s, t.rcol0, t.rcol1, red);
}

// Workgroups of one sweep: tiles_x * tiles_y tiles times nchunk z ranges of
// kchunk planes.
struct RbGrid {
    int tiles_x, tiles_y, nchunk, kchunk;
};

// Number of z ranges of a z-marching launch over `tiles` column tiles and nz
// planes: the count that minimises rounds x (planes per range + halo), where
// a round is `resident` workgroups (all that fit on the device at once) and
// `halo` the extra planes a workgroup steps through below its range; at
// least `minplanes` planes per range. A launch of 3.2 rounds takes as long
// as one of 4, so the count is chosen to fill the last round.
inline int z_ranges(long tiles, int nz, long resident, int halo, int minplanes) {
    const int maxc = nz / minplanes > 1 ? nz / minplanes : 1;
    resident = resident > 0 ? resident : 1;
    int best = 1;
    long best_cost = -1;
    for (int nc = 1; nc <= maxc; nc++) {
        const int kchunk = (nz + nc - 1) / nc;
        const int used = (nz + kchunk - 1) / kchunk;
        if (used != nc) continue;  // the same ranges as a smaller count
        const long rounds = (tiles * nc + resident - 1) / resident;
        const long cost = rounds * (kchunk + halo);
        if (best_cost < 0 || cost < best_cost) {
            best = nc;
            best_cost = cost;
        }
    }
    return best;
}

template <class D>
inline RbGrid rb_grid(int nx, int ny, int nz, long resident) {
    RbGrid g;
    g.tiles_x = (nx + D::TX - 1) / D::TX;
    g.tiles_y = (ny + D::TY - 1) / D::TY;
    const int nc = z_ranges((long)g.tiles_x * g.tiles_y, nz, resident, 3, 4);
    g.kchunk = (nz + nc - 1) / nc;
    g.nchunk = (nz + g.kchunk - 1) / g.kchunk;
    return g;
}

}  // namespace mg
}  // namespace bcg
