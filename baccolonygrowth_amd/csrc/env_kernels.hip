// env_kernels.hip - device side of the planar colony environment (gfx950).
//
// Field passes (one launch each, one thread per column or box):
//   density   GetDensity            ABM/Compute.cpp:120-253
//   height    GetHeight             ABM/Compute.cpp:256-341
//   shift     ShiftDensity          ABM/Compute.cpp:344-463
//   average   BoxAverage            ABM/Compute.cpp:495-605
//   height_average, smooth, normal  ABM/Compute.cpp:640-738, 470-493
// Nutrient update, one workgroup for the whole of
//   UpdateEnvArrayMG_OxygenAcetate  ABM/Multigrid.cpp:984-1125
// with barrier-separated phases in the order of the single-threaded
// reference: the nested agar multigrid (Multigrid::solve, :864-889), then
// 200 Jacobi sweeps over the colony and the colony-agar interface per outer
// iteration. Every update is a pure function of the previous phase's arrays,
// so the points of a phase run in parallel; the multigrid relaxation keeps
// the reference's order where it matters (colours by x parity, z ascending
// within a column, one thread per column).
//
// Arithmetic follows the reference expression by expression (operand order
// and grouping), compiled without contraction, so that results are the
// reference's to the last bit except where a libm function differs (log in
// the pH output fields).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "env_kernels.h"

namespace bcg {
namespace env {

// std::max / std::min as the reference's <algorithm> calls evaluate them.
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ void flag(int* e, int bit) { atomicOr(e, bit); }

// ------------------------------------------------------------ field passes
__global__ void k_env_init_bounds(FieldArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.bounds[0] = a.BX;  // minx
        a.bounds[1] = 0;     // maxx
        a.bounds[2] = 1;     // miny (BoxY)
        a.bounds[3] = 0;     // maxy
        a.bounds[4] = 0;     // maxz
    }
}

__device__ __forceinline__ void note_box(const FieldArgs& a, int x, int z) {
    atomicMin(&a.bounds[0], x);
    atomicMax(&a.bounds[1], x);
    atomicMin(&a.bounds[2], 0);
    atomicMax(&a.bounds[3], 0);
    atomicMax(&a.bounds[4], z);
}

// GetDensity: per column, boxes bottom to top; per box the cells in bin
// order. Density = sum(L R)/BoxLength^2, type fractions, footprint 1/0,
// per-sub-column maximum top height (starting at cellRadius).
__global__ void k_env_density(FieldArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.BX) return;
    double hgt[MAX_REF];
    for (int r = 0; r < a.REF; r++) hgt[r] = a.cellRadius;
    const double Vbox = a.BoxLength * a.BoxLength;
    for (int z = 0; z < a.BZ; z++) {
        const int b = x * a.BZ + z;
        const int s = a.start[b], e = a.start[b + 1];
        if (e > s) note_box(a, x, z);
        double dens = 0.0, d1 = 0.0, d2 = 0.0;
        for (int c = s; c < e; c++) {
            const int id = a.sorted[c];
            const double* h = a.hot + (size_t)id * 16;
            const double cmx = (h[0] + h[3]) * 0.5;
            int i = (int)((cmx / a.BoxLength - (double)(x - a.BX / 2)) * (double)a.REF);
            i = i < 0 ? 0 : (i > a.REF - 1 ? a.REF - 1 : i);
            const double top = smax(h[2], h[5]);
            hgt[i] = smax(hgt[i], top);
            const double LR = h[12] * h[13];
            const int t = a.type[id];
            if (t == 1) d1 += LR;
            else if (t == 2) d2 += LR;
            else flag(a.errflag, ERR_CELLTYPE);
            dens += LR;
        }
        if (dens == 0) {
            a.rough1[b] = 0;
            a.rough2[b] = 0;
            a.rough[b] = 0;
            a.inside[b] = 0;
        } else {
            a.rough1[b] = d1 / dens;
            a.rough2[b] = d2 / dens;
            a.rough[b] = dens / Vbox;
            a.inside[b] = 1.0;
        }
    }
    for (int r = 0; r < a.REF; r++) a.height[(x * a.REF + r) * a.NHY] = hgt[r];
}

// GetHeight: per column from the top down, stopping after the fourth
// occupied box; only the boxes visited enter the bounding box.
__global__ void k_env_height(FieldArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.BX) return;
    double hgt[MAX_REF];
    for (int r = 0; r < a.REF; r++) hgt[r] = a.cellRadius;
    int occupied = 0;
    for (int z = a.BZ - 1; z >= 0; z--) {
        const int b = x * a.BZ + z;
        const int s = a.start[b], e = a.start[b + 1];
        if (e > s) {
            note_box(a, x, z);
            occupied++;
        }
        for (int c = s; c < e; c++) {
            const int id = a.sorted[c];
            const double* h = a.hot + (size_t)id * 16;
            const double cmx = (h[0] + h[3]) * 0.5;
            int i = (int)((cmx - (double)(x - a.BX / 2) * a.BoxLength) / (a.BoxLength / (double)a.REF));
            i = i < 0 ? 0 : (i > a.REF - 1 ? a.REF - 1 : i);
            hgt[i] = smax(hgt[i], smax(h[2], h[5]));
        }
        if (occupied > 3) break;
    }
    for (int r = 0; r < a.REF; r++) a.height[(x * a.REF + r) * a.NHY] = hgt[r];
}

// Halo of the footprint: empty boxes next to an occupied one (x or z) get
// 0.4; at z = 0 the box below does not exist. 0.4 < 0.5 never qualifies a
// neighbour, so the in-place pass equals the reference's serial one.
__global__ void k_env_halo(FieldArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.BX * a.BZ) return;
    const int x = b / a.BZ, z = b - x * a.BZ;
    const double* I = a.inside;
    if (!(I[b] < 0.5)) return;
    if (z > 0 && x > 0 && x < a.BX - 1 && z < a.BZ - 1) {
        if (I[b + a.BZ] > 0.5 || I[b - a.BZ] > 0.5 || I[b + 1] > 0.5 || I[b - 1] > 0.5) a.inside[b] = 0.4;
    } else if (z == 0 && x > 0 && x < a.BX - 1) {
        if (I[b + a.BZ] > 0.5 || I[b - a.BZ] > 0.5 || I[b + 1] > 0.5) a.inside[b] = 0.4;
    }
}

__global__ void k_env_pad_bounds(FieldArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.bounds[0] -= 5;
        a.bounds[1] += 5;
        a.bounds[4] += 5;
    }
}

// ShiftDensity with BoxY = 1: the reference's eight colony loops and four
// wall loops overwrite each other (its y = 0 and y = BoxY-1 rows coincide);
// each box gets the value of the last loop that writes it, including the
// wall loops' reuse of the type-2 temporary of the first wall loop.
__global__ void k_env_shift(FieldArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.BX * a.BZ) return;
    const int BX1 = a.BX - 1, BZ1 = a.BZ - 1;
    const int x = b / a.BZ, z = b - x * a.BZ;
    const double* Ds[3] = {a.rough, a.rough1, a.rough2};
    double* Out[3] = {a.rshift, a.rshift1, a.rshift2};
    for (int s = 0; s < 3; s++) {
        const double* D = Ds[s];
        auto d = [&](int xx, int zz) { return D[xx * a.BZ + zz]; };
        double v;
        if (x < BX1 && z < BZ1) v = (d(x, z) + d(x, z + 1) + d(x + 1, z) + d(x + 1, z + 1)) / 4.0;
        else if (x < BX1) v = (d(x, BZ1) + d(x + 1, BZ1)) / 2.0;
        else if (z < BZ1) v = (d(BX1, z) + d(BX1, z)) / 2.0;
        else v = d(BX1, BZ1);
        Out[s][b] = v;
    }
    if (z == 0) {
        auto d = [&](const double* D, int xx) { return D[xx * a.BZ]; };
        if (x < BX1) {
            a.rwshift[x] = (d(a.rough, x) + d(a.rough, x + 1)) / 2.0;
            a.rwshift1[x] = (d(a.rough1, x) + d(a.rough1, x + 1)) / 2.0;
            const int xl = a.BX - 2;  // last iteration of the first wall loop
            a.rwshift2[x] = (d(a.rough2, xl) + d(a.rough2, xl) + d(a.rough2, xl + 1) + d(a.rough2, xl + 1)) / 4.0;
        } else {
            a.rwshift[x] = d(a.rough, BX1);
            a.rwshift1[x] = d(a.rough1, BX1);
            a.rwshift2[x] = d(a.rough2, BX1);
        }
    }
}

// BoxAverage: the wall row under the colony boxes, then a FilterLen^2 x-z
// Gaussian with the reflection rule below the wall, divided by the density
// threshold and capped at 1, on x in [minx, maxx), z in [0, maxz]. Boxes
// outside that window are 0 (the reference leaves them uninitialised).
__global__ void k_env_box_average(FieldArgs a, const double* rough, const double* rwall, double* out,
                                  double* outwall, const double* filter3d) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int NZ = a.BZ + 1;  // wall row + colony boxes
    if (b >= a.BX * NZ) return;
    const int x = b / NZ, k = b - x * NZ;
    const int minx = a.bounds[0], maxx = a.bounds[1], maxz = a.bounds[4];
    double res = 0.0;
    if (x >= minx && x < maxx && k <= maxz) {
        const int FD = a.FD, FL = 2 * a.FD + 1;
        auto in = [&](int i, int kk) -> double {
            if (i < 0 || i >= a.BX || kk < 0 || kk > a.BZ) {
                flag(a.errflag, ERR_RANGE);
                return 0.0;
            }
            return kk == 0 ? rwall[i] : rough[i * a.BZ + kk - 1];
        };
        double avg = 0.0;
        for (int iF = -FD; iF <= FD; iF++)
            for (int kF = -FD; kF <= FD; kF++) {
                const int i = x + iF, kk = k + kF;
                const double F = filter3d[(iF + FD) * FL + (kF + FD)];
                if (kk < 0) {
                    const double v0 = in(i, 0);
                    if (v0 > 0.00001) avg += F * 1.0;
                    else avg += F * v0;
                } else {
                    avg += F * in(i, kk);
                }
            }
        const double v = avg / a.density_threshold;
        res = smin(1.0, v);
    }
    if (k == 0) outwall[x] = res;
    else out[x * a.BZ + k - 1] = res;
}

// Height_Average: each sub-column averaged with its lower neighbours
// (x-1 first, then x+1), over [minx*ref, (maxx+1)*ref).
__global__ void k_env_height_average(FieldArgs a) {
    const int i = a.bounds[0] * a.REF + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (a.bounds[1] + 1) * a.REF) return;
    const int NHX = a.BX * a.REF;
    if (i - 1 < 0 || i + 1 >= NHX) {
        flag(a.errflag, ERR_RANGE);
        return;
    }
    const double H0 = a.height[i * a.NHY];
    double Ht = H0;
    int count = 1;
    for (int n = -1; n <= 1; n += 2) {
        const double H = a.height[(i + n) * a.NHY];
        if (H < H0) {
            Ht += H;
            count++;
        }
    }
    a.rheight[i * a.NHY] = Ht / count;
}

// Smooth: box filter of RoughHeight into Height over the same window.
__global__ void k_env_smooth(FieldArgs a, const double* filter1d) {
    const int i = a.bounds[0] * a.REF + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (a.bounds[1] + 1) * a.REF) return;
    const int NHX = a.BX * a.REF;
    if (i - a.FD < 0 || i + a.FD >= NHX) {
        flag(a.errflag, ERR_RANGE);
        return;
    }
    double avg = 0.0;
    for (int f = -a.FD; f <= a.FD; f++) avg += filter1d[f + a.FD] * a.rheight[(i + f) * a.NHY];
    a.height[i * a.NHY] = avg;
}

// GetSurfaceNormal: central difference of the height, normal (dx, 0, -1)/|.|.
__global__ void k_env_normal(FieldArgs a) {
    const int i = a.bounds[0] * a.REF + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (a.bounds[1] + 1) * a.REF) return;
    const int NHX = a.BX * a.REF;
    if (i - 1 < 0 || i + 1 >= NHX) {
        flag(a.errflag, ERR_RANGE);
        return;
    }
    const double d_x = (a.height[(i + 1) * a.NHY] - a.height[(i - 1) * a.NHY]) / (2.0 * a.BoxLength / a.REF);
    const double d_z = -1.0;
    const double l_n = sqrt(d_x * d_x + d_z * d_z);
    double* N = a.normal + (size_t)i * a.NHY * 3;
    N[0] = d_x / l_n;
    N[1] = 0;
    N[2] = d_z / l_n;
}

__global__ void k_env_growth_to_mech(const bcg_env_cell* colony, int n, double* gr1, double* gr2) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    gr1[b] = colony[b].GrowthRate1;
    gr2[b] = colony[b].GrowthRate2;
}

// ======================================================== nutrient update
namespace {

constexpr int COARSE_MAX = 64;  // points of the coarsest level relaxed in LDS
// Loads in flight per column walk of the relaxation: the one-workgroup
// passes (many walks a lane, L2-resident levels) and the device-wide passes
// on the wide levels (one walk a lane, deep columns that stream from HBM).
constexpr int RING_WG = 6;
#ifndef BCG_RING_WIDE
#define BCG_RING_WIDE 16
#endif
constexpr int RING_WIDE = BCG_RING_WIDE;

// Orders one wave's LDS writes before its later LDS reads (the lanes of a
// wave run in lockstep; the fences keep the compiler from moving accesses).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Global-memory doubles: data pointers of EnvArgs read from LDS are plain
// (flat) pointers; cast to this type, their loads and stores are global ones.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gdouble* G(double* p) { return (gdouble*)p; }
__device__ __forceinline__ const gdouble* G(const double* p) { return (const gdouble*)p; }

struct Geo {  // one multigrid level as the loops of Multigrid.cpp see it
    int nx, nz, nxp, nzp, d;
    double h2;  // BoxLength^2 d^2
};

// The three multigrid species of one point (the multigrid never touches
// WasteAgar). Each level stores them as three blocks C, O2, acetate, each
// [z][x] with x fastest, so that the lanes of a wave, which take
// neighbouring x, read neighbouring words.
struct AgaRef {
    gdouble& CarbonAgar;
    gdouble& OxygenAgar;
    gdouble& AcetateAgar;
};

// Multigrid arrays: (x extent, z extent) per level from the constructor
// (ABM/Multigrid.cpp:5-33).
struct MG {
    const EnvArgs& a;
    int AL;  // agarLevels = maxLevelsMG
    double dt;
    gdouble* phantom;  // this thread's 3 values for reads past the end of an array
    __device__ MG(const EnvArgs& args, double dt_, gdouble* ph) : a(args), AL(args.ML), dt(dt_), phantom(ph) {}

    __device__ int xext(int l) const { return a.mgx[l]; }  // x extent - 1
    __device__ Geo geo(int l, int sub) const {  // sub: loop extents BoxX*d0/d/sub
        Geo g;
        g.d = 1 << l;
        const int d0 = 1 << AL;
        g.nx = min(a.BX, a.BX * d0 / g.d / 2);
        g.nz = min(a.BZA, a.BZA * d0 / g.d / 2);
        g.nxp = sub == 1 ? min(a.BX, a.BX * d0 / g.d) : min(a.BX, a.BX * d0 / g.d / 4);
        g.nzp = sub == 1 ? min(a.BZA, a.BZA * d0 / g.d) : min(a.BZA, a.BZA * d0 / g.d / 4);
        g.h2 = a.BoxLength * a.BoxLength * g.d * g.d;
        return g;
    }
    // Element (i, k) as the reference addresses it, by the flat index
    // x*(nz+1) + z of Array3D::Index: a z past the top of a level reads the
    // next column. A flat index outside the array (the reference reads past
    // the end of the coarsest error array when BoxZAgar < 2^(8 - maxLevelsMG))
    // reads zero.
    __device__ __forceinline__ AgaRef at(double* const* arr, int l, int i, int k) const {
        const int X1 = a.mgx[l] + 1, Z1 = a.mgz[l] + 1, n = X1 * Z1;
        int p;
        if (i >= 0 && i < X1 && k >= 0 && k < Z1) {
            p = k * X1 + i;
        } else {
            const int f = i * Z1 + k;
            if (f < 0 || f >= n) {
                phantom[0] = phantom[1] = phantom[2] = 0.0;
                return {phantom[0], phantom[1], phantom[2]};
            }
            const int ii = f / Z1;
            p = (f - ii * Z1) * X1 + ii;
        }
        gdouble* b = G(arr[l]);
        return {b[p], b[p + n], b[p + 2 * n]};
    }
    __device__ __forceinline__ AgaRef A(int l, int i, int k) const { return at(a.aga, l, i, k); }
    __device__ __forceinline__ AgaRef R(int l, int i, int k) const { return at(a.res, l, i, k); }
    __device__ __forceinline__ AgaRef E(int l, int i, int k) const { return at(a.err, l, i, k); }
    // insideColonyDen(x, 0, 0); x past the last box reads as 0
    __device__ double inside0(int x) const { return (x >= 0 && x < a.BX) ? G(a.insT)[x] : 0.0; }
};

// One species of one level, [z][x]: element (i, k) at p[k * X1 + i] (global
// memory, or LDS for the coarsest level).
template <class T>
struct Col {
    T* p;
    int X1;
};

// relaxationAgarElementwise (ABM/Multigrid.cpp:161-249) down one column
// (k ascending) for one species: the reference's branches and operand
// order, with e(k-1) carried in a register and the loads that do not depend
// on it (R, its term RT = r h^2 / dt / D, the two neighbour columns of the
// other colour, the old e(k+1)) prefetched RING steps ahead; nn = h^2 / dt /
// D + 2. Nothing else writes this column during a
// colour pass, so the values match the sequential k loop exactly.
template <int RING, class T>
__device__ __forceinline__ void relax_column(const Col<T> E, const Col<T> R, const Col<T> RT, int i, int nx, int nz,
                                             bool isO, double nn, bool outer, bool dir0) {
    const bool edge = (i == 0 || i == nx);
    if (edge && !outer && !dir0) return;  // nothing in this column changes
    const int X1 = E.X1;
    const int il = edge ? (i == 0 ? 1 : i - 1) : i - 1;  // E(i-1, .) or the inner column
    const int ih = edge ? i : i + 1;                     // E(i+1, .), interior only
    // e(k+1) past the top is the next column's k = 0 (the flat index), and
    // past the last column nothing (read as 0)
    const int inext = min(i + 1, nx);
    const bool next_ok = i + 1 <= nx;
    double qr[RING], qt[RING], ql[RING], qh[RING], qn[RING];
    // loads without branches (so that they stay in flight across steps):
    // past the column end and on the edge an in-range element is loaded and
    // not used
    auto fetch = [&](int k, int j) {
        const int kk = min(k, nz);
        qr[j] = R.p[kk * X1 + i];
        qt[j] = RT.p[kk * X1 + i];
        ql[j] = E.p[kk * X1 + il];
        qh[j] = E.p[kk * X1 + ih];
        qn[j] = E.p[kk < nz ? (kk + 1) * X1 + i : inext];
    };
#pragma unroll
    for (int j = 0; j < RING; j++) fetch(j, j);
    double ek = E.p[i], em1 = 0.0;
    // steps k0 .. k0+RING-1 use ring slots 0 .. RING-1 (fixed registers: a
    // load stays in flight until its slot is used RING steps later). The
    // step body has no branches, so the loads' completion can be counted
    // exactly; steps past the column end store the last value again.
    for (int k0 = 0; k0 <= nz; k0 += RING) {
#pragma unroll
        for (int j = 0; j < RING; j++) {
            const int k = k0 + j;
            const double r = qr[j], t = qt[j], L = ql[j], H = qh[j], ekp1 = (k < nz || next_ok) ? qn[j] : 0.0;
            fetch(k + RING, j);
            // keep the loads here: the scheduler would otherwise sink them to
            // the step before their use, and every step would wait for memory
            __builtin_amdgcn_sched_barrier(0);
            // the numerator of the branch taken, in the reference's order of
            // operations - corners t + (p + q); edges t + e + p/2 + q/2 or
            // t + p/2 + q/2 + e; interior t + p/2 + q/2 + r/2 + s/2 - chosen
            // without branches; then one division
            const bool bot = k == 0, top = !bot && k == nz;
            const double n2 = t + (L + (bot ? ekp1 : em1));
            const double n3e = t + 0.5 * em1 + 0.5 * ekp1 + L;
            const double e3 = bot ? ekp1 : em1, p3 = bot ? L : H, q3 = bot ? H : L;
            const double n3 = t + e3 + 0.5 * p3 + 0.5 * q3;
            const double n4 = t + 0.5 * em1 + 0.5 * ekp1 + 0.5 * L + 0.5 * H;
            const double num = edge ? ((bot || top) ? n2 : n3e) : ((bot || top) ? n3 : n4);
            const double q = num / nn;
            // oxygen keeps r on the boundary rows and columns; a column or
            // a top point outside the outer level keeps its value
            const bool keepR = bot ? (dir0 || isO) : (edge && isO);
            const bool keep = !bot && !outer && (top || edge);
            const bool live = k <= nz;
            const double v = live ? (keep ? ek : (keepR ? r : q)) : em1;
            E.p[min(k, nz) * X1 + i] = v;
            em1 = v;
            ek = ekp1;
        }
    }
}

// The restricted residual terms: (Rp + 0.5*S*D*dt*4/h2 - e)/2 (f4 = true)
// or (Rp + 0.5*S*D*dt/h2*4 - e)*0.5 (f4 = false), S summed by the caller.
__device__ __forceinline__ double rr_a(double Rp, double S, double D, double dt, double h2, double e) {
    return (Rp + 0.5 * S * D * dt * 4 / h2 - e) / 2;
}
__device__ __forceinline__ double rr_b(double Rp, double S, double D, double dt, double h2, double e) {
    return (Rp + 0.5 * S * D * dt * 4 / h2 - e) * 0.5;
}
__device__ __forceinline__ double rr_c(double Rp, double S, double D, double dt, double h2, double e) {
    return (Rp + 0.5 * S * D * dt / h2 * 4 - e) * 0.5;
}

// Species accessors for the three multigrid species.
struct SpC {
    __device__ static gdouble& v(const AgaRef& g) { return g.CarbonAgar; }
};
struct SpO {
    __device__ static gdouble& v(const AgaRef& g) { return g.OxygenAgar; }
};
struct SpA {
    __device__ static gdouble& v(const AgaRef& g) { return g.AcetateAgar; }
};

// restrictionAgaElementwise (ABM/Multigrid.cpp:270-405).
__device__ __forceinline__ void restrict_point(const MG m, int l, int i, int k, const Geo g) {
    const EnvArgs& a = m.a;
    const double dt = m.dt, h2 = g.h2;
    const int nx = g.nx, nz = g.nz, nxp = g.nxp, nzp = g.nzp;
    const int i0 = min(max((i - nx / 2) * g.d + a.BX / 2, 0), a.BX);
    const int ip = 2 * i - nx + nxp / 2, kp = 2 * k;
    const int lp = l - 1;
    const double DC = a.DiffAgarCarbon, DO = a.DiffAgarO2, DA = a.DiffAgarAce;
    const AgaRef out = m.R(l, i, k);
    auto E = [&](int ii, int kk) -> AgaRef { return m.E(lp, ii, kk); };
    auto Rp = [&](int ii, int kk) -> AgaRef { return m.R(lp, ii, kk); };
    auto Ap = [&](int ii, int kk) -> AgaRef { return m.A(lp, ii, kk); };
    auto Al = [&](int ii, int kk) -> AgaRef { return m.A(l, ii, kk); };
    if (i == 0 || i == nx || k == nz) {
        if (l >= m.AL) {
            if (k == nz && i > 0 && i < nx) {
                out.CarbonAgar = rr_a(Rp(ip, kp).CarbonAgar,
                                      E(ip + 1, kp).CarbonAgar + E(ip - 1, kp).CarbonAgar + 2 * E(ip, kp - 1).CarbonAgar -
                                          4 * E(ip, kp).CarbonAgar,
                                      DC, dt, h2, E(ip, kp).CarbonAgar);
                out.OxygenAgar = rr_a(Rp(ip, kp).OxygenAgar,
                                      E(2 * i + 1, 2 * k).OxygenAgar + E(2 * i - 1, 2 * k).OxygenAgar +
                                          2 * E(2 * i, 2 * k - 1).OxygenAgar - 4 * E(2 * i, 2 * k).OxygenAgar,
                                      DO, dt, h2, E(2 * i, 2 * k).OxygenAgar);
                out.AcetateAgar = rr_a(Rp(ip, kp).AcetateAgar,
                                       E(2 * i + 1, 2 * k).AcetateAgar + E(2 * i - 1, 2 * k).AcetateAgar +
                                           2 * E(2 * i, 2 * k - 1).AcetateAgar - 4 * E(2 * i, 2 * k).AcetateAgar,
                                       DA, dt, h2, E(2 * i, 2 * k).AcetateAgar);
            } else {
                // i == 0 or i == nx: the x neighbour is on the inside, doubled
                const int s = i == 0 ? 1 : -1;
                double SC, SA;
                if (k == 0) {
                    SC = 2 * E(ip + s, kp).CarbonAgar + 2 * E(ip, kp + 1).CarbonAgar - 4 * E(ip, kp).CarbonAgar;
                    SA = 2 * E(2 * i + s, 2 * k).AcetateAgar + 2 * E(2 * i, 2 * k + 1).AcetateAgar -
                         4 * E(2 * i, 2 * k).AcetateAgar;
                } else if (k == nz) {
                    SC = 2 * E(ip + s, kp).CarbonAgar + 2 * E(ip, kp - 1).CarbonAgar - 4 * E(ip, kp).CarbonAgar;
                    SA = 2 * E(2 * i + s, 2 * k).AcetateAgar + 2 * E(2 * i, 2 * k - 1).AcetateAgar -
                         4 * E(2 * i, 2 * k).AcetateAgar;
                } else {
                    SC = 2 * E(ip + s, kp).CarbonAgar + E(ip, kp + 1).CarbonAgar + E(ip, kp - 1).CarbonAgar -
                         4 * E(ip, kp).CarbonAgar;
                    SA = 2 * E(2 * i + s, 2 * k).AcetateAgar + E(2 * i, 2 * k + 1).AcetateAgar +
                         E(2 * i, 2 * k - 1).AcetateAgar - 4 * E(2 * i, 2 * k).AcetateAgar;
                }
                out.CarbonAgar = rr_a(Rp(ip, kp).CarbonAgar, SC, DC, dt, h2, E(ip, kp).CarbonAgar);
                out.OxygenAgar = Rp(ip, kp).OxygenAgar - E(ip, kp).OxygenAgar;
                out.AcetateAgar = rr_a(Rp(ip, kp).AcetateAgar, SA, DA, dt, h2, E(2 * i, 2 * k).AcetateAgar);
            }
        }
    } else if (k == 0) {
        if (m.inside0(i0) > 0) {
            out.CarbonAgar = 0;
            out.OxygenAgar = 0;
            out.AcetateAgar = 0;
        } else if (ip > 0 && ip < nxp) {
            out.CarbonAgar = rr_b(Rp(ip, 0).CarbonAgar,
                                  E(ip + 1, 0).CarbonAgar + E(ip - 1, 0).CarbonAgar + 2 * E(ip, 1).CarbonAgar -
                                      4 * E(ip, 0).CarbonAgar,
                                  DC, dt, h2, E(ip, 0).CarbonAgar);
            out.AcetateAgar = rr_b(Rp(ip, 0).AcetateAgar,
                                   E(ip + 1, 0).AcetateAgar + E(ip - 1, 0).AcetateAgar + 2 * E(ip, 1).AcetateAgar -
                                       4 * E(ip, 0).AcetateAgar,
                                   DA, dt, h2, E(ip, 0).AcetateAgar);
            out.OxygenAgar = (Rp(ip, 0).OxygenAgar - E(ip, 0).OxygenAgar) * 0.5;
        } else if (l < m.AL && ip >= 0 && ip <= nxp && (ip == 0 || ip == nxp)) {
            if (ip == 0) {
                out.CarbonAgar = (Ap(ip + 2, 0).CarbonAgar + 0.5 * E(ip + 2, 0).CarbonAgar + Al(i - 1, 0).CarbonAgar +
                                  2 * Al(i, 1).CarbonAgar - 4 * Al(i, 0).CarbonAgar) *
                                 DC * dt / h2;
                out.AcetateAgar = (Ap(ip + 2, 0).AcetateAgar + 0.5 * E(ip + 2, 0).AcetateAgar +
                                   Al(i - 1, 0).AcetateAgar + 2 * Al(i, 1).AcetateAgar - 4 * Al(i, 0).AcetateAgar) *
                                  DA * dt / h2;
            } else {
                out.CarbonAgar = (Al(i + 1, 0).CarbonAgar + 0.5 * E(ip - 2, 0).CarbonAgar + Ap(ip - 2, 0).CarbonAgar +
                                  2 * Al(i, 1).CarbonAgar - 4 * Al(i, 0).CarbonAgar) *
                                 dt * DC / h2;
                out.AcetateAgar = (Al(i + 1, 0).AcetateAgar + 0.5 * E(ip - 2, 0).AcetateAgar +
                                   Ap(ip - 2, 0).AcetateAgar + 2 * Al(i, 1).AcetateAgar - 4 * Al(i, 0).AcetateAgar) *
                                  dt * DA / h2;
            }
            out.OxygenAgar = a.maxO2_Agar - Al(i, 0).OxygenAgar;
        }
    } else if (l < m.AL && ip >= 0 && ip <= nxp && kp <= nzp) {
        // composite boundary of the nested region (five-point residual of Aga)
        auto res5 = [&](double D, double p, double q, double r, double s, double c) {
            return (p + q + r + s - 4 * c) * D * dt / h2;
        };
        if (kp == nzp) {
            if (ip == 0 || ip == nxp) {
                out.CarbonAgar = res5(DC, Al(i + 1, k).CarbonAgar, Al(i - 1, k).CarbonAgar, Al(i, k + 1).CarbonAgar,
                                      Al(i, k - 1).CarbonAgar, Al(i, k).CarbonAgar);
                out.OxygenAgar = res5(DO, Al(i + 1, k).OxygenAgar, Al(i - 1, k).OxygenAgar, Al(i, k + 1).OxygenAgar,
                                      Al(i, k - 1).OxygenAgar, Al(i, k).OxygenAgar);
                out.AcetateAgar = res5(DA, Al(i + 1, k).AcetateAgar, Al(i - 1, k).AcetateAgar,
                                       Al(i, k + 1).AcetateAgar, Al(i, k - 1).AcetateAgar, Al(i, k).AcetateAgar);
            } else {
                out.CarbonAgar = (Al(i + 1, k).CarbonAgar + Al(i - 1, k).CarbonAgar + Al(i, k + 1).CarbonAgar +
                                  Ap(ip, kp - 2).CarbonAgar + 0.5 * E(ip, kp - 2).CarbonAgar - 4 * Al(i, k).CarbonAgar) *
                                 DC * dt / h2;
                out.OxygenAgar = (Al(i + 1, k).OxygenAgar + Al(i - 1, k).OxygenAgar + Al(i, k + 1).OxygenAgar +
                                  Ap(ip, kp - 2).OxygenAgar + 0.5 * E(ip, kp - 2).OxygenAgar - 4 * Al(i, k).OxygenAgar) *
                                 DO * dt / h2;
                out.AcetateAgar =
                    (Al(i + 1, k).AcetateAgar + Al(i - 1, k).AcetateAgar + Al(i, k + 1).AcetateAgar +
                     Ap(ip, kp - 2).AcetateAgar + 0.5 * E(ip, kp - 2).AcetateAgar - 4 * Al(i, k).AcetateAgar) *
                    DA * dt / h2;
            }
        } else if (ip == 0) {
            out.CarbonAgar = (Ap(ip + 2, kp).CarbonAgar + 0.5 * E(ip + 2, kp).CarbonAgar + Al(i - 1, k).CarbonAgar +
                              Al(i, k + 1).CarbonAgar + Al(i, k - 1).CarbonAgar - 4 * Al(i, k).CarbonAgar) *
                             DC * dt / h2;
            out.OxygenAgar = (Ap(ip + 2, kp).OxygenAgar + 0.5 * E(ip + 2, kp).OxygenAgar + Al(i - 1, k).OxygenAgar +
                              Al(i, k + 1).OxygenAgar + Al(i, k - 1).OxygenAgar - 4 * Al(i, k).OxygenAgar) *
                             DO * dt / h2;
            out.AcetateAgar = (Ap(ip + 2, kp).AcetateAgar + 0.5 * E(ip + 2, kp).AcetateAgar +
                               Al(i - 1, k).AcetateAgar + Al(i, k + 1).AcetateAgar + Al(i, k - 1).AcetateAgar -
                               4 * Al(i, k).AcetateAgar) *
                              DA * dt / h2;
        } else if (ip == nxp) {
            out.CarbonAgar = (Al(i + 1, k).CarbonAgar + 0.5 * E(ip - 2, kp).CarbonAgar + Ap(ip - 2, kp).CarbonAgar +
                              Al(i, k + 1).CarbonAgar + Al(i, k - 1).CarbonAgar - 4 * Al(i, k).CarbonAgar) *
                             DC * dt / h2;
            out.OxygenAgar = (Al(i + 1, k).OxygenAgar + 0.5 * E(ip - 2, kp).OxygenAgar + Ap(ip - 2, kp).OxygenAgar +
                              Al(i, k + 1).OxygenAgar + Al(i, k - 1).OxygenAgar - 4 * Al(i, k).OxygenAgar) *
                             DO * dt / h2;
            out.AcetateAgar = (Al(i + 1, k).AcetateAgar + 0.5 * E(ip - 2, kp).AcetateAgar +
                               Ap(ip - 2, kp).AcetateAgar + Al(i, k + 1).AcetateAgar + Al(i, k - 1).AcetateAgar -
                               4 * Al(i, k).AcetateAgar) *
                              DA * dt / h2;
        } else {
            out.CarbonAgar = rr_c(Rp(ip, kp).CarbonAgar,
                                  E(ip + 1, kp).CarbonAgar + E(ip - 1, kp).CarbonAgar + E(ip, kp + 1).CarbonAgar +
                                      E(ip, kp - 1).CarbonAgar - 4 * E(ip, kp).CarbonAgar,
                                  DC, dt, h2, E(ip, kp).CarbonAgar);
            out.OxygenAgar = rr_c(Rp(ip, kp).OxygenAgar,
                                  E(ip + 1, kp).OxygenAgar + E(ip - 1, kp).OxygenAgar + E(ip, kp + 1).OxygenAgar +
                                      E(ip, kp - 1).OxygenAgar - 4 * E(ip, kp).OxygenAgar,
                                  DO, dt, h2, E(ip, kp).OxygenAgar);
            out.AcetateAgar = rr_c(Rp(ip, kp).AcetateAgar,
                                   E(ip + 1, kp).AcetateAgar + E(ip - 1, kp).AcetateAgar + E(ip, kp + 1).AcetateAgar +
                                       E(ip, kp - 1).AcetateAgar - 4 * E(ip, kp).AcetateAgar,
                                   DA, dt, h2, E(ip, kp).AcetateAgar);
        }
    } else if (l >= m.AL) {
        out.CarbonAgar = rr_c(Rp(ip, kp).CarbonAgar,
                              E(ip + 1, kp).CarbonAgar + E(ip - 1, kp).CarbonAgar + E(ip, kp + 1).CarbonAgar +
                                  E(ip, kp - 1).CarbonAgar - 4 * E(ip, kp).CarbonAgar,
                              DC, dt, h2, E(ip, kp).CarbonAgar);
        out.OxygenAgar = rr_c(Rp(ip, kp).OxygenAgar,
                              E(ip + 1, kp).OxygenAgar + E(ip - 1, kp).OxygenAgar + E(ip, kp + 1).OxygenAgar +
                                  E(ip, kp - 1).OxygenAgar - 4 * E(ip, kp).OxygenAgar,
                              DO, dt, h2, E(ip, kp).OxygenAgar);
        out.AcetateAgar = rr_c(Rp(ip, kp).AcetateAgar,
                               E(ip + 1, kp).AcetateAgar + E(ip - 1, kp).AcetateAgar + E(ip, kp + 1).AcetateAgar +
                                   E(ip, kp - 1).AcetateAgar - 4 * E(ip, kp).AcetateAgar,
                               DA, dt, h2, E(ip, kp).AcetateAgar);
    }
}

// interpolationAgaElementwise (ABM/Multigrid.cpp:427-467): bilinear, added.
__device__ void interp_point(const MG m, int l, int i, int k, const Geo g) {
    const AgaRef e = m.E(l, i, k);
    const int nx = g.nx, nxp = g.nxp;
    const int ip = nxp / 2 - nx / 4 + (i % 2 == 0 ? i / 2 : (i - 1) / 2);
    auto C = [&](int ii, int kk) -> AgaRef { return m.E(l + 1, ii, kk); };
    if (i % 2 == 0) {
        if (k % 2 == 0) {
            const int kp = k / 2;
            e.CarbonAgar += C(ip, kp).CarbonAgar;
            e.OxygenAgar += C(ip, kp).OxygenAgar;
            e.AcetateAgar += C(ip, kp).AcetateAgar;
        } else {
            const int kp = (k - 1) / 2;
            e.CarbonAgar += 0.5 * (C(ip, kp).CarbonAgar + C(ip, kp + 1).CarbonAgar);
            e.OxygenAgar += 0.5 * (C(ip, kp).OxygenAgar + C(ip, kp + 1).OxygenAgar);
            e.AcetateAgar += 0.5 * (C(ip, kp).AcetateAgar + C(ip, kp + 1).AcetateAgar);
        }
    } else {
        if (k % 2 == 0) {
            const int kp = k / 2;
            e.CarbonAgar += 0.5 * (C(ip, kp).CarbonAgar + C(ip + 1, kp).CarbonAgar);
            e.OxygenAgar += 0.5 * (C(ip, kp).OxygenAgar + C(ip + 1, kp).OxygenAgar);
            e.AcetateAgar += 0.5 * (C(ip, kp).AcetateAgar + C(ip + 1, kp).AcetateAgar);
        } else {
            const int kp = (k - 1) / 2;
            e.CarbonAgar +=
                0.25 * (C(ip, kp).CarbonAgar + C(ip, kp + 1).CarbonAgar + C(ip + 1, kp).CarbonAgar + C(ip + 1, kp + 1).CarbonAgar);
            e.OxygenAgar +=
                0.25 * (C(ip, kp).OxygenAgar + C(ip, kp + 1).OxygenAgar + C(ip + 1, kp).OxygenAgar + C(ip + 1, kp + 1).OxygenAgar);
            e.AcetateAgar += 0.25 * (C(ip, kp).AcetateAgar + C(ip, kp + 1).AcetateAgar + C(ip + 1, kp).AcetateAgar +
                                     C(ip + 1, kp + 1).AcetateAgar);
        }
    }
}

// updateSolAgaElementwise (ABM/Multigrid.cpp:539-562): u += e, clamped.
__device__ void update_sol_point(const MG m, int l, int i, int k, const Geo g) {
    const EnvArgs& a = m.a;
    const AgaRef r = m.R(l, i, k);
    r.CarbonAgar = 0;
    r.OxygenAgar = 0;
    r.AcetateAgar = 0;
    const int ip = 2 * i - g.nx + g.nxp / 2, kp = 2 * k;
    const AgaRef e = m.E(l, i, k);
    if (l < m.AL && (l == 0 || ip <= 0 || ip >= g.nxp || kp >= g.nzp)) {
        const AgaRef u = m.A(l, i, k);
        u.CarbonAgar += e.CarbonAgar;
        u.OxygenAgar += e.OxygenAgar;
        u.AcetateAgar += e.AcetateAgar;
        u.CarbonAgar = smax(smin(u.CarbonAgar, a.maxCarbon), 0.0);
        u.OxygenAgar = smax(smin(u.OxygenAgar, a.maxO2_Agar), 0.0);
        u.AcetateAgar = smax(u.AcetateAgar, 0.0);
    }
    e.CarbonAgar = 0;
    e.OxygenAgar = 0;
    e.AcetateAgar = 0;
}

// updateResAgaElementwise (ABM/Multigrid.cpp:564-689): residual dt D Lap u
// of the nested levels with the composite-boundary neighbours.
__device__ __forceinline__ void update_res_point(const MG m, int l, int i, int k, const Geo g) {
    const EnvArgs& a = m.a;
    if (l >= m.AL) return;
    const double dt = m.dt, h2 = g.h2;
    const int nx = g.nx, nz = g.nz, nxp = g.nxp, nzp = g.nzp;
    const int ip = 2 * i - nx + nxp / 2, kp = 2 * k;
    const double DC = a.DiffAgarCarbon, DO = a.DiffAgarO2, DA = a.DiffAgarAce;
    const AgaRef out = m.R(l, i, k);
    auto U = [&](int ii, int kk) -> AgaRef { return m.A(l, ii, kk); };
    auto Up = [&](int ii, int kk) -> AgaRef { return m.A(l - 1, ii, kk); };
    if (i == 0 || i == nx) {
        if (l == m.AL - 1) {
            const int s = i == 0 ? 1 : -1;
            if (k == 0) {
                out.CarbonAgar = (2 * U(i + s, k).CarbonAgar + 2 * U(i, k + 1).CarbonAgar - 4 * U(i, k).CarbonAgar) * DC * dt / h2;
                out.OxygenAgar = 0;
                out.AcetateAgar =
                    (2 * U(i + s, k).AcetateAgar + 2 * U(i, k + 1).AcetateAgar - 4 * U(i, k).AcetateAgar) * DA * dt / h2;
            } else if (k == nz) {
                out.CarbonAgar = (2 * U(i + s, k).CarbonAgar + 2 * U(i, k - 1).CarbonAgar - 4 * U(i, k).CarbonAgar) * DC * dt / h2;
                out.OxygenAgar = 0;
                out.AcetateAgar =
                    (2 * U(i + s, k).AcetateAgar + 2 * U(i, k - 1).AcetateAgar - 4 * U(i, k).AcetateAgar) * DA * dt / h2;
            } else {
                out.CarbonAgar = (2 * U(i + s, k).CarbonAgar + U(i, k - 1).CarbonAgar + U(i, k + 1).CarbonAgar -
                                  4 * U(i, k).CarbonAgar) *
                                 DC * dt / h2;
                out.OxygenAgar = 0;
                out.AcetateAgar = (2 * U(i + s, k).AcetateAgar + U(i, k - 1).AcetateAgar + U(i, k + 1).AcetateAgar -
                                   4 * U(i, k).AcetateAgar) *
                                  DA * dt / h2;
            }
        }
    } else if (k == nz) {
        if (l == m.AL - 1) {
            out.CarbonAgar = (U(i + 1, k).CarbonAgar + U(i - 1, k).CarbonAgar + 2 * U(i, k - 1).CarbonAgar -
                              4 * U(i, k).CarbonAgar) *
                             dt * DC / h2;
            // the reference's oxygen term has no dt here
            out.OxygenAgar = (U(i + 1, k).OxygenAgar + U(i - 1, k).OxygenAgar + 2 * U(i, k - 1).OxygenAgar -
                              4 * U(i, k).OxygenAgar) *
                             DO / h2;
            out.AcetateAgar = (U(i + 1, k).AcetateAgar + U(i - 1, k).AcetateAgar + 2 * U(i, k - 1).AcetateAgar -
                               4 * U(i, k).AcetateAgar) *
                              dt * DA / h2;
        }
    } else if (k == 0) {
        if (l == 0) {
            if (a.inside[i * a.BZ] > 0) {
                out.CarbonAgar = 0;
                out.OxygenAgar = 0;
                out.AcetateAgar = 0;
            } else {
                out.CarbonAgar = (U(i + 1, k).CarbonAgar + U(i - 1, k).CarbonAgar + 2 * U(i, k + 1).CarbonAgar -
                                  4 * U(i, k).CarbonAgar) *
                                 dt * DC / h2;
                out.OxygenAgar = a.maxO2_Agar - U(i, k).OxygenAgar;
                out.AcetateAgar = (U(i + 1, k).AcetateAgar + U(i - 1, k).AcetateAgar + 2 * U(i, k + 1).AcetateAgar -
                                   4 * U(i, k).AcetateAgar) *
                                  dt * DA / h2;
            }
        } else if (!(ip > 0 && ip < nxp)) {
            if (ip == 0) {
                out.CarbonAgar = (Up(ip + 2, kp).CarbonAgar + U(i - 1, k).CarbonAgar + 2 * U(i, k + 1).CarbonAgar -
                                  4 * U(i, k).CarbonAgar) *
                                 dt * DC / h2;
                out.OxygenAgar = a.maxO2_Agar - U(i, k).OxygenAgar;
                out.AcetateAgar = (Up(ip + 2, kp).AcetateAgar + U(i - 1, k).AcetateAgar +
                                   2 * U(i, k + 1).AcetateAgar - 4 * U(i, k).AcetateAgar) *
                                  dt * DA / h2;
            } else if (ip == nxp) {
                out.CarbonAgar = (Up(ip - 2, kp).CarbonAgar + U(i + 1, k).CarbonAgar + 2 * U(i, k + 1).CarbonAgar -
                                  4 * U(i, k).CarbonAgar) *
                                 dt * DC / h2;
                out.OxygenAgar = a.maxO2_Agar - U(i, k).OxygenAgar;
                out.AcetateAgar = (Up(ip - 2, kp).AcetateAgar + U(i + 1, k).AcetateAgar +
                                   2 * U(i, k + 1).AcetateAgar - 4 * U(i, k).AcetateAgar) *
                                  dt * DA / h2;
            } else {
                out.CarbonAgar = (U(i + 1, k).CarbonAgar + U(i - 1, k).CarbonAgar + 2 * U(i, k + 1).CarbonAgar -
                                  4 * U(i, k).CarbonAgar) *
                                 dt * DC / h2;
                out.OxygenAgar = a.maxO2_Agar - U(i, k).OxygenAgar;
                out.AcetateAgar = (U(i + 1, k).AcetateAgar + U(i - 1, k).AcetateAgar + 2 * U(i, k + 1).AcetateAgar -
                                   4 * U(i, k).AcetateAgar) *
                                  dt * DA / h2;
            }
        }
    } else {
        // interior rows: (p + q + r + s - 4c) dt D / h2 with the composite neighbours
        auto lap = [&](double p, double q, double r, double s, double c, double D) {
            return (p + q + r + s - 4 * c) * dt * D / h2;
        };
        auto put = [&](auto P, auto Q, auto Rr, auto S) {
            out.CarbonAgar = lap(P(SpC()), Q(SpC()), Rr(SpC()), S(SpC()), U(i, k).CarbonAgar, DC);
            out.OxygenAgar = lap(P(SpO()), Q(SpO()), Rr(SpO()), S(SpO()), U(i, k).OxygenAgar, DO);
            out.AcetateAgar = lap(P(SpA()), Q(SpA()), Rr(SpA()), S(SpA()), U(i, k).AcetateAgar, DA);
        };
        auto u = [&](int ii, int kk) { return [&, ii, kk](auto sp) { return decltype(sp)::v(U(ii, kk)); }; };
        auto up = [&](int ii, int kk) { return [&, ii, kk](auto sp) { return decltype(sp)::v(Up(ii, kk)); }; };
        if (l > 0) {
            if (kp == nzp && ip > 0 && ip < nxp) put(u(i + 1, k), u(i - 1, k), up(ip, kp - 2), u(i, k + 1));
            else if (ip == 0 && kp < nzp) put(up(ip + 2, kp), u(i - 1, k), u(i, k - 1), u(i, k + 1));
            else if (ip == nxp && kp < nzp) put(u(i + 1, k), up(ip - 2, kp), u(i, k - 1), u(i, k + 1));
            else if (ip <= 0 || ip >= nxp || kp >= nzp) put(u(i + 1, k), u(i - 1, k), u(i, k - 1), u(i, k + 1));
        } else {
            put(u(i + 1, k), u(i - 1, k), u(i, k - 1), u(i, k + 1));
        }
    }
}

// NestSolAgaElementwise (ABM/Multigrid.cpp:711-724): inject the finer nested
// level where it covers this one.
__device__ void nest_point(const MG m, int l, int i, int k, const Geo g) {
    const int ip = 2 * i - g.nx + g.nxp / 2, kp = 2 * k;
    if (l < m.AL && ip > 0 && ip < g.nxp && kp < g.nzp) {
        const AgaRef u = m.A(l, i, k);
        const AgaRef f = m.A(l - 1, ip, kp);
        u.CarbonAgar = f.CarbonAgar;
        u.OxygenAgar = f.OxygenAgar;
        u.AcetateAgar = f.AcetateAgar;
    }
}

// acetateBufferSolve (ABM/Nutrients.cpp:806-855): Newton on the buffer
// equilibrium; writes pH, Hydrox, Acetic, Buffer of the point, returns the
// undissociated acetic acid.
__device__ double acetate_buffer(const EnvArgs& a, double Cac, bcg_env_cell& p) {
    const double aH = -(a.Cb * a.Kb) / (1e-4 + a.Kb);
    double x = 1e-7, y = Cac, z = a.Cb;
    for (int it = 0; it < 100; it++) {
        double J0[4] = {2 * x + y + z + (a.Kw / a.c1) + aH, x, x, -(x * x + x * (y + z + (a.Kw / a.c1) + aH) - a.Kw)};
        double J1[4] = {y, 2 * y + x + z + a.Ka + aH, y, -(y * y + y * (x + z + (a.Ka) + aH) - a.Ka * Cac)};
        double J2[4] = {z, z, 2 * z + x + y + a.Kb + aH, -(z * z + z * (x + y + (a.Kb) + aH) - a.Kb * a.Cb)};
        const double m1 = J1[0] / J0[0], m2 = J2[0] / J0[0];
        for (int i = 0; i < 4; i++) {
            J1[i] -= m1 * J0[i];
            J2[i] -= m2 * J0[i];
        }
        const double m3 = J2[1] / J1[1];
        for (int i = 0; i < 4; i++) J2[i] -= m3 * J1[i];
        const double s2 = J2[3] / J2[2];
        const double s1 = J1[3] / J1[1] - (J1[2] / J1[1]) * s2;
        const double s0 = J0[3] / J0[0] - (J0[2] / J0[0]) * s2 - (J0[1] / J0[0]) * s1;
        double mx = 0;
        if (fabs(s0) > mx) mx = fabs(s0);
        if (fabs(s1) > mx) mx = fabs(s1);
        if (fabs(s2) > mx) mx = fabs(s2);
        x += s0;
        y += s1;
        z += s2;
        if (mx < 1e-7) break;
    }
    p.pH = -log((x + y + z + aH) / 1000) / log(10.0);
    p.Hydrox = x;
    p.Acetic = y;
    p.Buffer = z;
    return (Cac - y);
}

// Per-update constants of the uptake kinetics (ABM/Nutrients.cpp:244-266).
// Working copies of the colony (fields C, W, O2, acetate in the order of
// bcg_aga, each [z][x]) and of interface or agar rows ([field][x]) during an
// update: the lanes of a wave take neighbouring x.
enum { FC = 0, FW = 1, FO = 2, FA = 3 };
// A colony buffer: field f of box (x, z) at p[f * NF + z * W + x - x0] -
// the whole grid in global memory (x0 = 0, W = BoxX), or the window the
// sweeps update plus its one-box halo in LDS.
template <class T>
struct ColS {
    T* p;
    int x0, W, NF;
    __device__ T& operator()(int f, int x, int z) const { return p[f * NF + z * W + x - x0]; }
};
struct RowS {
    gdouble* p;
    int BX;
    __device__ gdouble& operator()(int f, int x) const { return p[f * BX + x]; }
};

// UpdateWallMultigrid_OxygenAcetate (ABM/Nutrients.cpp:509-659) at level l,
// box x, from its operands: inside the colony (level 0) from the colony box
// above, e, and the agar row below, g; elsewhere from the x term (Cxx, Wxx,
// Axx), the point's previous values Wx and g. out: the interface row of
// level l of this buffer.
__device__ __forceinline__ void wall_update(const EnvArgs& a, const RowS out, int l, int x, bool inside,
                                            const double (&e)[4], const double (&g)[4], const double (&Wx)[4],
                                            double Cxx, double Wxx, double Axx) {
    if (inside) {
        const double eC = e[FC], eW = e[FW], eO = e[FO], eA = e[FA];
        const double Cnew = (a.DiffColonyCarbon * eC + a.DiffAgarCarbon * g[FC]) / (a.DiffColonyCarbon + a.DiffAgarCarbon);
        const double Onew = (a.DiffColonyO2 * eO + a.DiffAgarO2 * g[FO]) / (a.DiffColonyO2 + a.DiffAgarO2);
        out(FC, x) = smax(0.0, smin(Cnew, a.maxCarbon));
        out(FO, x) = smax(0.0, smin(Onew, a.maxO2));
        const double CAcenew = (a.DiffColonyAce * eA + a.DiffAgarAce * g[FA]) / (a.DiffColonyAce + a.DiffAgarAce);
        out(FA, x) = smax(0.0, CAcenew);
        const double Wnew = (a.DiffColony * eW + a.DiffAgar * g[FW]) / (a.DiffColony + a.DiffAgar);
        out(FW, x) = smax(0.0, Wnew);
        return;
    }
    const double hl = a.BoxLength * a.BoxLength * (double)(1 << (2 * l));
    const double Wzz = 2.0 * (g[FW] - Wx[FW]) / hl;
    const double Czz = 2.0 * (g[FC] - Wx[FC]) / hl;
    const double Azz = 2.0 * (g[FA] - Wx[FA]) / hl;
    const double Cnew = Wx[FC] + a.DiffAgarCarbon * (Cxx + Czz) * a.Cdt;
    const double Onew = a.maxO2;
    out(FC, x) = smax(0.0, smin(Cnew, a.maxCarbon));
    out(FO, x) = smax(0.0, smin(Onew, a.maxO2));
    const double CAcenew = Wx[FA] + a.DiffAgarAce * (Axx + Azz) * a.Cdt;
    out(FA, x) = smax(0.0, CAcenew);
    const double Wnew = Wx[FW] + a.DiffAgar * (Wxx + Wzz) * a.Cdt;
    out(FW, x) = smax(0.0, Wnew);
}

// UpdateEnvironment_OxygenAcetate (ABM/Nutrients.cpp:200-506), non-cross-
// feeding chemistry, one colony box. The reference also writes, in every
// sweep, fields that no later sweep reads: the pathway growth rates and
// maintenance terms of the box (into Env) and the acetate buffer state
// (acetateBufferSolve, into the previous buffer; its first call, on the box's
// own acetate, is overwritten by the second, on the vertex average). Those
// are computed when `full` is set: in the last two sweeps of an update, whose
// writes are the ones that remain in the two buffers.
//
// A box of row z = 0 also updates the interface point below it
// (UpdateWallMultigrid at level 0, x inside the swept window): in the
// reference's order that point comes before the box when dir = 1 (it reads
// the box's value of the sweep before) and after it when dir = -1 (the value
// just written). g0: the first agar row of the previous buffer; out0: the
// interface row of this buffer (both level 0).
template <class CS>
__device__ __forceinline__ void colony_point(const EnvArgs& a, const ChemConst& cc, const CS cur, const CS prev, const RowS pw,
                             bcg_env_cell* Env, bcg_env_cell* prevEnv, int x, int z, bool full, int dir,
                             const RowS g0, const RowS out0) {
    const int BX = a.BX;
    const gdouble* insT = G(a.insT);
    const gdouble* dshT = G(a.dshT);
    // Every operand of the box is loaded before any decision, so that the
    // loads overlap instead of waiting behind each branch. Below z = 0 the
    // loads read row 0 (values not used there); 1 <= x <= BX - 2 and
    // z <= BZ - 2 for every box the sweeps update.
    const int zd = z > 0 ? z - 1 : 0;
    const double Dp = dshT[z * BX + x];
    const double ins = insT[z * BX + x];
    const double Ixp = insT[z * BX + x + 1], Ixm = insT[z * BX + x - 1];
    const double Izp = insT[(z + 1) * BX + x], Izm = insT[zd * BX + x];
    // prev at (x, z), (x - 1, z), (x + 1, z), (x, z + 1), (x, z - 1),
    // (x - 1, z - 1); the wall row below at x - 1, x and x + 1; for z = 0,
    // the agar row below and this box's value of the sweep before
    double c[4], xm[4], xp[4], zp[4], zm[4], xmzm[4], wm[4], w0[4], wp[4], gg[4], co[4], nv[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        c[f] = prev(f, x, z);
        xm[f] = prev(f, x - 1, z);
        xp[f] = prev(f, x + 1, z);
        zp[f] = prev(f, x, z + 1);
        zm[f] = prev(f, x, zd);
        xmzm[f] = prev(f, x - 1, zd);
        wm[f] = pw(f, x - 1);
        w0[f] = pw(f, x);
        wp[f] = pw(f, x + 1);
        gg[f] = g0(f, x);
        co[f] = cur(f, x, 0);
    }
    // the sweeps update boxes with a density shift only
    const bool upd = Dp > DBL_EPSILON;
    bool wrote = false;
    if (upd && ins > 0.5) {
        // fC, fW and fI of the box itself enter nothing (fI = 0)
        double Cgr = 0, Wgr = 0, Agr = 0, Ogr = 0;
        const double mgrCarAers = cc.mgrCarAers, mgrCarAnas = cc.mgrCarAnas, mgrAceAers = cc.mgrAceAers;
        const double g1s = cc.g1s, g2s = cc.g2s, g3s = cc.g3s;
        if (x != 0) {
            // vertex averages of the four boxes (the wall row below z = 0)
            if (z == 0) {
                Cgr = (wm[FC] + xm[FC] + w0[FC] + c[FC]) / 4.0;
                Ogr = (wm[FO] + xm[FO] + w0[FO] + c[FO]) / 4.0;
                Agr = (wm[FA] + xm[FA] + w0[FA] + c[FA]) / 4.0;
            } else {
                Cgr = (xmzm[FC] + xm[FC] + zm[FC] + c[FC]) / 4.0;
                Ogr = (xmzm[FO] + xm[FO] + zm[FO] + c[FO]) / 4.0;
                Agr = (xmzm[FA] + xm[FA] + zm[FA] + c[FA]) / 4.0;
            }
        }
        bcg_env_cell& out = Env[x * a.BZ + z];
        if (full) (void)acetate_buffer(a, Agr, prevEnv[x * a.BZ + z]);
        const double fCgr_aer = Cgr / (Cgr + a.KC_aer);
        const double fCgr_ana = Cgr / (Cgr + a.KC_ana);
        const double fAgr = Agr / (Agr + a.KA);
        const double fOgr = Ogr / (Ogr + a.KO);
        const double fIgr = 0;
        (void)Wgr;
        const double l1 = smax((a.maxGrowthRateCarAer + mgrCarAers) * fCgr_aer - mgrCarAers, 0.0);
        const double l2 = smax((a.maxGrowthRateCarAna + mgrCarAnas) * fCgr_ana - mgrCarAnas, 0.0);
        const double l3 = smax((a.maxGrowthRateAceAer + mgrAceAers) * fAgr - mgrAceAers, 0.0);
        const double inh = smax(1 - fIgr, 0.0);
        const double qO = (a.qO2Car * l1 * fCgr_aer + a.qO2Ace * l3 * (1 - fCgr_aer)) * fOgr * inh +
                          a.qO2Carnt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr +
                          a.qO2Acent * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
        const double qC = (a.qCarAer * l1 * fCgr_aer * fOgr + a.qCarAna * l2 * (1 - fOgr)) * inh +
                          a.qCarAernt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr +
                          a.qCarAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
        const double qA = (a.qAceAer * l3 * (1 - fCgr_aer) * fOgr) * inh +
                          a.qAceAernt * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
        const double pA = (a.pAceAer * l1 * fCgr_aer * fOgr + a.pAceAna * l2 * (1 - fOgr)) * inh +
                          a.pAceAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
        if (full) {
            out.qCarmnt = a.qCarAernt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr + a.qCarAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
            out.qAcemnt = a.qAceAernt * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
            out.pAcemnt = a.pAceAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
            out.qOxymnt = a.qO2Carnt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr + a.qO2Acent * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
            // l1..l3 are recomputed by the reference with the same operands: same values
            out.GrCarAer = l1 * fCgr_aer * fOgr * inh;
            out.GrCarAna = l2 * (1 - fOgr) * inh;
            out.GrAceAer = l3 * (1 - fCgr_aer) * fOgr * inh;
            out.GrowthRate1 = out.GrCarAer + out.GrCarAna + out.GrAceAer;
            out.GrowthRate2 = out.GrowthRate1;
        }
        const double pW = a.Carbon_to_Waste * qC;
        const double qW = 0.0;
        if (!(x > 2)) flag(a.errflag, ERR_COLONY_EDGE);
        const double BL2 = a.BoxLength * a.BoxLength;
        const double cC = c[FC], cW = c[FW], cO = c[FO], cA = c[FA];
        double Czz, Wzz, Azz, Ozz;
        if (z == 0) {
            Wzz = (zp[FW] + w0[FW] - 2.0 * cW) / BL2;
            Czz = (zp[FC] + w0[FC] - 2.0 * cC) / BL2;
            Azz = (zp[FA] + w0[FA] - 2.0 * cA) / BL2;
            Ozz = (zp[FO] + w0[FO] - 2.0 * cO) / BL2;
        } else {
            Wzz = (zp[FW] + zm[FW] - 2.0 * cW) / BL2;
            Czz = (zp[FC] + zm[FC] - 2.0 * cC) / BL2;
            Azz = (zp[FA] + zm[FA] - 2.0 * cA) / BL2;
            Ozz = (zp[FO] + zm[FO] - 2.0 * cO) / BL2;
        }
        const double Cxx = (xp[FC] + xm[FC] - 2.0 * cC) / BL2;
        const double Wxx = (xp[FW] + xm[FW] - 2.0 * cW) / BL2;
        const double Axx = (xp[FA] + xm[FA] - 2.0 * cA) / BL2;
        const double Oxx = (xp[FO] + xm[FO] - 2.0 * cO) / BL2;
        const double Cnew = cC + (a.DiffColonyCarbon * (Cxx + Czz) - qC * 0.25 * Dp) * a.Cdt;
        const double CAcenew = cA + (a.DiffColonyAce * (Axx + Azz) + pA * 0.25 * Dp - qA * 0.25 * Dp) * a.Cdt;
        const double Onew = cO + (a.DiffColonyO2 * (Oxx + Ozz) - qO * 0.25 * Dp) * a.Cdt;
        nv[FC] = smax(0.0, smin(Cnew, a.maxCarbon));
        nv[FO] = smax(0.0, smin(Onew, a.maxO2));
        nv[FA] = smax(0.0, CAcenew);
        const double Wnew = cW + (a.DiffColony * (Wxx + Wzz) + pW * Dp - qW * Dp) * a.Cdt;
        nv[FW] = smax(0.0, Wnew);
        wrote = true;
    } else if (upd && ins > 0.3) {
        // colony surface: mean of the neighbouring colony boxes (x, z; at
        // z = 0 without the box below)
        const bool top = z > 0;
        const int ne = (Ixp > 0.5) + (Ixm > 0.5) + (Izp > 0.5) + (top ? (Izm > 0.5) : 0);
        auto mean = [&](int f) -> double {
            double s = xp[f] * (Ixp > 0.5) + xm[f] * (Ixm > 0.5) + zp[f] * (Izp > 0.5);
            if (top) s = s + zm[f] * (Izm > 0.5);
            return s / ne;
        };
        const double Cnew = mean(FC);
        const double Wnew = mean(FW);
        const double CAcenew = mean(FA);
        nv[FC] = smax(0.0, smin(Cnew, a.maxCarbon));
        nv[FO] = smax(0.0, smin(a.maxO2, a.maxO2));
        nv[FA] = smax(0.0, CAcenew);
        nv[FW] = smax(0.0, Wnew);
        wrote = true;
    }
    if (wrote) {
        cur(FC, x, z) = nv[FC];
        cur(FO, x, z) = nv[FO];
        cur(FA, x, z) = nv[FA];
        cur(FW, x, z) = nv[FW];
    }
    if (z == 0) {
        // the interface point below: inside the colony from this box (the
        // value before this sweep when dir = 1, after it when dir = -1),
        // elsewhere the interior form of its x term (1 <= x <= BX - 2)
        const bool after = dir == -1 && wrote;
        double e[4];
#pragma unroll
        for (int f = 0; f < 4; f++) e[f] = after ? nv[f] : co[f];
        const double hl = a.BoxLength * a.BoxLength * (double)(1 << (2 * 0));
        const double Wxx = (wm[FW] + wp[FW] - 2 * w0[FW]) / hl;
        const double Cxx = (wm[FC] + wp[FC] - 2 * w0[FC]) / hl;
        const double Axx = (wm[FA] + wp[FA] - 2 * w0[FA]) / hl;
        wall_update(a, out0, 0, x, ins > 0.5, e, gg, w0, Cxx, Wxx, Axx);
    }
}

// The interface point (l, x) with its operands read here. env: the colony
// being written this sweep; row: the first agar row of the previous buffer;
// wal / pwal: the interface rows of this and of the previous buffer (levels
// 4 BX apart).
template <class CS>
__device__ __forceinline__ void wall_point(const EnvArgs& a, const CS env, double* row, double* wal, double* pwal,
                                           int l, int x) {
    const int BX = a.BX;
    const RowS out{G(wal) + l * 4 * BX, BX}, W{G(pwal) + l * 4 * BX, BX};
    const RowS gr{G(row) + l * 4 * BX, BX};
    double e[4] = {0, 0, 0, 0}, g[4], Wx[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        g[f] = gr(f, x);
        Wx[f] = W(f, x);
    }
    const bool inside = l == 0 && G(a.insT)[x] > 0.5;
    if (inside) {
#pragma unroll
        for (int f = 0; f < 4; f++) e[f] = env(f, x, 0);
        wall_update(a, out, l, x, true, e, g, Wx, 0, 0, 0);
        return;
    }
    const double hl = a.BoxLength * a.BoxLength * (double)(1 << (2 * l));
    // covered interior of a coarser level: the reference leaves the x term
    // uninitialised there (values overwritten by the next copyFromMG); 0 here
    double Cxx = 0, Wxx = 0, Axx = 0;
    if (l > 0 && x >= BX / 4 - 1 && x <= BX * 3 / 4) {
        const RowS Wf{G(pwal) + (l - 1) * 4 * BX, BX};
        if (x == BX / 4 - 1) {
            Wxx = (W(FW, x - 1) + Wf(FW, 0) - 2 * W(FW, x)) / hl;
            Cxx = (W(FC, x - 1) + Wf(FC, 0) - 2 * W(FC, x)) / hl;
            Axx = (W(FA, x - 1) + Wf(FA, 0) - 2 * W(FA, x)) / hl;
        } else if (x == BX * 3 / 4) {
            Wxx = (W(FW, x + 1) + Wf(FW, BX - 2) - 2 * W(FW, x)) / hl;
            Cxx = (W(FC, x + 1) + Wf(FC, BX - 2) - 2 * W(FC, x)) / hl;
            Axx = (W(FA, x + 1) + Wf(FA, BX - 2) - 2 * W(FA, x)) / hl;
        }
    } else if (l < a.ML - 1 && (x == 0 || x == BX - 1)) {
        // coarser level outside: y parity 0 (BoxY = 1), weights 1 and 0
        const int ly = 0;
        const RowS Wc{G(pwal) + (l + 1) * 4 * BX, BX};
        const double w1 = 1.0 / (ly + 1), w2 = ly * 1.0 / (ly + 1);
        if (x == 0) {
            const int c = BX / 4 - 1;
            Wxx = (W(FW, x + 2) + w1 * Wc(FW, c) + w2 * Wc(FW, c) - 2 * W(FW, x)) / 4 / hl;
            Cxx = (W(FC, x + 2) + w1 * Wc(FC, c) + w2 * Wc(FC, c) - 2 * W(FC, x)) / 4 / hl;
            Axx = (W(FA, x + 2) + w1 * Wc(FA, c) + w2 * Wc(FA, c) - 2 * W(FA, x)) / 4 / hl;
        } else {
            const int c = BX * 3 / 4;
            Wxx = (W(FW, x - 1) + w1 * Wc(FW, c) + w2 * Wc(FW, c) - 2 * W(FW, x)) / hl;
            Cxx = (W(FC, x - 1) + w1 * Wc(FC, c) + w2 * Wc(FC, c) - 2 * W(FC, x)) / hl;
            Axx = (W(FA, x - 1) + w1 * Wc(FA, c) + w2 * Wc(FA, c) - 2 * W(FA, x)) / hl;
        }
    } else if (l == a.ML - 1 && (x == 0 || x == BX - 1)) {
        const int n = x == 0 ? x + 1 : x - 1;
        Wxx = (W(FW, n) - 1.0 * W(FW, x)) / hl;
        Cxx = (W(FC, n) - 1.0 * W(FC, x)) / hl;
        Axx = (W(FA, n) - 1.0 * W(FA, x)) / hl;
    } else {
        Wxx = (W(FW, x - 1) + W(FW, x + 1) - 2 * W(FW, x)) / hl;
        Cxx = (W(FC, x - 1) + W(FC, x + 1) - 2 * W(FC, x)) / hl;
        Axx = (W(FA, x - 1) + W(FA, x + 1) - 2 * W(FA, x)) / hl;
    }
    wall_update(a, out, l, x, false, e, g, Wx, Cxx, Wxx, Axx);
}

// Diffusivity of multigrid species sp (C, O2, acetate) and the
// relaxation's constants of level l.
__device__ __forceinline__ double species_D(const EnvArgs& a, int sp) {
    return sp == 0 ? a.DiffAgarCarbon : (sp == 1 ? a.DiffAgarO2 : a.DiffAgarAce);
}
__device__ __forceinline__ double level_h2(const EnvArgs& a, int l) {
    return a.BoxLength * a.BoxLength * (double)(1 << (2 * l));
}

// The residual term of the relaxation, r h^2 / dt / D, of every point of
// level l (the reference forms it in each relaxation; the same operations
// give the same value), entries t0, t0 + stride, ...
__device__ __forceinline__ void rterm_level(const EnvArgs& a, const MG m, int l, int t0, int stride) {
    const int npts = (a.mgx[l] + 1) * (a.mgz[l] + 1);
    const double h2 = level_h2(a, l), dt = m.dt;
    const gdouble* R = G(a.res[l]);
    gdouble* T = G(a.rt[l]);
    for (int f = t0; f < 3 * npts; f += stride) {
        const int sp = f / npts;
        T[f] = R[f] * h2 / dt / species_D(a, sp);
    }
}

// One colour pass of relaxationAgarAtLevel (ABM/Multigrid.cpp:135-159): the
// columns of parity `par`, each k ascending; (species, column) pairs t0,
// t0 + stride, ... of this thread. Columns of one parity read only columns
// of the other: the pass may be spread over any number of threads.
template <int RING>
__device__ __forceinline__ void relax_pass(const EnvArgs& a, const MG m, int l, int par, int t0, int stride) {
    const Geo g = m.geo(l, 1);
    const int X1 = a.mgx[l] + 1, npts = X1 * (a.mgz[l] + 1);
    const bool outer = l >= m.AL - 1;
    const double h2 = level_h2(a, l);
    const int d = 1 << l;
    const gdouble* insT = G(a.insT);
    gdouble* E = G(a.err[l]);
    gdouble* R = G(a.res[l]);
    gdouble* T = G(a.rt[l]);
    const int ncol = par <= g.nx ? (g.nx - par) / 2 + 1 : 0;
    for (int w = t0; w < 3 * ncol; w += stride) {
        const int sp = w / ncol, i = par + 2 * (w - sp * ncol);
        const int i0 = min(max((i - g.nx / 2) * d + a.BX / 2, 0), a.BX - 1);
        const double nn = h2 / m.dt / species_D(a, sp) + 2;
        relax_column<RING>(Col<gdouble>{E + sp * npts, X1}, Col<gdouble>{R + sp * npts, X1},
                           Col<gdouble>{T + sp * npts, X1}, i, g.nx, g.nz, sp == 1, nn, outer, insT[i0] > 0);
    }
}

// relaxationAgarAtLevel: parity 0, then parity 1, in one workgroup.
__device__ __forceinline__ void relax_level(const EnvArgs& a, const MG m, int l, int tid) {
    for (int par = 0; par < 2; par++) {
        relax_pass<RING_WG>(a, m, l, par, tid, ENV_THREADS);
        __syncthreads();
    }
}

// The 100 relaxations of the coarsest level (Multigrid::Vcycle): in wave 0,
// out of LDS, when the level fits (no workgroup barrier between the colour
// passes).
__device__ __forceinline__ void coarse_solve(const EnvArgs& a, const MG m, int tid, double* cs_e, double* cs_r,
                                             double* cs_t) {
    const int l = MG_LEVELS - 1;
    const Geo g = m.geo(l, 1);
    const int X1 = g.nx + 1, npts = X1 * (g.nz + 1);
    if (npts > COARSE_MAX || 3 * (g.nx / 2 + 1) > 64) {
        for (int j = 0; j < 100; j++) relax_level(a, m, l, tid);
        return;
    }
    if (tid < 64) {
        gdouble* E = G(a.err[l]);
        const gdouble* R = G(a.res[l]);
        const gdouble* T = G(a.rt[l]);
        for (int f = tid; f < npts; f += 64)
            for (int sp = 0; sp < 3; sp++) {
                cs_e[sp * COARSE_MAX + f] = E[sp * npts + f];
                cs_r[sp * COARSE_MAX + f] = R[sp * npts + f];
                cs_t[sp * COARSE_MAX + f] = T[sp * npts + f];
            }
        wave_sync();
        const bool outer = l >= m.AL - 1;
        const double h2 = level_h2(a, l);
        const int d = 1 << l;
        const gdouble* insT = G(a.insT);
        for (int j = 0; j < 100; j++) {
            for (int par = 0; par < 2; par++) {
                const int ncol = par <= g.nx ? (g.nx - par) / 2 + 1 : 0;
                if (tid < 3 * ncol) {
                    const int sp = tid / ncol, i = par + 2 * (tid - sp * ncol);
                    const int i0 = min(max((i - g.nx / 2) * d + a.BX / 2, 0), a.BX - 1);
                    const double nn = h2 / m.dt / species_D(a, sp) + 2;
                    relax_column<RING_WG>(Col<double>{cs_e + sp * COARSE_MAX, X1}, Col<double>{cs_r + sp * COARSE_MAX, X1},
                                          Col<double>{cs_t + sp * COARSE_MAX, X1}, i, g.nx, g.nz, sp == 1, nn, outer,
                                          insT[i0] > 0);
                }
                wave_sync();
            }
        }
        for (int f = tid; f < npts; f += 64)
            for (int sp = 0; sp < 3; sp++) E[sp * npts + f] = cs_e[sp * COARSE_MAX + f];
    }
    __syncthreads();
}

// The 100 relaxations of the coarsest level by the whole workgroup out of
// LDS (each species [z][x]), for
// levels too large for the one-wave path of coarse_solve (lds: 3 x 3 x npts
// doubles, E, R, then RT).
__device__ __forceinline__ void coarse_solve_lds(const EnvArgs& a, const MG m, int tid, double* lds) {
    const int l = MG_LEVELS - 1;
    const Geo g = m.geo(l, 1);
    const int X1 = a.mgx[l] + 1, npts = X1 * (a.mgz[l] + 1);
    double* cs_e = lds;
    double* cs_r = lds + 3 * npts;
    double* cs_t = lds + 6 * npts;
    gdouble* E = G(a.err[l]);
    const gdouble* R = G(a.res[l]);
    const gdouble* T = G(a.rt[l]);
    for (int f = tid; f < 3 * npts; f += ENV_THREADS) {
        cs_e[f] = E[f];
        cs_r[f] = R[f];
        cs_t[f] = T[f];
    }
    __syncthreads();
    const bool outer = l >= m.AL - 1;
    const double h2 = level_h2(a, l);
    const int d = 1 << l;
    const gdouble* insT = G(a.insT);
    for (int j = 0; j < 100; j++) {
        for (int par = 0; par < 2; par++) {
            const int ncol = par <= g.nx ? (g.nx - par) / 2 + 1 : 0;
            for (int w = tid; w < 3 * ncol; w += ENV_THREADS) {
                const int sp = w / ncol, i = par + 2 * (w - sp * ncol);
                const int i0 = min(max((i - g.nx / 2) * d + a.BX / 2, 0), a.BX - 1);
                const double nn = h2 / m.dt / species_D(a, sp) + 2;
                relax_column<RING_WG>(Col<double>{cs_e + sp * npts, X1}, Col<double>{cs_r + sp * npts, X1},
                                      Col<double>{cs_t + sp * npts, X1}, i, g.nx, g.nz, sp == 1, nn, outer,
                                      insT[i0] > 0);
            }
            __syncthreads();
        }
    }
    for (int f = tid; f < 3 * npts; f += ENV_THREADS) E[f] = cs_e[f];
    __syncthreads();
}

// The 100 relaxations of the coarsest level with one wave per species (the
// three error equations do not couple) and one lane per column, out of LDS
// (lds: E, R, RT of the three species, 3 x 3 x npts doubles): the colour
// passes of a species need only its own wave's synchronisation, and each
// lane's column constants are formed once. Requires nx + 1 <= 64.
__device__ __forceinline__ void coarse_solve_waves(const EnvArgs& a, const MG m, int tid, double* lds) {
    const int l = MG_LEVELS - 1;
    const Geo g = m.geo(l, 1);
    const int X1 = a.mgx[l] + 1, npts = X1 * (a.mgz[l] + 1);
    double* cs_e = lds;
    double* cs_r = lds + 3 * npts;
    double* cs_t = lds + 6 * npts;
    gdouble* E = G(a.err[l]);
    const gdouble* R = G(a.res[l]);
    const gdouble* T = G(a.rt[l]);
    for (int f = tid; f < 3 * npts; f += ENV_THREADS) {
        cs_e[f] = E[f];
        cs_r[f] = R[f];
        cs_t[f] = T[f];
    }
    __syncthreads();
    const int sp = tid / 64, i = tid - 64 * sp;
    if (sp < 3) {
        const bool col = i <= g.nx;
        const bool outer = l >= m.AL - 1;
        const int d = 1 << l;
        const int i0 = min(max((i - g.nx / 2) * d + a.BX / 2, 0), a.BX - 1);
        const bool dir0 = col && G(a.insT)[i0] > 0;
        const double nn = level_h2(a, l) / m.dt / species_D(a, sp) + 2;
        const Col<double> ce{cs_e + sp * npts, X1}, cr{cs_r + sp * npts, X1}, ct{cs_t + sp * npts, X1};
        for (int j = 0; j < 100; j++) {
            for (int par = 0; par < 2; par++) {
                // a one-row level takes no ring of loads ahead
                if (col && (i & 1) == par) {
                    if (g.nz == 0)
                        relax_column<1>(ce, cr, ct, i, g.nx, g.nz, sp == 1, nn, outer, dir0);
                    else
                        relax_column<RING_WG>(ce, cr, ct, i, g.nx, g.nz, sp == 1, nn, outer, dir0);
                }
                wave_sync();
            }
        }
    }
    __syncthreads();
    for (int f = tid; f < 3 * npts; f += ENV_THREADS) E[f] = cs_e[f];
    __syncthreads();
}

// Every point of a level, x fastest across the lanes (the passes that use it
// write each point from values no other point of the pass writes).
template <class F>
__device__ __forceinline__ void over(const Geo g, int tid, F fn) {
    const int n = (g.nx + 1) * (g.nz + 1);
    for (int w = tid; w < n; w += ENV_THREADS) {
        const int k = w / (g.nx + 1);
        fn(w - k * (g.nx + 1), k);
    }
}

// Multigrid::reset's zeroing of the error (all levels) and of the coarse
// residuals.
__device__ __forceinline__ void reset_zero(const MG m, int tid) {
    for (int l = 0; l < MG_LEVELS; l++) {
        over(m.geo(l, 1), tid, [=](int i, int k) {
            const AgaRef e = m.E(l, i, k);
            e.CarbonAgar = 0;
            e.OxygenAgar = 0;
            e.AcetateAgar = 0;
            if (l > 0) {
                const AgaRef r = m.R(l, i, k);
                r.CarbonAgar = 0;
                r.OxygenAgar = 0;
                r.AcetateAgar = 0;
            }
        });
    }
    __syncthreads();
}

// 200 Jacobi sweeps over the colony and the interface rows
// (UpdateEnvArrayMG_OxygenAcetate, ABM/Multigrid.cpp:1015-1112), direction
// alternating with the sweep count. Single-threaded order: z ascending starts
// with the interface rows (which then read the colony row of two sweeps ago),
// z descending ends with them. The only interface point that reads a colony
// box written in the same sweep is (level 0, x), reading box (x, 0): the
// thread of box (x, 0) updates both, in that order, so one barrier per sweep
// suffices. c0 / c1: the colony buffers (Environment, oldEnvironment).
// One sweep, the boxes and rows t0, t0 + stride, ... of this thread: the
// colony window first (index w < ncol * nrow), then the interface rows of
// every nested level outside it. sw: which buffer is Current; dir: +1 when
// the sweep count is even.
template <class T>
__device__ __forceinline__ void sweep_body(const EnvArgs& a, const ChemConst& cc, bool full, int dir, int sw, T* c0,
                                           T* c1, int x0, int W, int NF, int t0, int stride) {
    const int BX = a.BX, AL = a.ML;
    const int minX = a.bounds[0], maxX = a.bounds[1], maxH = a.bounds[4];
    const int ncol = maxX - minX + 1, nrow = maxH + 1;
    const ColS<T> cur{sw ? c1 : c0, x0, W, NF}, prev{sw ? c0 : c1, x0, W, NF};
    const RowS pw0{G(a.swall[sw ^ 1]), BX}, g0{G(a.srow[sw ^ 1]), BX}, out0{G(a.swall[sw]), BX};
    const int nw = ncol * nrow;
    for (int w = t0; w < nw + AL * BX; w += stride) {
        if (w < nw) {
            const int z = w / ncol, x = minX + (w - z * ncol);
            // (with the interface point below a box of row 0)
            colony_point(a, cc, cur, prev, pw0, a.colony[sw], a.colony[sw ^ 1], x, z, full, dir, g0, out0);
        } else {
            const int q = w - nw, l = q / BX, x = q - l * BX;
            if (l == 0 && x >= minX && x <= maxX) continue;
            wall_point(a, cur, a.srow[sw ^ 1], a.swall[sw], a.swall[sw ^ 1], l, x);
        }
    }
}

template <class T>
__device__ __forceinline__ void sweeps(const EnvArgs& a, const ChemConst& cc, int outer, int tid, T* c0, T* c1, int x0,
                                       int W, int NF, int& sw, int& count) {
    for (int j = 0; j < 200; j++) {
        const int dir = (count % 2 == 0) ? 1 : -1;
        const bool full = outer == a.n_outer - 1 && j >= 198;
        sweep_body(a, cc, full, dir, sw, c0, c1, x0, W, NF, tid, ENV_THREADS);
        __syncthreads();
        sw ^= 1;
        count++;
    }
}

}  // namespace

// The whole UpdateEnvArrayMG_OxygenAcetate in one workgroup.
__global__ void __launch_bounds__(ENV_THREADS) k_env_nutrients(EnvArgs args) {
    // the pointer tables are indexed by level and buffer at run time: keep
    // them in LDS rather than in per-lane scratch (data pointers are cast to
    // gdouble* where used, so that accesses are global, not flat)
    __shared__ EnvArgs a;
    __shared__ double cs_e[3 * COARSE_MAX], cs_r[3 * COARSE_MAX], cs_t[3 * COARSE_MAX];
    extern __shared__ double lds[];  // the colony window of both buffers, when it fits
    const int tid = threadIdx.x;
    if (tid == 0) a = args;
    __syncthreads();
    const ChemConst cc = a.cc;
    const MG m(a, 100.0, G(a.phantom) + 3 * tid);  // dt = 100 (ABM/Multigrid.cpp:993)
    const int AL = a.ML;
    const int BX = a.BX, BZ = a.BZ, NB = BX * BZ, RW = 4 * BX;

    // working copies: inside / density shift [z][x]; colony fields; the
    // interface rows and the first agar row of each nested level
    for (int w = tid; w < NB; w += ENV_THREADS) {
        const int z = w / BX, x = w - z * BX, b = x * BZ + z;
        G(a.insT)[w] = a.inside[b];
        G(a.dshT)[w] = a.dshift[b];
        for (int bb = 0; bb < 2; bb++) {
            const bcg_env_cell& c = a.colony[bb][b];
            gdouble* d = G(a.scol[bb]);
            d[FC * NB + w] = c.Carbon;
            d[FW * NB + w] = c.Waste;
            d[FO * NB + w] = c.Oxygen;
            d[FA * NB + w] = c.Acetate;
        }
    }
    for (int w = tid; w < AL * BX; w += ENV_THREADS) {
        const int l = w / BX, x = w - l * BX;
        for (int bb = 0; bb < 2; bb++) {
            const bcg_aga& q = a.wall[bb][l][x];
            const bcg_aga& r = a.agar[bb][l][x * a.BZA];
            gdouble* dw = G(a.swall[bb]) + l * RW;
            gdouble* dr = G(a.srow[bb]) + l * RW;
            dw[FC * BX + x] = q.CarbonAgar;
            dw[FW * BX + x] = q.WasteAgar;
            dw[FO * BX + x] = q.OxygenAgar;
            dw[FA * BX + x] = q.AcetateAgar;
            dr[FC * BX + x] = r.CarbonAgar;
            dr[FW * BX + x] = r.WasteAgar;
            dr[FO * BX + x] = r.OxygenAgar;
            dr[FA * BX + x] = r.AcetateAgar;
        }
    }
    // the window [minX-1, maxX+1] x [0, maxH+1] (the boxes the sweeps update
    // and the halo they read), fields C, W, O2, acetate, if the launch sized
    // the LDS for these bounds
    const bool use_lds = a.lds_colony && a.lds_w == a.bounds[1] - a.bounds[0] + 3 &&
                         a.lds_nf == a.lds_w * (a.bounds[4] + 2);
    if (use_lds) {
        const int x0 = a.bounds[0] - 1, W = a.lds_w, H = a.lds_nf / W;
        for (int w = tid; w < W * H; w += ENV_THREADS) {
            const int z = w / W, x = x0 + (w - z * W);
            for (int bb = 0; bb < 2; bb++) {
                const bcg_env_cell& c = a.colony[bb][x * BZ + z];
                double* d = lds + bb * 4 * a.lds_nf;
                d[FC * a.lds_nf + w] = c.Carbon;
                d[FW * a.lds_nf + w] = c.Waste;
                d[FO * a.lds_nf + w] = c.Oxygen;
                d[FA * a.lds_nf + w] = c.Acetate;
            }
        }
    }
    __syncthreads();

    int count = 0;
    int sw = 0;  // 0: Current = Environment / FieldAgar / FieldWall
    for (int outer = 0; outer < a.n_outer; outer++) {
        // --- Multigrid::solve(CurrentAgar, CurrentWall, PreviousAgar, PreviousWall, 3, 100, 100)
        // copyToMG: colony footprint from the wall, and the last column of the
        // outermost nested level from its first
        {
            const RowS cw{G(a.swall[sw]), BX};
            const gdouble* insT = G(a.insT);
            for (int w = tid; w < BX - 1; w += ENV_THREADS) {
                const int i = w + 1;
                if (insT[i] > 0) {
                    const AgaRef u = m.A(0, i, 0);
                    u.CarbonAgar = cw(FC, i);
                    u.OxygenAgar = cw(FO, i);
                    u.AcetateAgar = cw(FA, i);
                }
            }
        }
        for (int k = tid; k <= a.BZA; k += ENV_THREADS) {
            const AgaRef u = m.A(AL - 1, BX, k);
            const AgaRef s = m.A(AL - 1, 0, k);
            u.CarbonAgar = s.CarbonAgar;
            u.OxygenAgar = s.OxygenAgar;
            u.AcetateAgar = s.AcetateAgar;
        }
        __syncthreads();
        // setup: reset, updateSolAga, updateResAga
        reset_zero(m, tid);
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = m.geo(l, 1);
            over(g, tid, [=](int i, int k) { update_sol_point(m, l, i, k, g); });
        }
        __syncthreads();
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = m.geo(l, 1);
            over(g, tid, [=](int i, int k) { update_res_point(m, l, i, k, g); });
        }
        __syncthreads();
        rterm_level(a, m, 0, tid, ENV_THREADS);
        __syncthreads();
        for (int cyc = 0; cyc < 3; cyc++) {
            for (int l = 0; l < MG_LEVELS - 1; l++) {
                for (int j = 0; j < 3; j++) relax_level(a, m, l, tid);
                const Geo g = m.geo(l + 1, 1);
                over(g, tid, [=](int i, int k) { restrict_point(m, l + 1, i, k, g); });
                __syncthreads();
                rterm_level(a, m, l + 1, tid, ENV_THREADS);
                __syncthreads();
            }
            coarse_solve(a, m, tid, cs_e, cs_r, cs_t);
            for (int l = MG_LEVELS - 2; l >= 0; l--) {
                const Geo g = m.geo(l, 2);
                over(g, tid, [=](int i, int k) { interp_point(m, l, i, k, g); });
                __syncthreads();
                for (int j = 0; j < 3; j++) relax_level(a, m, l, tid);
            }
        }
        // reset: updateSolAga, rewriteSolBdryAga, updateResAga, zeroing, NestSolAga
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = m.geo(l, 1);
            over(g, tid, [=](int i, int k) { update_sol_point(m, l, i, k, g); });
        }
        __syncthreads();
        for (int l = 0; l < AL - 1; l++) {
            const Geo g = m.geo(l, 2);
            const int nx = g.nx, nz = g.nz, nx2 = g.nxp;
            for (int i = tid; i <= nx; i += ENV_THREADS) {
                const int i2 = nx2 / 2 + i / 2 - nx / 4, i2p1 = nx2 / 2 + (i + 1) / 2 - nx / 4;
                const AgaRef u = m.A(l, i, nz);
                const AgaRef p = m.A(l + 1, i2, nz / 2), q = m.A(l + 1, i2p1, nz / 2);
                u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
                u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
                u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
            }
            for (int k = tid; k < nz; k += ENV_THREADS) {
                for (int side = 0; side < 2; side++) {
                    const int i = side ? nx : 0;
                    const int i2 = nx2 / 2 + i / 2 - nx / 4;
                    const AgaRef u = m.A(l, i, k);
                    const AgaRef p = m.A(l + 1, i2, k / 2), q = m.A(l + 1, i2, (k + 1) / 2);
                    u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
                    u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
                    u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
                }
            }
            __syncthreads();
        }
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = m.geo(l, 1);
            over(g, tid, [=](int i, int k) { update_res_point(m, l, i, k, g); });
        }
        __syncthreads();
        reset_zero(m, tid);
        for (int l = 1; l <= AL; l++) {
            const Geo g = m.geo(l, 1);
            over(g, tid, [=](int i, int k) { nest_point(m, l, i, k, g); });
            __syncthreads();
        }
        // copyFromMG into both buffers of the nested levels: the interface
        // rows (k = 0) here, the agar (k >= 1) when the update ends (nothing
        // reads it before then but the first agar row, kept here too)
        for (int l = 0; l < AL; l++) {
            for (int w = tid; w < 2 * BX; w += ENV_THREADS) {
                const int k = w / BX, i = w - k * BX;
                const AgaRef s = m.A(l, i, k);
                for (int bb = 0; bb < 2; bb++) {
                    const RowS d{G(k == 0 ? a.swall[bb] : a.srow[bb]) + l * RW, BX};
                    d(FC, i) = s.CarbonAgar;
                    d(FO, i) = s.OxygenAgar;
                    d(FA, i) = s.AcetateAgar;
                }
            }
        }
        __syncthreads();

        // --- 200 Jacobi sweeps (the colony from LDS when its window fits)
        if (use_lds)
            sweeps<double>(a, cc, outer, tid, lds, lds + 4 * a.lds_nf, a.bounds[0] - 1, a.lds_w, a.lds_nf, sw, count);
        else
            sweeps<gdouble>(a, cc, outer, tid, G(a.scol[0]), G(a.scol[1]), 0, BX, NB, sw, count);
    }

    // write the working copies back: colony fields and interface rows of
    // both buffers, and the agar of the last copyFromMG
    if (use_lds) {
        const int x0 = a.bounds[0] - 1, W = a.lds_w, H = a.lds_nf / W;
        for (int w = tid; w < W * H; w += ENV_THREADS) {
            const int z = w / W, x = x0 + (w - z * W);
            for (int bb = 0; bb < 2; bb++) {
                bcg_env_cell& c = a.colony[bb][x * BZ + z];
                const double* d = lds + bb * 4 * a.lds_nf;
                c.Carbon = d[FC * a.lds_nf + w];
                c.Waste = d[FW * a.lds_nf + w];
                c.Oxygen = d[FO * a.lds_nf + w];
                c.Acetate = d[FA * a.lds_nf + w];
            }
        }
    } else {
        for (int w = tid; w < NB; w += ENV_THREADS) {
            const int x = w / BZ, z = w - x * BZ, s = z * BX + x;
            for (int bb = 0; bb < 2; bb++) {
                bcg_env_cell& c = a.colony[bb][w];
                const gdouble* d = G(a.scol[bb]);
                c.Carbon = d[FC * NB + s];
                c.Waste = d[FW * NB + s];
                c.Oxygen = d[FO * NB + s];
                c.Acetate = d[FA * NB + s];
            }
        }
    }
    for (int w = tid; w < AL * BX; w += ENV_THREADS) {
        const int l = w / BX, x = w - l * BX;
        for (int bb = 0; bb < 2; bb++) {
            bcg_aga& q = a.wall[bb][l][x];
            const gdouble* dw = G(a.swall[bb]) + l * RW;
            q.CarbonAgar = dw[FC * BX + x];
            q.WasteAgar = dw[FW * BX + x];
            q.OxygenAgar = dw[FO * BX + x];
            q.AcetateAgar = dw[FA * BX + x];
        }
    }
    if (a.n_outer > 0) {
        for (int l = 0; l < AL; l++)
            for (int w = tid; w < BX * a.BZA; w += ENV_THREADS) {
                const int i = w / a.BZA, k = w - i * a.BZA + 1;
                const AgaRef s = m.A(l, i, k);
                for (int bb = 0; bb < 2; bb++) {
                    bcg_aga& d = a.agar[bb][l][w];
                    d.CarbonAgar = s.CarbonAgar;
                    d.OxygenAgar = s.OxygenAgar;
                    d.AcetateAgar = s.AcetateAgar;
                }
            }
    }
}

// ------------------------------------------------ the update in phases
// UpdateEnvArrayMG_OxygenAcetate as a sequence of launches, in the order of
// the one-workgroup kernel above: every barrier of that kernel is a launch
// boundary here, so each phase computes exactly what the same phase
// computes there, spread over the device. The host (bcg_env_nutrients)
// records one outer iteration as a graph and replays it.

namespace {

__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int gstride() { return gridDim.x * blockDim.x; }

// Every point of a level across the grid, x fastest.
template <class F>
__device__ __forceinline__ void over_grid(const Geo g, F fn) {
    const int n = (g.nx + 1) * (g.nz + 1);
    for (int w = gtid(); w < n; w += gstride()) {
        const int k = w / (g.nx + 1);
        fn(w - k * (g.nx + 1), k);
    }
}

__device__ __forceinline__ MG grid_mg(const EnvArgs& a) { return MG(a, 100.0, G(a.phantom) + 3 * gtid()); }

// The phase kernels read their arguments from one device copy per update:
// the pointer tables are indexed by level at run time, which from memory is
// a uniform load (by value, the struct would be copied to scratch memory).

}  // namespace

// Working copies (as the start of k_env_nutrients): insideColonyDen and
// DensityShiftP [z][x], colony fields C, W, O2, acetate of both buffers, the
// interface rows and first agar rows of both buffers.
__global__ void __launch_bounds__(PH_THREADS) k_nut_begin(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const int BX = a.BX, BZ = a.BZ, NB = BX * BZ, RW = 4 * BX, AL = a.ML;
    for (int w = gtid(); w < NB; w += gstride()) {
        const int z = w / BX, x = w - z * BX, b = x * BZ + z;
        G(a.insT)[w] = a.inside[b];
        G(a.dshT)[w] = a.dshift[b];
        for (int bb = 0; bb < 2; bb++) {
            const bcg_env_cell& c = a.colony[bb][b];
            gdouble* d = G(a.scol[bb]);
            d[FC * NB + w] = c.Carbon;
            d[FW * NB + w] = c.Waste;
            d[FO * NB + w] = c.Oxygen;
            d[FA * NB + w] = c.Acetate;
        }
    }
    for (int w = gtid(); w < AL * BX; w += gstride()) {
        const int l = w / BX, x = w - l * BX;
        for (int bb = 0; bb < 2; bb++) {
            const bcg_aga& q = a.wall[bb][l][x];
            const bcg_aga& r = a.agar[bb][l][x * a.BZA];
            gdouble* dw = G(a.swall[bb]) + l * RW;
            gdouble* dr = G(a.srow[bb]) + l * RW;
            dw[FC * BX + x] = q.CarbonAgar;
            dw[FW * BX + x] = q.WasteAgar;
            dw[FO * BX + x] = q.OxygenAgar;
            dw[FA * BX + x] = q.AcetateAgar;
            dr[FC * BX + x] = r.CarbonAgar;
            dr[FW * BX + x] = r.WasteAgar;
            dr[FO * BX + x] = r.OxygenAgar;
            dr[FA * BX + x] = r.AcetateAgar;
        }
    }
}

// copyToMG: the colony footprint from the wall; the last column of the
// outermost nested level from its first (the two touch different points).
__global__ void __launch_bounds__(PH_THREADS) k_nut_copy_to_mg(const EnvArgs* __restrict__ args, int sw) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const int BX = a.BX, AL = a.ML;
    const RowS cw{G(a.swall[sw]), BX};
    const gdouble* insT = G(a.insT);
    for (int w = gtid(); w < BX - 1; w += gstride()) {
        const int i = w + 1;
        if (insT[i] > 0) {
            const AgaRef u = m.A(0, i, 0);
            u.CarbonAgar = cw(FC, i);
            u.OxygenAgar = cw(FO, i);
            u.AcetateAgar = cw(FA, i);
        }
    }
    for (int k = gtid(); k <= a.BZA; k += gstride()) {
        const AgaRef u = m.A(AL - 1, BX, k);
        const AgaRef s = m.A(AL - 1, 0, k);
        u.CarbonAgar = s.CarbonAgar;
        u.OxygenAgar = s.OxygenAgar;
        u.AcetateAgar = s.AcetateAgar;
    }
}

// Multigrid::reset's zeroing: the error on every level, the residual on the
// coarse ones.
__global__ void __launch_bounds__(PH_THREADS) k_nut_reset_zero(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    for (int l = 0; l < MG_LEVELS; l++) {
        over_grid(m.geo(l, 1), [=](int i, int k) {
            const AgaRef e = m.E(l, i, k);
            e.CarbonAgar = 0;
            e.OxygenAgar = 0;
            e.AcetateAgar = 0;
            if (l > 0) {
                const AgaRef r = m.R(l, i, k);
                r.CarbonAgar = 0;
                r.OxygenAgar = 0;
                r.AcetateAgar = 0;
            }
        });
    }
}

// updateSolAga of every nested level (each reads and writes its own level).
__global__ void __launch_bounds__(PH_THREADS) k_nut_update_sol(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    for (int l = 0; l <= a.ML - 1; l++) {
        const Geo g = m.geo(l, 1);
        over_grid(g, [=](int i, int k) { update_sol_point(m, l, i, k, g); });
    }
}

// updateResAga of every nested level (writes residuals, reads solutions).
__global__ void __launch_bounds__(PH_THREADS) k_nut_update_res(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    for (int l = 0; l <= a.ML - 1; l++) {
        const Geo g = m.geo(l, 1);
        over_grid(g, [=](int i, int k) { update_res_point(m, l, i, k, g); });
    }
}

// One colour pass of the relaxation of level l.
__global__ void __launch_bounds__(PH_RELAX_THREADS) k_nut_relax(const EnvArgs* __restrict__ args, int l, int par) {
    const EnvArgs& a = *args;
    relax_pass<RING_WIDE>(a, grid_mg(a), l, par, gtid(), gstride());
}

// Level l's residual restricted from level l - 1.
__global__ void __launch_bounds__(PH_THREADS) k_nut_restrict(const EnvArgs* __restrict__ args, int l) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const Geo g = m.geo(l, 1);
    over_grid(g, [=](int i, int k) { restrict_point(m, l, i, k, g); });
}

// Level l + 1's error interpolated onto level l.
__global__ void __launch_bounds__(PH_THREADS) k_nut_interp(const EnvArgs* __restrict__ args, int l) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const Geo g = m.geo(l, 2);
    over_grid(g, [=](int i, int k) { interp_point(m, l, i, k, g); });
}

// The relaxation's residual term of level l (after its residual changed).
__global__ void __launch_bounds__(PH_THREADS) k_nut_rterm(const EnvArgs* __restrict__ args, int l) {
    const EnvArgs& a = *args;
    rterm_level(a, grid_mg(a), l, gtid(), gstride());
}

// The V-cycle from level lc down and back up to it, in one workgroup (the
// levels below lc are narrow: a barrier costs less than a launch there).
__global__ void __launch_bounds__(ENV_THREADS) k_nut_vcycle_tail(EnvArgs args, int lc, int coarse_mode) {
    __shared__ EnvArgs a;
    __shared__ double cs_e[3 * COARSE_MAX], cs_r[3 * COARSE_MAX], cs_t[3 * COARSE_MAX];
    extern __shared__ double lds[];  // the coarsest level, when coarse_mode > 0
    const int tid = threadIdx.x;
    if (tid == 0) a = args;
    __syncthreads();
    const MG m(a, 100.0, G(a.phantom) + 3 * tid);
    for (int l = lc; l < MG_LEVELS - 1; l++) {
        for (int j = 0; j < 3; j++) relax_level(a, m, l, tid);
        const Geo g = m.geo(l + 1, 1);
        over(g, tid, [=](int i, int k) { restrict_point(m, l + 1, i, k, g); });
        __syncthreads();
        rterm_level(a, m, l + 1, tid, ENV_THREADS);
        __syncthreads();
    }
    if (coarse_mode == 2)
        coarse_solve_waves(a, m, tid, lds);
    else if (coarse_mode == 1)
        coarse_solve_lds(a, m, tid, lds);
    else
        coarse_solve(a, m, tid, cs_e, cs_r, cs_t);
    for (int l = MG_LEVELS - 2; l >= lc; l--) {
        const Geo g = m.geo(l, 2);
        over(g, tid, [=](int i, int k) { interp_point(m, l, i, k, g); });
        __syncthreads();
        for (int j = 0; j < 3; j++) relax_level(a, m, l, tid);
    }
}

// rewriteSolBdryAga of level l from level l + 1.
__global__ void __launch_bounds__(PH_THREADS) k_nut_rewrite(const EnvArgs* __restrict__ args, int l) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const Geo g = m.geo(l, 2);
    const int nx = g.nx, nz = g.nz, nx2 = g.nxp;
    for (int i = gtid(); i <= nx; i += gstride()) {
        const int i2 = nx2 / 2 + i / 2 - nx / 4, i2p1 = nx2 / 2 + (i + 1) / 2 - nx / 4;
        const AgaRef u = m.A(l, i, nz);
        const AgaRef p = m.A(l + 1, i2, nz / 2), q = m.A(l + 1, i2p1, nz / 2);
        u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
        u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
        u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
    }
    for (int w = gtid(); w < 2 * nz; w += gstride()) {
        const int side = w / nz, k = w - side * nz;
        const int i = side ? nx : 0;
        const int i2 = nx2 / 2 + i / 2 - nx / 4;
        const AgaRef u = m.A(l, i, k);
        const AgaRef p = m.A(l + 1, i2, k / 2), q = m.A(l + 1, i2, (k + 1) / 2);
        u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
        u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
        u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
    }
}

// NestSolAga of level l from level l - 1.
__global__ void __launch_bounds__(PH_THREADS) k_nut_nest(const EnvArgs* __restrict__ args, int l) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const Geo g = m.geo(l, 1);
    over_grid(g, [=](int i, int k) { nest_point(m, l, i, k, g); });
}

// copyFromMG into both buffers' interface rows (k = 0) and first agar rows.
__global__ void __launch_bounds__(PH_THREADS) k_nut_copy_from_mg(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const int BX = a.BX, RW = 4 * BX;
    for (int w = gtid(); w < a.ML * 2 * BX; w += gstride()) {
        const int l = w / (2 * BX), r = w - l * 2 * BX, k = r / BX, i = r - k * BX;
        const AgaRef s = m.A(l, i, k);
        for (int bb = 0; bb < 2; bb++) {
            const RowS d{G(k == 0 ? a.swall[bb] : a.srow[bb]) + l * RW, BX};
            d(FC, i) = s.CarbonAgar;
            d(FO, i) = s.OxygenAgar;
            d(FA, i) = s.AcetateAgar;
        }
    }
}

// One sweep over the whole device, the colony in global memory.
__global__ void __launch_bounds__(PH_THREADS) k_nut_sweep(const EnvArgs* __restrict__ args, int full, int dir, int sw) {
    const EnvArgs& a = *args;
    const ChemConst cc = a.cc;
    const int NB = a.BX * a.BZ;
    sweep_body<gdouble>(a, cc, full != 0, dir, sw, G(a.scol[0]), G(a.scol[1]), 0, a.BX, NB, gtid(), gstride());
}

// The 200 sweeps of one outer iteration in one workgroup, with the colony
// window [minX-1, maxX+1] x [0, maxH+1] in LDS when the launch sized it
// (loaded from and stored back to the global working copy).
__global__ void __launch_bounds__(ENV_THREADS) k_nut_sweeps_wg(EnvArgs args, int last) {
    __shared__ EnvArgs a;
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    if (tid == 0) a = args;
    __syncthreads();
    const ChemConst cc = a.cc;
    const int BX = a.BX, NB = BX * a.BZ;
    const bool use_lds = a.lds_colony && a.lds_w == a.bounds[1] - a.bounds[0] + 3 &&
                         a.lds_nf == a.lds_w * (a.bounds[4] + 2);
    int sw = 0, count = 0;  // 200 sweeps an outer iteration: both return to these
    const int outer = last ? a.n_outer - 1 : 0;
    if (use_lds) {
        const int x0 = a.bounds[0] - 1, W = a.lds_w, H = a.lds_nf / W;
        for (int w = tid; w < W * H; w += ENV_THREADS) {
            const int z = w / W, x = x0 + (w - z * W);
            for (int bb = 0; bb < 2; bb++)
                for (int f = 0; f < 4; f++) lds[(bb * 4 + f) * a.lds_nf + w] = G(a.scol[bb])[f * NB + z * BX + x];
        }
        __syncthreads();
        sweeps<double>(a, cc, outer, tid, lds, lds + 4 * a.lds_nf, x0, W, a.lds_nf, sw, count);
        for (int w = tid; w < W * H; w += ENV_THREADS) {
            const int z = w / W, x = x0 + (w - z * W);
            for (int bb = 0; bb < 2; bb++)
                for (int f = 0; f < 4; f++) G(a.scol[bb])[f * NB + z * BX + x] = lds[(bb * 4 + f) * a.lds_nf + w];
        }
    } else {
        sweeps<gdouble>(a, cc, outer, tid, G(a.scol[0]), G(a.scol[1]), 0, BX, NB, sw, count);
    }
}

// Write the working copies back (as the end of k_env_nutrients, from the
// global working copy): colony fields and interface rows of both buffers,
// and the agar of the last copyFromMG.
__global__ void __launch_bounds__(PH_THREADS) k_nut_end(const EnvArgs* __restrict__ args) {
    const EnvArgs& a = *args;
    const MG m = grid_mg(a);
    const int BX = a.BX, BZ = a.BZ, NB = BX * BZ, RW = 4 * BX, AL = a.ML;
    for (int w = gtid(); w < NB; w += gstride()) {
        const int x = w / BZ, z = w - x * BZ, s = z * BX + x;
        for (int bb = 0; bb < 2; bb++) {
            bcg_env_cell& c = a.colony[bb][w];
            const gdouble* d = G(a.scol[bb]);
            c.Carbon = d[FC * NB + s];
            c.Waste = d[FW * NB + s];
            c.Oxygen = d[FO * NB + s];
            c.Acetate = d[FA * NB + s];
        }
    }
    for (int w = gtid(); w < AL * BX; w += gstride()) {
        const int l = w / BX, x = w - l * BX;
        for (int bb = 0; bb < 2; bb++) {
            bcg_aga& q = a.wall[bb][l][x];
            const gdouble* dw = G(a.swall[bb]) + l * RW;
            q.CarbonAgar = dw[FC * BX + x];
            q.WasteAgar = dw[FW * BX + x];
            q.OxygenAgar = dw[FO * BX + x];
            q.AcetateAgar = dw[FA * BX + x];
        }
    }
    if (a.n_outer > 0) {
        for (int l = 0; l < AL; l++)
            for (int w = gtid(); w < BX * a.BZA; w += gstride()) {
                const int i = w / a.BZA, k = w - i * a.BZA + 1;
                const AgaRef s = m.A(l, i, k);
                for (int bb = 0; bb < 2; bb++) {
                    bcg_aga& d = a.agar[bb][l][w];
                    d.CarbonAgar = s.CarbonAgar;
                    d.OxygenAgar = s.OxygenAgar;
                    d.AcetateAgar = s.AcetateAgar;
                }
            }
    }
}

}  // namespace env
}  // namespace bcg
