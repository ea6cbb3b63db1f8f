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

struct Geo {  // one multigrid level as the loops of Multigrid.cpp see it
    int nx, nz, nxp, nzp, d;
    double h2;  // BoxLength^2 d^2
};

// Multigrid arrays: (x extent, z extent) per level from the constructor
// (ABM/Multigrid.cpp:5-33); index x*(nz+1) + z.
struct MG {
    const EnvArgs& a;
    int AL;  // agarLevels = maxLevelsMG
    double dt;
    __device__ MG(const EnvArgs& args, double dt_) : a(args), AL(args.ML), dt(dt_) {}

    __device__ int xext(int l) const {  // x extent - 1
        int nx = a.BX;
        for (int i = AL; i <= l; i++) nx /= 2;
        return nx;
    }
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
    // Element (i, k) with the reference's flat index x*(nz+1) + z: a z past
    // the top of a level reads the next column, as Array3D::Index does. A
    // flat index outside the array (the reference reads past the end of the
    // coarsest error array when BoxZAgar < 2^(8 - maxLevelsMG)) reads zero.
    mutable bcg_aga phantom;
    __device__ bcg_aga& at(bcg_aga* const* arr, int l, int i, int k) const {
        const int X = xext(l), Z = a.mgz[l];
        const long f = (long)i * (Z + 1) + k;
        if (f < 0 || f >= (long)(X + 1) * (Z + 1)) {
            flag(a.errflag, ERR_PHANTOM);
            phantom.CarbonAgar = phantom.WasteAgar = phantom.OxygenAgar = phantom.AcetateAgar = 0.0;
            return phantom;
        }
        return arr[l][f];
    }
    __device__ bcg_aga& A(int l, int i, int k) const { return at(a.aga, l, i, k); }
    __device__ bcg_aga& R(int l, int i, int k) const { return at(a.res, l, i, k); }
    __device__ bcg_aga& E(int l, int i, int k) const { return at(a.err, l, i, k); }
    // insideColonyDen(x, 0, 0); x past the last box reads as 0
    __device__ double inside0(int x) const { return (x >= 0 && x < a.BX) ? a.inside[x * a.BZ] : 0.0; }
};

// (t + (p + q))            relaxation at corners
__device__ __forceinline__ double g2(double t, double p, double q, double den) { return (t + (p + q)) / den; }
// (t + e + 0.5 p + 0.5 q)
__device__ __forceinline__ double g3a(double t, double e, double p, double q, double den) {
    return (t + e + 0.5 * p + 0.5 * q) / den;
}
// (t + 0.5 p + 0.5 q + e)
__device__ __forceinline__ double g3b(double t, double p, double q, double e, double den) {
    return (t + 0.5 * p + 0.5 * q + e) / den;
}
// (t + 0.5 (p + q + r + s)) spelled term by term
__device__ __forceinline__ double g4(double t, double p, double q, double r, double s, double den) {
    return (t + 0.5 * p + 0.5 * q + 0.5 * r + 0.5 * s) / den;
}

// relaxationAgarElementwise (ABM/Multigrid.cpp:161-249).
__device__ void relax_point(const MG& m, int l, int i, int k, int nx, int nz) {
    const EnvArgs& a = m.a;
    const double h2 = a.BoxLength * a.BoxLength * (double)(1 << (2 * l));
    const int d = 1 << l;
    const double dt = m.dt;
    const double DC = a.DiffAgarCarbon, DO = a.DiffAgarO2, DA = a.DiffAgarAce;
    bcg_aga& e = m.E(l, i, k);
    const bcg_aga r = m.R(l, i, k);
    const double tC = r.CarbonAgar * h2 / dt / DC, nC = h2 / dt / DC + 2;
    const double tA = r.AcetateAgar * h2 / dt / DA, nA = h2 / dt / DA + 2;
    const bool outer = l >= m.AL - 1;
    if (k == 0) {
        const int i0 = min(max((i - nx / 2) * d + a.BX / 2, 0), a.BX - 1);
        if (a.inside[i0 * a.BZ] > 0) {
            e.CarbonAgar = r.CarbonAgar;
            e.OxygenAgar = r.OxygenAgar;
            e.AcetateAgar = r.AcetateAgar;
        } else if (i == 0 || i == nx) {
            if (outer) {
                const int in = i == 0 ? i + 1 : i - 1;
                e.CarbonAgar = g2(tC, m.E(l, in, 0).CarbonAgar, m.E(l, i, 1).CarbonAgar, nC);
                e.OxygenAgar = r.OxygenAgar;
                e.AcetateAgar = g2(tA, m.E(l, in, 0).AcetateAgar, m.E(l, i, 1).AcetateAgar, nA);
            }
        } else {
            e.CarbonAgar = g3a(tC, m.E(l, i, 1).CarbonAgar, m.E(l, i - 1, 0).CarbonAgar, m.E(l, i + 1, 0).CarbonAgar, nC);
            e.OxygenAgar = r.OxygenAgar;
            e.AcetateAgar =
                g3a(tA, m.E(l, i, 1).AcetateAgar, m.E(l, i - 1, 0).AcetateAgar, m.E(l, i + 1, 0).AcetateAgar, nA);
        }
    } else if (k == nz) {
        if (outer) {
            if (i == 0 || i == nx) {
                const int in = i == 0 ? i + 1 : i - 1;
                e.CarbonAgar = g2(tC, m.E(l, in, k).CarbonAgar, m.E(l, i, k - 1).CarbonAgar, nC);
                e.OxygenAgar = r.OxygenAgar;
                e.AcetateAgar = g2(tA, m.E(l, in, k).AcetateAgar, m.E(l, i, k - 1).AcetateAgar, nA);
            } else {
                const double tO = r.OxygenAgar * h2 / dt / DO, nO = h2 / dt / DO + 2;
                e.CarbonAgar = g3a(tC, m.E(l, i, k - 1).CarbonAgar, m.E(l, i + 1, k).CarbonAgar,
                                   m.E(l, i - 1, k).CarbonAgar, nC);
                e.OxygenAgar = g3a(tO, m.E(l, i, k - 1).OxygenAgar, m.E(l, i + 1, k).OxygenAgar,
                                   m.E(l, i - 1, k).OxygenAgar, nO);
                e.AcetateAgar = g3a(tA, m.E(l, i, k - 1).AcetateAgar, m.E(l, i + 1, k).AcetateAgar,
                                    m.E(l, i - 1, k).AcetateAgar, nA);
            }
        }
    } else if (i == 0 || i == nx) {
        if (outer) {
            const int in = i == 0 ? i + 1 : i - 1;
            e.CarbonAgar = g3b(tC, m.E(l, i, k - 1).CarbonAgar, m.E(l, i, k + 1).CarbonAgar, m.E(l, in, k).CarbonAgar, nC);
            e.OxygenAgar = r.OxygenAgar;
            e.AcetateAgar =
                g3b(tA, m.E(l, i, k - 1).AcetateAgar, m.E(l, i, k + 1).AcetateAgar, m.E(l, in, k).AcetateAgar, nA);
        }
    } else {
        const double tO = r.OxygenAgar * h2 / dt / DO, nO = h2 / dt / DO + 2;
        e.CarbonAgar = g4(tC, m.E(l, i, k - 1).CarbonAgar, m.E(l, i, k + 1).CarbonAgar, m.E(l, i - 1, k).CarbonAgar,
                          m.E(l, i + 1, k).CarbonAgar, nC);
        e.OxygenAgar = g4(tO, m.E(l, i, k - 1).OxygenAgar, m.E(l, i, k + 1).OxygenAgar, m.E(l, i - 1, k).OxygenAgar,
                          m.E(l, i + 1, k).OxygenAgar, nO);
        e.AcetateAgar = g4(tA, m.E(l, i, k - 1).AcetateAgar, m.E(l, i, k + 1).AcetateAgar,
                           m.E(l, i - 1, k).AcetateAgar, m.E(l, i + 1, k).AcetateAgar, nA);
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
    __device__ static double& v(bcg_aga& g) { return g.CarbonAgar; }
};
struct SpO {
    __device__ static double& v(bcg_aga& g) { return g.OxygenAgar; }
};
struct SpA {
    __device__ static double& v(bcg_aga& g) { return g.AcetateAgar; }
};

// restrictionAgaElementwise (ABM/Multigrid.cpp:270-405).
__device__ void restrict_point(const MG& m, int l, int i, int k, const Geo& g) {
    const EnvArgs& a = m.a;
    const double dt = m.dt, h2 = g.h2;
    const int nx = g.nx, nz = g.nz, nxp = g.nxp, nzp = g.nzp;
    const int i0 = min(max((i - nx / 2) * g.d + a.BX / 2, 0), a.BX);
    const int ip = 2 * i - nx + nxp / 2, kp = 2 * k;
    const int lp = l - 1;
    const double DC = a.DiffAgarCarbon, DO = a.DiffAgarO2, DA = a.DiffAgarAce;
    bcg_aga& out = m.R(l, i, k);
    auto E = [&](int ii, int kk) -> bcg_aga& { return m.E(lp, ii, kk); };
    auto Rp = [&](int ii, int kk) -> bcg_aga& { return m.R(lp, ii, kk); };
    auto Ap = [&](int ii, int kk) -> bcg_aga& { return m.A(lp, ii, kk); };
    auto Al = [&](int ii, int kk) -> bcg_aga& { return m.A(l, ii, kk); };
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
__device__ void interp_point(const MG& m, int l, int i, int k, const Geo& g) {
    bcg_aga& e = m.E(l, i, k);
    const int nx = g.nx, nxp = g.nxp;
    const int ip = nxp / 2 - nx / 4 + (i % 2 == 0 ? i / 2 : (i - 1) / 2);
    auto C = [&](int ii, int kk) -> bcg_aga& { return m.E(l + 1, ii, kk); };
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
__device__ void update_sol_point(const MG& m, int l, int i, int k, const Geo& g) {
    const EnvArgs& a = m.a;
    bcg_aga& r = m.R(l, i, k);
    r.CarbonAgar = 0;
    r.OxygenAgar = 0;
    r.AcetateAgar = 0;
    const int ip = 2 * i - g.nx + g.nxp / 2, kp = 2 * k;
    bcg_aga& e = m.E(l, i, k);
    if (l < m.AL && (l == 0 || ip <= 0 || ip >= g.nxp || kp >= g.nzp)) {
        bcg_aga& u = m.A(l, i, k);
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
__device__ void update_res_point(const MG& m, int l, int i, int k, const Geo& g) {
    const EnvArgs& a = m.a;
    if (l >= m.AL) return;
    const double dt = m.dt, h2 = g.h2;
    const int nx = g.nx, nz = g.nz, nxp = g.nxp, nzp = g.nzp;
    const int ip = 2 * i - nx + nxp / 2, kp = 2 * k;
    const double DC = a.DiffAgarCarbon, DO = a.DiffAgarO2, DA = a.DiffAgarAce;
    bcg_aga& out = m.R(l, i, k);
    auto U = [&](int ii, int kk) -> bcg_aga& { return m.A(l, ii, kk); };
    auto Up = [&](int ii, int kk) -> bcg_aga& { return m.A(l - 1, ii, kk); };
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
__device__ void nest_point(const MG& m, int l, int i, int k, const Geo& g) {
    const int ip = 2 * i - g.nx + g.nxp / 2, kp = 2 * k;
    if (l < m.AL && ip > 0 && ip < g.nxp && kp < g.nzp) {
        bcg_aga& u = m.A(l, i, k);
        const bcg_aga& f = m.A(l - 1, ip, kp);
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

// UpdateEnvironment_OxygenAcetate (ABM/Nutrients.cpp:200-506), non-cross-
// feeding chemistry, one colony box.
__device__ void colony_point(const EnvArgs& a, bcg_env_cell* Env, bcg_env_cell* prev, const bcg_aga* prevWal, int x,
                             int z) {
    const int BZ = a.BZ;
    const int b = x * BZ + z;
    const double ins = a.inside[b];
    auto P = [&](int xx, int zz) -> bcg_env_cell& { return prev[xx * BZ + zz]; };
    auto I = [&](int xx, int zz) { return a.inside[xx * BZ + zz]; };
    bcg_env_cell& out = Env[b];
    if (ins > 0.5) {
        const double C = P(x, z).Carbon, W = P(x, z).Waste, A = P(x, z).Acetate, O = P(x, z).Oxygen;
        (void)acetate_buffer(a, A, P(x, z));
        const double fC = C / (C + a.KC_aer);
        const double fW = W / (W + a.KW);
        (void)fC;
        (void)fW;
        (void)O;
        const double fI = 0;
        double Cgr = 0, Wgr = 0, Agr = 0, Ogr = 0;
        double mgrCarAers, mgrCarAnas, mgrAceAers, g1s, g2s, g3s;
        if (a.maxGrowthRateCarAer != 0.0) {
            mgrCarAers = a.qCarAernt / a.qCarAer;
            g1s = a.KC_aer * (mgrCarAers / a.maxGrowthRateCarAer);
        } else {
            mgrCarAers = 0;
            g1s = 1e6;
        }
        if (a.maxGrowthRateCarAna != 0.0) {
            mgrCarAnas = a.qCarAnant / a.qCarAna;
            g2s = a.KC_ana * (mgrCarAnas / a.maxGrowthRateCarAna);
        } else {
            mgrCarAnas = 0;
            g2s = 1e6;
        }
        if (a.maxGrowthRateAceAer != 0.0) {
            mgrAceAers = a.qAceAernt / a.qAceAer;
            g3s = a.KA * (mgrAceAers / a.maxGrowthRateAceAer);
        } else {
            mgrAceAers = 0;
            g3s = 1e6;
        }
        if (x != 0) {
            // vertex averages of the four boxes (the wall row below z = 0)
            if (z == 0) {
                const bcg_aga &wm = prevWal[x - 1], &w0 = prevWal[x];
                Cgr = (wm.CarbonAgar + P(x - 1, z).Carbon + w0.CarbonAgar + P(x, z).Carbon) / 4.0;
                Wgr = (wm.WasteAgar + P(x - 1, z).Waste + w0.WasteAgar + P(x, z).Waste) / 4.0;
                Ogr = (wm.OxygenAgar + P(x - 1, z).Oxygen + w0.OxygenAgar + P(x, z).Oxygen) / 4.0;
                Agr = (wm.AcetateAgar + P(x - 1, z).Acetate + w0.AcetateAgar + P(x, z).Acetate) / 4.0;
            } else {
                Cgr = (P(x - 1, z - 1).Carbon + P(x - 1, z).Carbon + P(x, z - 1).Carbon + P(x, z).Carbon) / 4.0;
                Wgr = (P(x - 1, z - 1).Waste + P(x - 1, z).Waste + P(x, z - 1).Waste + P(x, z).Waste) / 4.0;
                Ogr = (P(x - 1, z - 1).Oxygen + P(x - 1, z).Oxygen + P(x, z - 1).Oxygen + P(x, z).Oxygen) / 4.0;
                Agr = (P(x - 1, z - 1).Acetate + P(x - 1, z).Acetate + P(x, z - 1).Acetate + P(x, z).Acetate) / 4.0;
            }
        }
        (void)acetate_buffer(a, Agr, P(x, z));
        const double fCgr_aer = Cgr / (Cgr + a.KC_aer);
        const double fCgr_ana = Cgr / (Cgr + a.KC_ana);
        const double fAgr = Agr / (Agr + a.KA);
        const double fOgr = Ogr / (Ogr + a.KO);
        const double fIgr = 0;
        (void)fI;
        (void)Wgr;
        double l1 = smax((a.maxGrowthRateCarAer + mgrCarAers) * fCgr_aer - mgrCarAers, 0.0);
        double l2 = smax((a.maxGrowthRateCarAna + mgrCarAnas) * fCgr_ana - mgrCarAnas, 0.0);
        double l3 = smax((a.maxGrowthRateAceAer + mgrAceAers) * fAgr - mgrAceAers, 0.0);
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
        out.qCarmnt = a.qCarAernt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr + a.qCarAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
        out.qAcemnt = a.qAceAernt * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
        out.pAcemnt = a.pAceAnant * smin((Cgr / g2s), 1.0) * (1 - fOgr);
        out.qOxymnt = a.qO2Carnt * smin((Cgr / g1s), 1.0) * fCgr_aer * fOgr + a.qO2Acent * smin((Agr / g3s), 1.0) * (1 - fCgr_aer) * fOgr;
        const double pW = a.Carbon_to_Waste * qC;
        const double qW = 0.0;
        l1 = smax((a.maxGrowthRateCarAer + mgrCarAers) * fCgr_aer - mgrCarAers, 0.0);
        l2 = smax((a.maxGrowthRateCarAna + mgrCarAnas) * fCgr_ana - mgrCarAnas, 0.0);
        l3 = smax((a.maxGrowthRateAceAer + mgrAceAers) * fAgr - mgrAceAers, 0.0);
        out.GrCarAer = l1 * fCgr_aer * fOgr * inh;
        out.GrCarAna = l2 * (1 - fOgr) * inh;
        out.GrAceAer = l3 * (1 - fCgr_aer) * fOgr * inh;
        out.GrowthRate1 = out.GrCarAer + out.GrCarAna + out.GrAceAer;
        out.GrowthRate2 = out.GrowthRate1;
        if (!(x > 2)) flag(a.errflag, ERR_COLONY_EDGE);
        const double BL2 = a.BoxLength * a.BoxLength;
        double Czz, Wzz, Azz, Ozz;
        const bcg_env_cell &c = P(x, z), &up = P(x, z + 1);
        if (z == 0) {
            const bcg_aga& w = prevWal[x];
            Wzz = (up.Waste + w.WasteAgar - 2.0 * c.Waste) / BL2;
            Czz = (up.Carbon + w.CarbonAgar - 2.0 * c.Carbon) / BL2;
            Azz = (up.Acetate + w.AcetateAgar - 2.0 * c.Acetate) / BL2;
            Ozz = (up.Oxygen + w.OxygenAgar - 2.0 * c.Oxygen) / BL2;
        } else {
            const bcg_env_cell& dn = P(x, z - 1);
            Wzz = (up.Waste + dn.Waste - 2.0 * c.Waste) / BL2;
            Czz = (up.Carbon + dn.Carbon - 2.0 * c.Carbon) / BL2;
            Azz = (up.Acetate + dn.Acetate - 2.0 * c.Acetate) / BL2;
            Ozz = (up.Oxygen + dn.Oxygen - 2.0 * c.Oxygen) / BL2;
        }
        const bcg_env_cell &xp = P(x + 1, z), &xm = P(x - 1, z);
        const double Cxx = (xp.Carbon + xm.Carbon - 2.0 * c.Carbon) / BL2;
        const double Wxx = (xp.Waste + xm.Waste - 2.0 * c.Waste) / BL2;
        const double Axx = (xp.Acetate + xm.Acetate - 2.0 * c.Acetate) / BL2;
        const double Oxx = (xp.Oxygen + xm.Oxygen - 2.0 * c.Oxygen) / BL2;
        const double Dp = a.dshift[b];
        const double Cnew = c.Carbon + (a.DiffColonyCarbon * (Cxx + Czz) - qC * 0.25 * Dp) * a.Cdt;
        const double CAcenew = c.Acetate + (a.DiffColonyAce * (Axx + Azz) + pA * 0.25 * Dp - qA * 0.25 * Dp) * a.Cdt;
        const double Onew = c.Oxygen + (a.DiffColonyO2 * (Oxx + Ozz) - qO * 0.25 * Dp) * a.Cdt;
        out.Carbon = smax(0.0, smin(Cnew, a.maxCarbon));
        out.Oxygen = smax(0.0, smin(Onew, a.maxO2));
        out.Acetate = smax(0.0, CAcenew);
        const double Wnew = c.Waste + (a.DiffColony * (Wxx + Wzz) + pW * Dp - qW * Dp) * a.Cdt;
        out.Waste = smax(0.0, Wnew);
    } else if (ins > 0.3) {
        // colony surface: mean of the neighbouring colony boxes (x, z; at
        // z = 0 without the box below)
        const bool top = z > 0;
        const int ne = (I(x + 1, z) > 0.5) + (I(x - 1, z) > 0.5) + (I(x, z + 1) > 0.5) + (top ? (I(x, z - 1) > 0.5) : 0);
        auto mean = [&](double bcg_env_cell::*f) {
            double s = P(x + 1, z).*f * (I(x + 1, z) > 0.5) + P(x - 1, z).*f * (I(x - 1, z) > 0.5) +
                       P(x, z + 1).*f * (I(x, z + 1) > 0.5);
            if (top) s = s + P(x, z - 1).*f * (I(x, z - 1) > 0.5);
            return s / ne;
        };
        const double Cnew = mean(&bcg_env_cell::Carbon);
        const double Wnew = mean(&bcg_env_cell::Waste);
        const double CAcenew = mean(&bcg_env_cell::Acetate);
        out.Carbon = smax(0.0, smin(Cnew, a.maxCarbon));
        out.Oxygen = smax(0.0, smin(a.maxO2, a.maxO2));
        out.Acetate = smax(0.0, CAcenew);
        out.Waste = smax(0.0, Wnew);
    }
}

// UpdateWallMultigrid_OxygenAcetate (ABM/Nutrients.cpp:509-659) at level l,
// box x: the colony-agar interface row.
__device__ void wall_point(const EnvArgs& a, const bcg_env_cell* Env, bcg_aga* const* prevAga, bcg_aga* const* Wal,
                           bcg_aga* const* prevWal, int l, int x) {
    const int BX = a.BX;
    bcg_aga& out = Wal[l][x];
    if (l == 0 && a.inside[x * a.BZ] > 0.5) {
        const bcg_env_cell& e = Env[x * a.BZ];
        const bcg_aga& g = prevAga[0][x * a.BZA];
        const double Cnew = (a.DiffColonyCarbon * e.Carbon + a.DiffAgarCarbon * g.CarbonAgar) / (a.DiffColonyCarbon + a.DiffAgarCarbon);
        const double Onew = (a.DiffColonyO2 * e.Oxygen + a.DiffAgarO2 * g.OxygenAgar) / (a.DiffColonyO2 + a.DiffAgarO2);
        out.CarbonAgar = smax(0.0, smin(Cnew, a.maxCarbon));
        out.OxygenAgar = smax(0.0, smin(Onew, a.maxO2));
        const double CAcenew = (a.DiffColonyAce * e.Acetate + a.DiffAgarAce * g.AcetateAgar) / (a.DiffColonyAce + a.DiffAgarAce);
        out.AcetateAgar = smax(0.0, CAcenew);
        const double Wnew = (a.DiffColony * e.Waste + a.DiffAgar * g.WasteAgar) / (a.DiffColony + a.DiffAgar);
        out.WasteAgar = smax(0.0, Wnew);
        return;
    }
    const double hl = a.BoxLength * a.BoxLength * (double)(1 << (2 * l));
    const bcg_aga* W = prevWal[l];
    // covered interior of a coarser level: the reference leaves the x term
    // uninitialised there (values overwritten by the next copyFromMG); 0 here
    double Cxx = 0, Wxx = 0, Axx = 0;
    if (l > 0 && x >= BX / 4 - 1 && x <= BX * 3 / 4) {
        const bcg_aga* Wf = prevWal[l - 1];
        if (x == BX / 4 - 1) {
            Wxx = (W[x - 1].WasteAgar + Wf[0].WasteAgar - 2 * W[x].WasteAgar) / hl;
            Cxx = (W[x - 1].CarbonAgar + Wf[0].CarbonAgar - 2 * W[x].CarbonAgar) / hl;
            Axx = (W[x - 1].AcetateAgar + Wf[0].AcetateAgar - 2 * W[x].AcetateAgar) / hl;
        } else if (x == BX * 3 / 4) {
            Wxx = (W[x + 1].WasteAgar + Wf[BX - 2].WasteAgar - 2 * W[x].WasteAgar) / hl;
            Cxx = (W[x + 1].CarbonAgar + Wf[BX - 2].CarbonAgar - 2 * W[x].CarbonAgar) / hl;
            Axx = (W[x + 1].AcetateAgar + Wf[BX - 2].AcetateAgar - 2 * W[x].AcetateAgar) / hl;
        }
    } else if (l < a.ML - 1 && (x == 0 || x == BX - 1)) {
        // coarser level outside: y parity 0 (BoxY = 1), weights 1 and 0
        const int ly = 0;
        const bcg_aga* Wc = prevWal[l + 1];
        const double w1 = 1.0 / (ly + 1), w2 = ly * 1.0 / (ly + 1);
        if (x == 0) {
            const bcg_aga &c1 = Wc[BX / 4 - 1], &c2 = Wc[BX / 4 - 1];
            Wxx = (W[x + 2].WasteAgar + w1 * c1.WasteAgar + w2 * c2.WasteAgar - 2 * W[x].WasteAgar) / 4 / hl;
            Cxx = (W[x + 2].CarbonAgar + w1 * c1.CarbonAgar + w2 * c2.CarbonAgar - 2 * W[x].CarbonAgar) / 4 / hl;
            Axx = (W[x + 2].AcetateAgar + w1 * c1.AcetateAgar + w2 * c2.AcetateAgar - 2 * W[x].AcetateAgar) / 4 / hl;
        } else {
            const bcg_aga &c1 = Wc[BX * 3 / 4], &c2 = Wc[BX * 3 / 4];
            Wxx = (W[x - 1].WasteAgar + w1 * c1.WasteAgar + w2 * c2.WasteAgar - 2 * W[x].WasteAgar) / hl;
            Cxx = (W[x - 1].CarbonAgar + w1 * c1.CarbonAgar + w2 * c2.CarbonAgar - 2 * W[x].CarbonAgar) / hl;
            Axx = (W[x - 1].AcetateAgar + w1 * c1.AcetateAgar + w2 * c2.AcetateAgar - 2 * W[x].AcetateAgar) / hl;
        }
    } else if (l == a.ML - 1 && (x == 0 || x == BX - 1)) {
        const int n = x == 0 ? x + 1 : x - 1;
        Wxx = (W[n].WasteAgar - 1.0 * W[x].WasteAgar) / hl;
        Cxx = (W[n].CarbonAgar - 1.0 * W[x].CarbonAgar) / hl;
        Axx = (W[n].AcetateAgar - 1.0 * W[x].AcetateAgar) / hl;
    } else {
        Wxx = (W[x - 1].WasteAgar + W[x + 1].WasteAgar - 2 * W[x].WasteAgar) / hl;
        Cxx = (W[x - 1].CarbonAgar + W[x + 1].CarbonAgar - 2 * W[x].CarbonAgar) / hl;
        Axx = (W[x - 1].AcetateAgar + W[x + 1].AcetateAgar - 2 * W[x].AcetateAgar) / hl;
    }
    const bcg_aga& g = prevAga[l][x * a.BZA];
    const double Wzz = 2.0 * (g.WasteAgar - W[x].WasteAgar) / hl;
    const double Czz = 2.0 * (g.CarbonAgar - W[x].CarbonAgar) / hl;
    const double Azz = 2.0 * (g.AcetateAgar - W[x].AcetateAgar) / hl;
    const double Cnew = W[x].CarbonAgar + a.DiffAgarCarbon * (Cxx + Czz) * a.Cdt;
    const double Onew = a.maxO2;
    out.CarbonAgar = smax(0.0, smin(Cnew, a.maxCarbon));
    out.OxygenAgar = smax(0.0, smin(Onew, a.maxO2));
    const double CAcenew = W[x].AcetateAgar + a.DiffAgarAce * (Axx + Azz) * a.Cdt;
    out.AcetateAgar = smax(0.0, CAcenew);
    const double Wnew = W[x].WasteAgar + a.DiffAgar * (Wxx + Wzz) * a.Cdt;
    out.WasteAgar = smax(0.0, Wnew);
}

}  // namespace

// The whole UpdateEnvArrayMG_OxygenAcetate in one workgroup.
__global__ void __launch_bounds__(ENV_THREADS) k_env_nutrients(EnvArgs args) {
    // the pointer tables are indexed by level and buffer at run time: keep
    // them in LDS rather than in per-lane scratch
    __shared__ EnvArgs a;
    const int tid = threadIdx.x;
    if (tid == 0) a = args;
    __syncthreads();
    const MG m(a, 100.0);  // dt = 100 (ABM/Multigrid.cpp:993)
    const int AL = a.ML;
    const int BX = a.BX;
    auto sync = [] { __syncthreads(); };
    // level geometry for the loops (sub 1: restriction/update; 2: interpolation)
    auto G1 = [&](int l) { return m.geo(l, 1); };
    auto G2 = [&](int l) { return m.geo(l, 2); };
    // one colour pass of relaxationAgarAtLevel: columns of parity `par`
    auto relax = [&](int l) {
        const Geo g = G1(l);
        for (int par = 0; par < 2; par++) {
            for (int c = tid; 2 * c + par <= g.nx; c += ENV_THREADS) {
                const int i = 2 * c + par;
                for (int k = 0; k <= g.nz; k++) relax_point(m, l, i, k, g.nx, g.nz);
            }
            sync();
        }
    };
    auto over = [&](const Geo& g, auto fn) {
        const int n = (g.nx + 1) * (g.nz + 1);
        for (int w = tid; w < n; w += ENV_THREADS) fn(w / (g.nz + 1), w % (g.nz + 1));
    };
    auto reset_zero = [&]() {
        for (int l = 0; l < MG_LEVELS; l++) {
            const Geo g = G1(l);
            over(g, [&](int i, int k) {
                bcg_aga& e = m.E(l, i, k);
                e.CarbonAgar = 0;
                e.OxygenAgar = 0;
                e.AcetateAgar = 0;
                if (l > 0) {
                    bcg_aga& r = m.R(l, i, k);
                    r.CarbonAgar = 0;
                    r.OxygenAgar = 0;
                    r.AcetateAgar = 0;
                }
            });
        }
        sync();
    };
    auto update_sol = [&]() {
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = G1(l);
            over(g, [&](int i, int k) { update_sol_point(m, l, i, k, g); });
        }
        sync();
    };
    auto update_res = [&]() {
        for (int l = 0; l <= AL - 1; l++) {
            const Geo g = G1(l);
            over(g, [&](int i, int k) { update_res_point(m, l, i, k, g); });
        }
        sync();
    };

    int count = 0;
    int sw = 0;  // 0: Current = Environment / FieldAgar / FieldWall
    for (int outer = 0; outer < a.n_outer; outer++) {
        bcg_aga* const* curAgar = a.agar[sw];
        bcg_aga* const* prevAgar = a.agar[sw ^ 1];
        bcg_aga* const* curWall = a.wall[sw];
        bcg_aga* const* prevWall = a.wall[sw ^ 1];
        // --- Multigrid::solve(CurrentAgar, CurrentWall, PreviousAgar, PreviousWall, 3, 100, 100)
        // copyToMG: colony footprint from the wall, and the last column of the
        // outermost nested level from its first
        for (int w = tid; w < BX - 1; w += ENV_THREADS) {
            const int i = w + 1;
            if (a.inside[i * a.BZ] > 0) {
                bcg_aga& u = m.A(0, i, 0);
                const bcg_aga& s = curWall[0][i];
                u.CarbonAgar = s.CarbonAgar;
                u.OxygenAgar = s.OxygenAgar;
                u.AcetateAgar = s.AcetateAgar;
            }
        }
        for (int k = tid; k <= a.BZA; k += ENV_THREADS) {
            bcg_aga& u = m.A(AL - 1, BX, k);
            const bcg_aga& s = m.A(AL - 1, 0, k);
            u.CarbonAgar = s.CarbonAgar;
            u.OxygenAgar = s.OxygenAgar;
            u.AcetateAgar = s.AcetateAgar;
        }
        sync();
        // setup
        reset_zero();
        update_sol();
        update_res();
        for (int cyc = 0; cyc < 3; cyc++) {
            for (int l = 0; l < MG_LEVELS - 1; l++) {
                for (int j = 0; j < 3; j++) relax(l);
                const Geo g = G1(l + 1);
                over(g, [&](int i, int k) { restrict_point(m, l + 1, i, k, g); });
                sync();
            }
            for (int j = 0; j < 100; j++) relax(MG_LEVELS - 1);
            for (int l = MG_LEVELS - 2; l >= 0; l--) {
                const Geo g = G2(l);
                over(g, [&](int i, int k) { interp_point(m, l, i, k, g); });
                sync();
                for (int j = 0; j < 3; j++) relax(l);
            }
        }
        // reset
        update_sol();
        for (int l = 0; l < AL - 1; l++) {
            const Geo g = G2(l);
            const int nx = g.nx, nz = g.nz, nx2 = g.nxp;
            for (int i = tid; i <= nx; i += ENV_THREADS) {
                const int i2 = nx2 / 2 + i / 2 - nx / 4, i2p1 = nx2 / 2 + (i + 1) / 2 - nx / 4;
                bcg_aga& u = m.A(l, i, nz);
                const bcg_aga &p = m.A(l + 1, i2, nz / 2), &q = m.A(l + 1, i2p1, nz / 2);
                u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
                u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
                u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
            }
            for (int k = tid; k < nz; k += ENV_THREADS) {
                for (int side = 0; side < 2; side++) {
                    const int i = side ? nx : 0;
                    const int i2 = nx2 / 2 + i / 2 - nx / 4;
                    bcg_aga& u = m.A(l, i, k);
                    const bcg_aga &p = m.A(l + 1, i2, k / 2), &q = m.A(l + 1, i2, (k + 1) / 2);
                    u.CarbonAgar = (p.CarbonAgar + q.CarbonAgar) / 2;
                    u.OxygenAgar = (p.OxygenAgar + q.OxygenAgar) / 2;
                    u.AcetateAgar = (p.AcetateAgar + q.AcetateAgar) / 2;
                }
            }
            sync();
        }
        update_res();
        reset_zero();
        for (int l = 1; l <= AL; l++) {
            const Geo g = G1(l);
            over(g, [&](int i, int k) { nest_point(m, l, i, k, g); });
            sync();
        }
        // copyFromMG into both buffers of the nested levels
        for (int l = 0; l < AL; l++) {
            for (int w = tid; w < BX * (a.BZA + 1); w += ENV_THREADS) {
                const int i = w / (a.BZA + 1), k = w % (a.BZA + 1);
                const bcg_aga& s = m.A(l, i, k);
                bcg_aga& d1 = k == 0 ? curWall[l][i] : curAgar[l][i * a.BZA + k - 1];
                bcg_aga& d2 = k == 0 ? prevWall[l][i] : prevAgar[l][i * a.BZA + k - 1];
                d1.CarbonAgar = s.CarbonAgar;
                d1.OxygenAgar = s.OxygenAgar;
                d1.AcetateAgar = s.AcetateAgar;
                d2.CarbonAgar = s.CarbonAgar;
                d2.OxygenAgar = s.OxygenAgar;
                d2.AcetateAgar = s.AcetateAgar;
            }
        }
        sync();

        // --- 200 Jacobi sweeps, direction alternating with the sweep count
        const int minX = a.bounds[0], maxX = a.bounds[1], maxH = a.bounds[4];
        const int ncol = maxX - minX + 1, nrow = maxH + 1;
        for (int j = 0; j < 200; j++) {
            const int dir = (count % 2 == 0) ? 1 : -1;
            bcg_env_cell* cur = a.colony[sw];
            bcg_env_cell* prev = a.colony[sw ^ 1];
            bcg_aga* const* pA = a.agar[sw ^ 1];
            bcg_aga* const* cW = a.wall[sw];
            bcg_aga* const* pW = a.wall[sw ^ 1];
            auto wall_phase = [&]() {
                for (int w = tid; w < a.ML * BX; w += ENV_THREADS) wall_point(a, cur, pA, cW, pW, w / BX, w % BX);
                sync();
            };
            auto colony_phase = [&]() {
                for (int w = tid; w < ncol * nrow; w += ENV_THREADS) {
                    const int x = minX + w / nrow, z = w % nrow;
                    if (a.dshift[x * a.BZ + z] > DBL_EPSILON) colony_point(a, cur, prev, pW[0], x, z);
                }
                sync();
            };
            // single-threaded order: z ascending starts with the interface
            // (which then reads the colony row of two sweeps ago), z
            // descending ends with it
            if (dir == 1) {
                wall_phase();
                colony_phase();
            } else {
                colony_phase();
                wall_phase();
            }
            sw ^= 1;
            count++;
        }
    }
}

}  // namespace env
}  // namespace bcg
