// mg_api.hip - host side of the 3-D nutrient solver (bcg_rd3d_* in include/bcg.h).
//
// Device state of one grid: the nutrient field c (FP64, ping-pong for the
// refinement update), the previous field, the uptake density q, the FP64
// right-hand side and coefficient a0 = 1 + dt q/(K + c_prev), and per
// multigrid level FP64 and FP32 copies of u, f, a0 and a residual buffer.
//
// Two solvers of the same implicit step, both to a relative residual tol:
//   mode 0  FP64 multigrid: V(nu1,nu2) cycles on A c = b, FP64 everywhere.
//   mode 1  mixed: FP64 iterative refinement; each outer step forms the
//           residual in FP64 (fused with the update), scales it into FP32
//           and runs `inner` FP32 V-cycles (red-black smoother and
//           transfers in FP32, coarsest level solved in FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bcg.h"
#include "ctx_internal.h"
#include "mech_kernels.h"
#include "mg_kernels.h"
#include "rbgs_fused.h"
#include "zmarch.h"

using namespace bcg;
using namespace bcg::mg;

namespace {

struct Level {
    int nx, ny, nz;
    int cx = 0, cy = 0, cz = 0;  // coarsening to the next level
    double hx, hy, hz;
    double lx, ly, lz;
    long n() const { return (long)nx * ny * nz; }
    double *u64 = nullptr, *f64 = nullptr, *a064 = nullptr, *r64 = nullptr;
    float *u32 = nullptr, *f32 = nullptr, *a032 = nullptr, *r32 = nullptr;
};

// Deterministic cloud-in-cell deposition: cells sorted by base vertex, each
// vertex gathers from its 8 base-vertex buckets in a fixed order.
// Sources given as centres (x, y, z per source) with their rates in `rate`.
struct PointSources {
    const double* cm;
    __device__ void centre(int c, double& x, double& y, double& z, double*) const {
        x = cm[3 * c];
        y = cm[3 * c + 1];
        z = cm[3 * c + 2];
    }
};

// Sources of a mechanics context's cells: centre (p + q) / 2 and rate
// `rate` x Length (hot record: p [0..2], q [3..5], Length [12]), written to
// `out` for the gather.
struct CellSources {
    const double* hot;
    double rate;
    __device__ void centre(int c, double& x, double& y, double& z, double* out) const {
        const double* h = hot + (size_t)c * bcg::HOT;
        x = (h[0] + h[3]) * 0.5;
        y = (h[1] + h[4]) * 0.5;
        z = (h[2] + h[5]) * 0.5;
        out[c] = rate * h[12];
    }
};

template <class Src>
__global__ void k_dep_key(Src src, int n, double ox, double oy, double oz, double inv_h, int nx, int ny, int nz,
                          int* __restrict__ key, double* __restrict__ frac, int* __restrict__ cnt,
                          double* __restrict__ rate_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double x, y, z;
    src.centre(c, x, y, z, rate_out);
    const double gx = (x - ox) * inv_h, gy = (y - oy) * inv_h, gz = (z - oz) * inv_h;
    const double fx = floor(gx), fy = floor(gy), fz = floor(gz);
    int k = -1;
    if (fx >= 0 && fx < nx - 1 && fy >= 0 && fy < ny - 1 && fz >= 0 && fz < nz - 1) {
        k = (int)(((long)fz * ny + (long)fy) * nx + (long)fx);
        frac[3 * c] = gx - fx;
        frac[3 * c + 1] = gy - fy;
        frac[3 * c + 2] = gz - fz;
        atomicAdd(&cnt[k], 1);
    }
    key[c] = k;
}

__global__ void k_dep_scatter(const int* __restrict__ key, int n, const int* __restrict__ start, int* __restrict__ fill,
                              int* __restrict__ sorted) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int k = key[c];
    if (k < 0) return;
    sorted[start[k] + atomicAdd(&fill[k], 1)] = c;
}

// The same placement from the bucket counts themselves, counted back down
// to zero (the next deposition needs no clearing pass).
__global__ void k_dep_scatter_dec(const int* __restrict__ key, int n, const int* __restrict__ start,
                                  int* __restrict__ cnt, int* __restrict__ sorted) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int k = key[c];
    if (k < 0) return;
    sorted[start[k] + atomicSub(&cnt[k], 1) - 1] = c;
}

__global__ void k_dep_sort(const int* __restrict__ start, long nn, int* __restrict__ sorted) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nn) return;
    const int s = start[b], e = start[b + 1];
    for (int i = s + 1; i < e; i++) {
        const int v = sorted[i];
        int j = i - 1;
        while (j >= s && sorted[j] > v) {
            sorted[j + 1] = sorted[j];
            j--;
        }
        sorted[j + 1] = v;
    }
}

// One vertex per thread: the sum over its 8 base-vertex buckets (planes k,
// then k-1; rows j, then j-1; columns i, then i-1; each bucket's cells in
// ascending order). The 8 bucket ranges are read first, all loads in flight
// at once (three starts a row serve both columns); a plane of buckets with
// no cell (its first and last start equal) is skipped as a whole, as a
// colony fills only the lowest planes of its grid. (A column walk that
// reuses each plane's bucket starts was slower.)
__global__ void __launch_bounds__(256) k_dep_gather(const int* __restrict__ start, const int* __restrict__ sorted,
                                                    const double* __restrict__ frac, const double* __restrict__ rate,
                                                    int nx, int ny, int nz, double inv_vol, double* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const long nxy = (long)nx * ny;
    int sb[8], eb[8];  // bucket ranges, index dz * 4 + dy * 2 + dx
#pragma unroll
    for (int dz = 0; dz <= 1; dz++) {
        const int bk = k - dz;
        bool zok = bk >= 0 && bk < nz - 1;
        if (zok) zok = start[bk * nxy] != start[(bk + 1) * nxy];
#pragma unroll
        for (int dy = 0; dy <= 1; dy++) {
            const int bj = j - dy;
            const bool yok = zok && bj >= 0 && bj < ny - 1;
            const bool x0 = yok && i < nx - 1;  // bucket column i
            const bool x1 = yok && i >= 1;      // bucket column i - 1
            const long b = ((long)bk * ny + bj) * nx + i;
            const int sm = x1 ? start[b - 1] : 0;
            const int s0 = (x0 || x1) ? start[b] : 0;
            const int sp = x0 ? start[b + 1] : 0;
            const int m = dz * 4 + dy * 2;
            sb[m] = s0;
            eb[m] = x0 ? sp : s0;
            sb[m + 1] = sm;
            eb[m + 1] = x1 ? s0 : sm;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int dz = m >> 2, dy = (m >> 1) & 1, dx = m & 1;
        for (int s = sb[m]; s < eb[m]; s++) {
            const int c = sorted[s];
            const double fx = frac[3 * c], fy = frac[3 * c + 1], fz = frac[3 * c + 2];
            const double wx = dx ? fx : 1.0 - fx, wy = dy ? fy : 1.0 - fy, wz = dz ? fz : 1.0 - fz;
            acc += rate[c] * (wx * wy * wz);
        }
    }
    q[((long)k * ny + j) * nx + i] = acc * inv_vol;
}

// Growth rate of every mechanics bin (x * BoxY * BoxZ + y * BoxZ + z) from
// the field at the bin centre: mu c / (K + c), c trilinear in the grid,
// positions clamped to it. Bin centres follow the bin address rule
// (x = int(x/BL + BoxX/2), y = int(y/BL + BoxY/2 + 0.5), z = int(z/BL)).
__global__ void k_growth_bins(const double* __restrict__ x, int nx, int ny, int nz, double ox, double oy, double oz,
                              double inv_h, int BoxX, int BoxY, int BoxZ, double BL, double mu, double K,
                              double* __restrict__ gr) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (long)BoxX * BoxY * BoxZ) return;
    const int xb = (int)(b / ((long)BoxY * BoxZ)), yb = (int)((b / BoxZ) % BoxY), zb = (int)(b % BoxZ);
    const double px = (xb - BoxX / 2 + 0.5) * BL, py = (yb - BoxY / 2) * BL, pz = (zb + 0.5) * BL;
    auto axis = [](double g, int n, int& i0, double& w) {
        g = g < 0.0 ? 0.0 : (g > n - 1 ? n - 1 : g);
        i0 = (int)g;
        if (i0 > n - 2) i0 = n - 2;
        w = g - i0;
    };
    int i, j, k;
    double wx, wy, wz;
    axis((px - ox) * inv_h, nx, i, wx);
    axis((py - oy) * inv_h, ny, j, wy);
    axis((pz - oz) * inv_h, nz, k, wz);
    const long sxy = (long)nx * ny;
    auto at = [&](int di, int dj, int dk) { return x[(k + dk) * sxy + (long)(j + dj) * nx + (i + di)]; };
    const double c00 = at(0, 0, 0) * (1 - wx) + at(1, 0, 0) * wx, c10 = at(0, 1, 0) * (1 - wx) + at(1, 1, 0) * wx;
    const double c01 = at(0, 0, 1) * (1 - wx) + at(1, 0, 1) * wx, c11 = at(0, 1, 1) * (1 - wx) + at(1, 1, 1) * wx;
    double c = (c00 * (1 - wy) + c10 * wy) * (1 - wz) + (c01 * (1 - wy) + c11 * wy) * wz;
    c = c > 0.0 ? c : 0.0;
    gr[b] = mu * c / (K + c);
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

struct bcg_rd3d {
    bcg_rd3d_params prm;
    hipStream_t stream = nullptr;
    std::vector<Level> L;
    long N = 0;
    double *x[2] = {nullptr, nullptr};  // field ping-pong
    int cur = 0;
    double *cprev = nullptr, *b = nullptr, *q = nullptr, *a0 = nullptr, *save = nullptr, *grad = nullptr;
    double *gpart = nullptr, *dout = nullptr, *dfield = nullptr;
    float* a0f = nullptr;
    double* part = nullptr;  // per-block partial sums
    long npart = 0;
    double* part2 = nullptr;  // second-stage partial sums
    int bottom = 0;           // first level handled by the single-workgroup bottom solver
    double *bot_u0 = nullptr, *bot_f0 = nullptr;  // FP64 level-0 copies when that level is the bottom
    double* dnorm = nullptr;
    double* hnorm = nullptr;  // pinned
    // deposition scratch
    int *dkey = nullptr, *dcnt = nullptr, *dstart = nullptr, *dsorted = nullptr, *scan_tot = nullptr;
    bool dcnt_zero = false;  // every bucket count is zero (set after a completed deposition)
    double *dfrac = nullptr, *dcm = nullptr, *drate = nullptr;
    int dcap = 0;
    hipEvent_t ev[5];  // [0..3] timing, [4] ordering against a mechanics context
    float last_ms[3] = {0, 0, 0};
    // workgroups resident on the device at once, for the z-range counts of
    // the z-marching launches: k_rbgs2 (FP64, FP32 narrow, FP32 wide tile),
    // the fine-level passes (residual, residual norm, refinement update),
    // k_zm_prolong and k_zm_restrict (FP64, FP32)
    long res_rb[4] = {1, 1, 1, 1}, res_zm = 1, res_zp[2] = {1, 1}, res_zr[2] = {1, 1};
    // bcg_rd3d_update_cells: its fixed-shape part (set-up, V-cycles, growth
    // rates) recorded as a HIP graph for one (mechanics context, V-cycle
    // count) and replayed
    bool a32_valid = false;  // FP32 copies of a0 match the FP64 ones
    hipGraphExec_t upd_graph = nullptr;
    const bcg_ctx* upd_cells = nullptr;
    int upd_vcycles = 0, upd_cur = -1;
    double upd_mu = 0, upd_K = 0;
    std::string err;
};

namespace {

int fail(bcg_rd3d* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(c, x)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(c, BCG_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(bcg_rd3d* c, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(c, BCG_ERR_HIP, "hipMalloc(%zu): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

dim3 grid2(const Level& v, int bx = 64, int by = 4) { return dim3(nblk(v.nx, bx), nblk(v.ny, by), v.nz); }
// 1-D launches over 64 x 4 x 1 tiles, remapped in the kernel (slab_tile)
dim3 slab_grid(int nx, int ny, int nz) { return dim3((unsigned)(nblk(nx, 64) * nblk(ny, 4) * nz)); }
dim3 slab_grid(const Level& v) { return slab_grid(v.nx, v.ny, v.nz); }
dim3 grid_rb(const Level& v) { return slab_grid((v.nx + 1) / 2, v.ny, v.nz); }

template <class T>
struct Arr;
template <>
struct Arr<double> {
    static double* u(Level& v) { return v.u64; }
    static double* f(Level& v) { return v.f64; }
    static double* a0(Level& v) { return v.a064; }
    static double* r(Level& v) { return v.r64; }
};
template <>
struct Arr<float> {
    static float* u(Level& v) { return v.u32; }
    static float* f(Level& v) { return v.f32; }
    static float* a0(Level& v) { return v.a032; }
    static float* r(Level& v) { return v.r32; }
};

// Levels of at most this many points go to the single-workgroup bottom
// solver: one barrier-separated phase there (~1-3 us) is cheaper than one
// launch of a small kernel. (Measured: taking 33 x 33 x 9 in as well, worked
// on from memory as it does not fit the LDS, made the 257 x 257 x 65 cycle
// slower, 0.60 against 0.58 ms for the coupled nutrient update.)
constexpr long BOTTOM_MAX_POINTS = 4096;

// Levels of at least this many points use the fused z-marching sweeps;
// below it the two-launch sweep has more parallelism per launch.
constexpr long FUSED_MIN_POINTS = 1L << 21;
constexpr long WIDE_MIN_POINTS = 1L << 24;  // FP32 levels that get the 62 x 30 sweep tile
// Levels of at most this many points take the prolongation inside their
// first red sweep (measured: 5.3 us against 4.7 + 4.6 us for the two
// launches on 65 x 65 x 17 and 33 x 33 x 9; 18 us against 5 + 6 us on
// 129 x 129 x 33, where the six interpolations a red point repeats cost more
// than the launch).
constexpr long PROLONG_FUSED_MAX_POINTS = 1L << 17;

// Blocks of the ||b||^2 reduction (k_sumsq): one partial sum each in c->part,
// which holds at least this many.
constexpr int SUMSQ_BLOCKS = 4096;

// Workgroups of `kernel` (block size `threads`) resident on the device at once.
template <class K>
long resident(K kernel, int threads) {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess) return 1;
    return std::max(1L, (long)nb * cus);
}

// `sweeps` red-black sweeps on u of level v. Fused sweeps go in pairs
// u -> r -> u (the residual buffer is free while smoothing); an odd last
// sweep, and all sweeps of small levels, use one launch per colour. With
// `zero`, u starts at zero: the first fused sweep reads no iterate (u need
// not be cleared); otherwise u is cleared first.
template <class T>
void smooth(bcg_rd3d* c, Level& v, int sweeps, bool zero = false) {
    T* u = Arr<T>::u(v);
    int s = 0;
    if (c->prm.smoother != BCG_RD3D_SMOOTH_SPLIT && v.n() >= FUSED_MIN_POINTS) {
        T* tmp = Arr<T>::r(v);
        // plane loads two steps ahead (measured ~4% faster than one step ahead on
        // the 513 x 513 x 129 level in both precisions)
        auto sweeps2 = [&](auto tile, long resident) {
            using D = decltype(tile);
            const RbGrid g = rb_grid<D>(v.nx, v.ny, v.nz, resident);
            const dim3 grid(g.tiles_x * g.tiles_y * g.nchunk);
            auto sweep = [&](const T* in, T* out) {
                hipLaunchKernelGGL((k_rbgs2<T, 2, D>), grid, dim3(D::THREADS), 0, c->stream, in, out, Arr<T>::f(v),
                                   Arr<T>::a0(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz, g.tiles_x, g.tiles_y,
                                   g.kchunk);
            };
            for (; s + 2 <= sweeps; s += 2) {
                sweep(zero && s == 0 ? nullptr : u, tmp);
                sweep(tmp, u);
            }
        };
        // the wide tile pays where there are enough tiles to fill the device
        // (measured on 513 x 513 x 129: FP32 143 vs 166 us a sweep; FP64, with
        // 87 KB of LDS and one workgroup a CU, 264 vs 280 us - taken only when
        // asked for, BCG_RD3D_SMOOTH_FUSED_WIDE; on 257 x 257 x 65 the narrow
        // tile is faster)
        const bool wide = v.n() >= WIDE_MIN_POINTS && (sizeof(T) == 4 || c->prm.smoother == BCG_RD3D_SMOOTH_FUSED_WIDE);
        if (wide) sweeps2(RbWide{}, c->res_rb[sizeof(T) == 4 ? 2 : 3]);
        else sweeps2(RbNarrow{}, c->res_rb[sizeof(T) == 4]);
    }
    if (zero && s == 0)
        hipLaunchKernelGGL((k_fill<T>), dim3(nblk(v.n(), 256)), dim3(256), 0, c->stream, u, v.n(), (T)0);
    for (; s < sweeps; s++)
        for (int color = 0; color < 2; color++)
            hipLaunchKernelGGL((k_rbgs<T>), grid_rb(v), dim3(64, 4), 0, c->stream, u, Arr<T>::f(v), Arr<T>::a0(v),
                               v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz, color);
}

// Split sweeps on a small level (one launch per colour), the first red sweep
// as `first` says: a plain sweep, from u = 0 (k_rbgs_zero), or fused with
// the prolongation from level w (k_rbgs_prolong). With no sweeps, u is
// cleared or the correction added, as the fused forms would have done.
enum class FirstRed { plain, zero, prolong };
template <class T>
void smooth_small(bcg_rd3d* c, Level& v, int sweeps, FirstRed first, Level* w = nullptr) {
    T* u = Arr<T>::u(v);
    if (sweeps == 0) {
        if (first == FirstRed::zero)
            hipLaunchKernelGGL((k_fill<T>), dim3(nblk(v.n(), 256)), dim3(256), 0, c->stream, u, v.n(), (T)0);
        else if (first == FirstRed::prolong)
            hipLaunchKernelGGL((k_prolong<T>), slab_grid(v), dim3(64, 4), 0, c->stream, u, Arr<T>::u(*w), v.nx, v.ny,
                               v.nz, w->nx, w->ny, v.cx, v.cy, v.cz);
        return;
    }
    for (int s = 0; s < sweeps; s++)
        for (int color = 0; color < 2; color++) {
            if (s == 0 && color == 0 && first == FirstRed::zero)
                hipLaunchKernelGGL((k_rbgs_zero<T>), grid_rb(v), dim3(64, 4), 0, c->stream, u, Arr<T>::f(v),
                                   Arr<T>::a0(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz);
            else if (s == 0 && color == 0 && first == FirstRed::prolong)
                hipLaunchKernelGGL((k_rbgs_prolong<T>), grid_rb(v), dim3(64, 4), 0, c->stream, u, Arr<T>::f(v),
                                   Arr<T>::a0(v), Arr<T>::u(*w), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz, w->nx,
                                   w->ny, v.cx, v.cy, v.cz);
            else
                hipLaunchKernelGGL((k_rbgs<T>), grid_rb(v), dim3(64, 4), 0, c->stream, u, Arr<T>::f(v), Arr<T>::a0(v),
                                   v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz, color);
        }
}

// The bottom levels by the single-workgroup FP64 solver, with level 0 on
// (u0, f0), or, in an FP32 cycle, read from and written back to (u32, f32)
// with (u0, f0) as working copies; `zero`: from u = 0.
void bottom_solve(bcg_rd3d* c, double* u0, double* f0, const float* f32, float* u32, bool zero) {
    BotLevels B;
    B.n = (int)c->L.size() - c->bottom;
    B.f32 = f32;
    B.u32 = u32;
    B.zero = zero ? 1 : 0;
    for (int i = 0; i < B.n; i++) {
        Level& v = c->L[c->bottom + i];
        B.L[i] = BotLevel{v.u64, v.f64, v.a064, v.r64, v.nx, v.ny, v.nz, v.cx, v.cy, v.cz, v.lx, v.ly, v.lz};
    }
    B.L[0].u = u0;
    B.L[0].f = f0;
    // the coarsest levels that fit go to LDS (u, f, a0 each + one residual
    // buffer of the largest of them)
    int lds_from = B.n;
    long pts = 0, bytes = 0;
    for (int i = B.n - 1; i >= 0; i--) {
        const long n = (long)B.L[i].nx * B.L[i].ny * B.L[i].nz;
        const long b = (3 * (pts + n) + n) * (long)sizeof(double);  // level i is the largest so far
        if (b > BOT_LDS_BYTES) break;
        pts += n;
        bytes = b;
        lds_from = i;
    }
    hipLaunchKernelGGL(k_bottom, dim3(1), dim3(1024), bytes, c->stream, B, c->prm.nu1, c->prm.nu2,
                       c->prm.coarse_sweeps, lds_from);
}

// One V-cycle at level l in precision T on (u_l, f_l), from u_l = 0 when
// `zero`; the levels from c->bottom down are one FP64 single-workgroup launch
// (on FP64 copies of u and f when T is float).
template <class T>
void vcycle(bcg_rd3d* c, int l, bool zero = false) {
    Level& v = c->L[l];
    if (l == c->bottom) {
        if (sizeof(T) == 4)
            bottom_solve(c, l ? v.u64 : c->bot_u0, l ? v.f64 : c->bot_f0, v.f32, v.u32, zero);
        else
            bottom_solve(c, v.u64, v.f64, nullptr, nullptr, zero);
        return;
    }
    Level& w = c->L[l + 1];
    if (c->prm.smoother != BCG_RD3D_SMOOTH_SPLIT && v.n() < FUSED_MIN_POINTS) {
        // small level: no clearing pass before the first sweep and, where the
        // level is small enough for its redundant interpolations to cost less
        // than a launch, the prolongation inside the first red sweep; the
        // same values as below
        smooth_small<T>(c, v, c->prm.nu1, zero ? FirstRed::zero : FirstRed::plain);
        hipLaunchKernelGGL((k_residual<T>), slab_grid(v), dim3(64, 4), 0, c->stream, Arr<T>::u(v), Arr<T>::f(v),
                           Arr<T>::a0(v), Arr<T>::r(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz);
        hipLaunchKernelGGL((k_restrict<T>), slab_grid(w), dim3(64, 4), 0, c->stream, Arr<T>::r(v), Arr<T>::f(w), v.nx,
                           v.ny, v.nz, w.nx, w.ny, w.nz, v.cx, v.cy, v.cz);
        vcycle<T>(c, l + 1, true);
        if (v.n() <= PROLONG_FUSED_MAX_POINTS) {
            smooth_small<T>(c, v, c->prm.nu2, FirstRed::prolong, &w);
        } else {
            hipLaunchKernelGGL((k_prolong<T>), slab_grid(v), dim3(64, 4), 0, c->stream, Arr<T>::u(v), Arr<T>::u(w),
                               v.nx, v.ny, v.nz, w.nx, w.ny, v.cx, v.cy, v.cz);
            smooth_small<T>(c, v, c->prm.nu2, FirstRed::plain);
        }
        return;
    }
    smooth<T>(c, v, c->prm.nu1, zero);
    // z-marching residual on the large levels
    if (v.n() >= FUSED_MIN_POINTS) {
        const ZmGrid g = zm_grid(v.nx, v.ny, v.nz, c->res_zm);
        if constexpr (sizeof(T) == 4)  // FP32: two planes a step (half the bytes a point of FP64)
            hipLaunchKernelGGL((k_zm_residual2<T>), dim3(g.blocks()), dim3(ZM_THREADS), 0, c->stream, Arr<T>::u(v),
                               Arr<T>::f(v), Arr<T>::a0(v), Arr<T>::r(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly,
                               (T)v.lz, g.tiles_x, g.tiles_y, g.kchunk);
        else
            hipLaunchKernelGGL((k_zm_residual<T>), dim3(g.blocks()), dim3(ZM_THREADS), 0, c->stream, Arr<T>::u(v),
                               Arr<T>::f(v), Arr<T>::a0(v), Arr<T>::r(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly,
                               (T)v.lz, g.tiles_x, g.tiles_y, g.kchunk);
    } else {
        hipLaunchKernelGGL((k_residual<T>), slab_grid(v), dim3(64, 4), 0, c->stream, Arr<T>::u(v), Arr<T>::f(v),
                           Arr<T>::a0(v), Arr<T>::r(v), v.nx, v.ny, v.nz, (T)v.lx, (T)v.ly, (T)v.lz);
    }
    if (v.n() >= WIDE_MIN_POINTS) {
        // z-marching: each fine plane's x-y sum serves both coarse planes it
        // touches (measured on the finest level only: 65 vs 78 us FP32, 100 vs
        // 104 us FP64 on 513 x 513 x 129; slower than the plain kernel below)
        const long cols = (long)nblk(w.nx, 64) * nblk(w.ny, 4);
        const int nc = z_ranges(cols, w.nz, c->res_zr[sizeof(T) == 4], 1, 4);
        const int kchunk = (w.nz + nc - 1) / nc;
        if (v.cx && v.cy && v.cz)
            hipLaunchKernelGGL((k_zm_restrict3<T>), slab_grid(w.nx, w.ny, (w.nz + kchunk - 1) / kchunk), dim3(256), 0,
                               c->stream, Arr<T>::r(v), Arr<T>::f(w), v.nx, v.ny, v.nz, w.nx, w.ny, w.nz, kchunk);
        else
            hipLaunchKernelGGL((k_zm_restrict<T>), slab_grid(w.nx, w.ny, (w.nz + kchunk - 1) / kchunk), dim3(256), 0,
                               c->stream, Arr<T>::r(v), Arr<T>::f(w), v.nx, v.ny, v.nz, w.nx, w.ny, w.nz, v.cx, v.cy,
                               v.cz, kchunk);
    } else {
        hipLaunchKernelGGL((k_restrict<T>), slab_grid(w), dim3(64, 4), 0, c->stream, Arr<T>::r(v), Arr<T>::f(w), v.nx,
                           v.ny, v.nz, w.nx, w.ny, w.nz, v.cx, v.cy, v.cz);
    }
    vcycle<T>(c, l + 1, true);
    if ((long)v.nx * v.ny * v.nz >= FUSED_MIN_POINTS) {
        // z-marching: a coarse plane's interpolant serves two fine planes
        // (short ranges: many columns in flight hide the latency of each
        // column's load-add-store chain)
        const long cols = (long)nblk(v.nx, 64) * nblk(v.ny, 4);
        const int nc = z_ranges(cols, v.nz, c->res_zp[sizeof(T) == 4], 1, 8);
        const int kchunk = (v.nz + nc - 1) / nc;
        hipLaunchKernelGGL((k_zm_prolong<T>), slab_grid(v.nx, v.ny, (v.nz + kchunk - 1) / kchunk),
                           dim3(256), 0, c->stream, Arr<T>::u(v), Arr<T>::u(w), v.nx, v.ny, v.nz, w.nx, w.ny, v.cx,
                           v.cy, v.cz, kchunk);
    } else {
        hipLaunchKernelGGL((k_prolong<T>), slab_grid(v), dim3(64, 4), 0, c->stream, Arr<T>::u(v), Arr<T>::u(w), v.nx, v.ny,
                           v.nz, w.nx, w.ny, v.cx, v.cy, v.cz);
    }
    smooth<T>(c, v, c->prm.nu2);
}

void drop_graphs(bcg_rd3d* c) {
    if (c->upd_graph) hipGraphExecDestroy(c->upd_graph);
    c->upd_graph = nullptr;
    c->upd_cells = nullptr;
}

int sum_partials(bcg_rd3d* c, const double* part, long nparts) {
    // fixed two-stage tree: 256 chunk sums, then one block; same order every call
    hipLaunchKernelGGL(k_sum_chunks, dim3(256), dim3(256), 0, c->stream, part, nparts, c->part2);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, c->part2, 256, c->dnorm);
    return 0;
}

int read_norm(bcg_rd3d* c, long nparts, double* out) {
    sum_partials(c, c->part, nparts);
    HIPCHK(c, hipMemcpyAsync(c->hnorm, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = std::sqrt(*c->hnorm);
    return 0;
}

long nparts_of(const Level& v) {
    dim3 g = grid2(v);
    return (long)g.x * g.y * g.z;
}

// a0 on every level (FP64 and FP32) and the right-hand side from c_prev.
// a0 on every level and b of the implicit step from c_prev (from x, copied
// into c_prev, when x is given); the FP32 copies of a0 that the mixed cycles
// use only with `f32` (else they are marked stale).
int setup_system(bcg_rd3d* c, bool f32, const double* x = nullptr) {
    Level& f = c->L[0];
    const long n = c->N;
    hipLaunchKernelGGL(k_setup_fine, dim3(nblk(n, 256)), dim3(256), 0, c->stream, x, c->cprev, c->q, c->prm.K,
                       c->prm.dt, c->prm.c_res, (long)f.nx * f.ny, f.a064, f32 ? f.a032 : nullptr, c->b, n);
    for (size_t l = 0; l + 1 < c->L.size(); l++) {
        Level& v = c->L[l];
        Level& w = c->L[l + 1];
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, slab_grid(w), dim3(64, 4), 0, c->stream, v.a064, w.a064, v.nx, v.ny, v.nz, w.nx,
                               w.ny, w.nz);
        };
        switch ((v.cx ? 4 : 0) | (v.cy ? 2 : 0) | (v.cz ? 1 : 0)) {
            case 0: launch(k_restrict_a0<double, 0, 0, 0>); break;
            case 1: launch(k_restrict_a0<double, 0, 0, 1>); break;
            case 2: launch(k_restrict_a0<double, 0, 1, 0>); break;
            case 3: launch(k_restrict_a0<double, 0, 1, 1>); break;
            case 4: launch(k_restrict_a0<double, 1, 0, 0>); break;
            case 5: launch(k_restrict_a0<double, 1, 0, 1>); break;
            case 6: launch(k_restrict_a0<double, 1, 1, 0>); break;
            default: launch(k_restrict_a0<double, 1, 1, 1>); break;
        }
        if (f32)
            hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(w.n(), 256)), dim3(256), 0, c->stream, w.a064, w.a032,
                               w.n());
    }
    c->a32_valid = f32;
    HIPCHK(c, hipGetLastError());
    return 0;
}

// The FP32 copies of a0 on every level, when the last set-up left them stale.
void ensure_a32(bcg_rd3d* c) {
    if (c->a32_valid) return;
    for (Level& v : c->L)
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(v.n(), 256)), dim3(256), 0, c->stream, v.a064, v.a032,
                           v.n());
    c->a32_valid = true;
}

}  // namespace

extern "C" {

int bcg_rd3d_default_params(bcg_rd3d_params* p) {
    if (!p) return BCG_ERR_ARG;
    memset(p, 0, sizeof *p);
    p->nx = 257;
    p->ny = 257;
    p->nz = 65;
    p->h = 2.0;
    p->origin[0] = -256.0;
    p->origin[1] = -256.0;
    p->origin[2] = -8.0;
    p->D = 600.0;
    p->dt = 10.0;
    p->K = 0.1;
    p->c_res = 20.0;
    p->nu1 = 2;
    p->nu2 = 2;
    p->coarse_sweeps = 20;
    p->max_levels = 12;
    p->inner = 2;
    p->smoother = BCG_RD3D_SMOOTH_FUSED;
    return BCG_OK;
}

int bcg_rd3d_create(const bcg_rd3d_params* p, bcg_rd3d** out) {
    if (!p || !out) return BCG_ERR_ARG;
    *out = nullptr;
    if (p->nx < 3 || p->ny < 3 || p->nz < 2 || !(p->h > 0) || !(p->D >= 0) || !(p->dt > 0) || p->nu1 < 0 ||
        p->nu2 < 0 || p->coarse_sweeps < 1 || p->max_levels < 1 || p->inner < 1 ||
        (p->smoother != BCG_RD3D_SMOOTH_FUSED && p->smoother != BCG_RD3D_SMOOTH_SPLIT &&
         p->smoother != BCG_RD3D_SMOOTH_FUSED_WIDE))
        return BCG_ERR_ARG;
    bcg_rd3d* c = new bcg_rd3d();
    c->prm = *p;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BCG_ERR_HIP;
    }
    // levels: halve every direction whose point count is 2m+1 (m >= 1)
    int nx = p->nx, ny = p->ny, nz = p->nz;
    double hx = p->h, hy = p->h, hz = p->h;
    for (int l = 0; l < p->max_levels; l++) {
        Level v;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.hx = hx;
        v.hy = hy;
        v.hz = hz;
        v.lx = p->dt * p->D / (hx * hx);
        v.ly = p->dt * p->D / (hy * hy);
        v.lz = p->dt * p->D / (hz * hz);
        v.cx = (nx >= 5 && (nx - 1) % 2 == 0);
        v.cy = (ny >= 5 && (ny - 1) % 2 == 0);
        v.cz = (nz >= 3 && (nz - 1) % 2 == 0);
        const bool more = (v.cx || v.cy || v.cz) && l + 1 < p->max_levels;
        if (!more) v.cx = v.cy = v.cz = 0;
        c->L.push_back(v);
        if (!more) break;
        if (v.cx) nx = (nx - 1) / 2 + 1, hx *= 2;
        if (v.cy) ny = (ny - 1) / 2 + 1, hy *= 2;
        if (v.cz) nz = (nz - 1) / 2 + 1, hz *= 2;
    }
    c->N = c->L[0].n();
    c->res_rb[0] = resident(k_rbgs2<double, 2, RbNarrow>, RbNarrow::THREADS);
    c->res_rb[1] = resident(k_rbgs2<float, 2, RbNarrow>, RbNarrow::THREADS);
    c->res_rb[2] = resident(k_rbgs2<float, 2, RbWide>, RbWide::THREADS);
    c->res_rb[3] = resident(k_rbgs2<double, 2, RbWide>, RbWide::THREADS);
    c->res_zm = std::min({resident(k_zm_residual<double>, ZM_THREADS), resident(k_zm_resnorm, ZM_THREADS),
                          resident(k_zm_ir, ZM_THREADS)});
    c->res_zp[0] = resident(k_zm_prolong<double>, 256);
    c->res_zp[1] = resident(k_zm_prolong<float>, 256);
    c->res_zr[0] = resident(k_zm_restrict<double>, 256);
    c->res_zr[1] = resident(k_zm_restrict<float>, 256);
    // the small coarse levels (at most BOT_MAX_LEVELS, each <= BOTTOM_MAX_POINTS)
    // form the bottom solver
    c->bottom = (int)c->L.size() - 1;
    while (c->bottom > 0 && c->L[c->bottom - 1].n() <= BOTTOM_MAX_POINTS &&
           (int)c->L.size() - (c->bottom - 1) <= BOT_MAX_LEVELS)
        c->bottom--;
    int rc = 0;
    auto A = [&](auto** ptr, size_t n) {
        if (!rc) rc = dalloc(c, ptr, n);
    };
    const size_t N = (size_t)c->N;
    A(&c->x[0], N);
    A(&c->x[1], N);
    A(&c->cprev, N);
    A(&c->b, N);
    A(&c->q, N);
    A(&c->save, N);
    A(&c->grad, N);
    for (size_t l = 0; l < c->L.size(); l++) {
        Level& v = c->L[l];
        const size_t n = (size_t)v.n();
        if (l == 0) {
            v.f64 = c->b;  // the FP64 path solves on the field itself
        } else {
            A(&v.u64, n);
            A(&v.f64, n);
        }
        A(&v.a064, n);
        A(&v.r64, n);
        A(&v.u32, n);
        A(&v.f32, n);
        A(&v.a032, n);
        A(&v.r32, n);
    }
    c->npart = std::max(nparts_of(c->L[0]), zm_grid(c->L[0].nx, c->L[0].ny, c->L[0].nz, c->res_zm).blocks());
    A(&c->part, (size_t)std::max(c->npart, (long)SUMSQ_BLOCKS));
    A(&c->part2, 256);
    if (c->bottom == 0) {
        A(&c->bot_u0, N);
        A(&c->bot_f0, N);
    }
    A(&c->gpart, 2 * (size_t)c->npart);
    A(&c->dnorm, 1);
    if (!rc && hipHostMalloc((void**)&c->hnorm, sizeof(double), hipHostMallocDefault) != hipSuccess) rc = BCG_ERR_HIP;
    for (int i = 0; i < 5 && !rc; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = BCG_ERR_HIP;
    if (!rc) {
        // c = c_res everywhere, no uptake
        hipLaunchKernelGGL((k_fill<double>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->x[0], c->N, p->c_res);
        hipLaunchKernelGGL((k_fill<double>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->cprev, c->N, p->c_res);
        hipLaunchKernelGGL((k_fill<double>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->q, c->N, 0.0);
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = BCG_ERR_HIP;
    }
    if (rc) {
        bcg_rd3d_destroy(c);
        return rc;
    }
    *out = c;
    return BCG_OK;
}

int bcg_rd3d_destroy(bcg_rd3d* c) {
    if (!c) return BCG_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graphs(c);
    std::vector<void*> ptrs = {c->x[0], c->x[1], c->cprev, c->b, c->q, c->save, c->grad, c->gpart, c->dout, c->dfield,
                               c->part, c->part2, c->bot_u0, c->bot_f0, c->dnorm, c->dkey, c->dcnt,
                               c->dstart, c->dsorted, c->scan_tot, c->dfrac, c->dcm, c->drate};
    for (size_t l = 0; l < c->L.size(); l++) {
        Level& v = c->L[l];
        if (l > 0) {
            ptrs.push_back(v.u64);
            ptrs.push_back(v.f64);
        }
        for (void* q : {(void*)v.a064, (void*)v.r64, (void*)v.u32, (void*)v.f32, (void*)v.a032, (void*)v.r32})
            ptrs.push_back(q);
    }
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (c->hnorm) hipHostFree(c->hnorm);
    for (int i = 0; i < 5; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return BCG_OK;
}

int bcg_rd3d_last_error(const bcg_rd3d* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return BCG_ERR_ARG;
    snprintf(buf, len, "%s", c->err.c_str());
    return BCG_OK;
}

int bcg_rd3d_levels(bcg_rd3d* c, int32_t* n, int32_t* dims, int32_t cap) {
    if (!c || !n) return BCG_ERR_ARG;
    *n = (int32_t)c->L.size();
    for (int l = 0; l < *n && l < cap && dims; l++) {
        dims[3 * l] = c->L[l].nx;
        dims[3 * l + 1] = c->L[l].ny;
        dims[3 * l + 2] = c->L[l].nz;
    }
    return BCG_OK;
}

int bcg_rd3d_set_field(bcg_rd3d* c, const double* f) {
    if (!c || !f) return BCG_ERR_ARG;
    HIPCHK(c, hipMemcpy(c->x[c->cur], f, sizeof(double) * c->N, hipMemcpyHostToDevice));
    return BCG_OK;
}

int bcg_rd3d_get_field(bcg_rd3d* c, double* f) {
    if (!c || !f) return BCG_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(f, c->x[c->cur], sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return BCG_OK;
}

int bcg_rd3d_set_uptake(bcg_rd3d* c, const double* q) {
    if (!c || !q) return BCG_ERR_ARG;
    HIPCHK(c, hipMemcpy(c->q, q, sizeof(double) * c->N, hipMemcpyHostToDevice));
    return BCG_OK;
}

int bcg_rd3d_get_uptake(bcg_rd3d* c, double* q) {
    if (!c || !q) return BCG_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(q, c->q, sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return BCG_OK;
}

}  // extern "C"

namespace {
// This context's next work follows the mechanics context's queued work (and
// whatever that context still waits for).
hipError_t join_cells(bcg_rd3d* c, bcg_ctx* cells) {
    hipError_t e = ctx_join(cells);
    if (e == hipSuccess) e = hipEventRecord(c->ev[4], cells->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev[4], 0);
    return e;
}

// The mechanics context's next kernel that reads the growth rates or writes
// cells follows the work queued here so far; its bins and lists, which do
// neither, run alongside it.
hipError_t release_cells(bcg_rd3d* c, bcg_ctx* cells) {
    hipError_t e = hipEventRecord(cells->after, c->stream);
    if (e == hipSuccess) cells->after_pending = true;
    return e;
}

// Room for n sources in the deposition scratch, allocated for `want` >= n
// (reallocation synchronises the device: a colony's scratch is sized for its
// maxCells once).
int dep_reserve(bcg_rd3d* c, int n, int want) {
    const long nn = c->N;
    if (n > c->dcap) {
        n = want > n ? want : n;
        for (void* q : {(void*)c->dkey, (void*)c->dsorted, (void*)c->dfrac, (void*)c->dcm, (void*)c->drate})
            if (q) hipFree(q);
        c->dkey = c->dsorted = nullptr;
        c->dfrac = c->dcm = c->drate = nullptr;
        int rc = 0;
        auto A = [&](auto** ptr, size_t m) {
            if (!rc) rc = dalloc(c, ptr, m);
        };
        A(&c->dkey, n);
        A(&c->dsorted, n);
        A(&c->dfrac, 3 * (size_t)n);
        A(&c->dcm, 3 * (size_t)n);
        A(&c->drate, n);
        if (rc) return rc;
        c->dcap = n;
    }
    if (!c->dcnt) {
        int rc = 0;
        auto A = [&](auto** ptr, size_t m) {
            if (!rc) rc = dalloc(c, ptr, m);
        };
        A(&c->dcnt, nn + 1);
        A(&c->dstart, nn + 1);
        A(&c->scan_tot, 1024 * 64);
        if (rc) return rc;
    }
    return 0;
}

// Deposition of n sources into q (stream-ordered); their rates are in
// c->drate (written there by the key kernel for cells).
template <class Src>
int dep_run(bcg_rd3d* c, Src src, int n, int32_t* n_inside) {
    Level& v = c->L[0];
    const long nn = c->N;
    // the counts return to zero in the scatter below; cleared here only after
    // a deposition that stopped before it
    if (!c->dcnt_zero) HIPCHK(c, hipMemsetAsync(c->dcnt, 0, sizeof(int) * (nn + 1), c->stream));
    c->dcnt_zero = n == 0;
    if (n > 0)
        hipLaunchKernelGGL(k_dep_key<Src>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, src, n, c->prm.origin[0],
                           c->prm.origin[1], c->prm.origin[2], 1.0 / c->prm.h, v.nx, v.ny, v.nz, c->dkey, c->dfrac,
                           c->dcnt, c->drate);
    // exclusive scan of the bucket counts (the mechanics' scan kernels)
    const long ns = nn + 1;
    const int nb = nblk(ns, TSCAN_ITEMS);
    if (nb > 1024 * 64) return fail(c, BCG_ERR_ARG, "grid too large for the deposition scan");
    hipLaunchKernelGGL(k_tscan_sums, dim3(nb), dim3(TSCAN_T), 0, c->stream, c->dcnt, (int)ns, c->scan_tot);
    hipLaunchKernelGGL(k_tscan_apply, dim3(nb), dim3(TSCAN_T), 0, c->stream, c->dcnt, (int)ns, c->scan_tot, c->dstart,
                       (int*)nullptr);
    if (n > 0) {
        hipLaunchKernelGGL(k_dep_scatter_dec, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->dkey, n, c->dstart,
                           c->dcnt, c->dsorted);
        c->dcnt_zero = true;
        hipLaunchKernelGGL(k_dep_sort, dim3(nblk(nn, 256)), dim3(256), 0, c->stream, c->dstart, nn, c->dsorted);
    }
    hipLaunchKernelGGL(k_dep_gather, grid2(v), dim3(64, 4), 0, c->stream, c->dstart, c->dsorted, c->dfrac, c->drate,
                       v.nx, v.ny, v.nz, 1.0 / (c->prm.h * c->prm.h * c->prm.h), c->q);
    HIPCHK(c, hipGetLastError());
    if (n_inside) {
        int tot = 0;
        HIPCHK(c, hipMemcpyAsync(&tot, c->dstart + nn, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *n_inside = tot;
    }
    return 0;
}
}  // namespace

extern "C" {

int bcg_rd3d_deposit(bcg_rd3d* c, const double* cm, const double* rate, int32_t n, int32_t* n_inside) {
    if (!c || n < 0 || (n > 0 && (!cm || !rate))) return BCG_ERR_ARG;
    int rc = dep_reserve(c, n, n);
    if (rc) return rc;
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(c->dcm, cm, 3 * sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->drate, rate, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    }
    if ((rc = dep_run(c, PointSources{c->dcm}, n, n_inside))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BCG_OK;
}

}  // extern "C"

// ---- deposition without allocation ---------------------------------------
// Scratch of bcg_rd3d_deposit_noalloc: one array per device, held until the
// process ends. Host code never passes `noalloc_scratch` itself to a kernel
// (on the host it names a stand-in, not device memory): it uses the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char noalloc_scratch[BCG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mutex;
void* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of this device's scratch array (looked up once per device,
// with that device current).
int noalloc_scratch_of(bcg_rd3d* c, unsigned char** out) {
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(c, BCG_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, BCG_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        noalloc_ptr[dev] = ptr;
    }
    *out = (unsigned char*)noalloc_ptr[dev];
    return 0;
}

__global__ void k_fill_int(int* __restrict__ a, long n, int v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int bcg_rd3d_deposit_noalloc(bcg_rd3d* c, const double* d_cm, const double* d_rate, int32_t n, double* d_q,
                             void* stream) {
    if (!c || n < 0 || !d_q || (n > 0 && (!d_cm || !d_rate))) return BCG_ERR_ARG;
    Level& v = c->L[0];
    const long nn = c->N;
    const long ns = nn + 1;
    const int nb = nblk(ns, SCAN_BLOCK_ITEMS);
    // scratch layout: key[n], frac[3n], sorted[n], cnt[nn+1], start[nn+1], fill[nn], scan totals[nb]
    const size_t o_key = 0;
    const size_t o_frac = o_key + align256(sizeof(int) * (size_t)n);
    const size_t o_sorted = o_frac + align256(sizeof(double) * 3 * (size_t)n);
    const size_t o_cnt = o_sorted + align256(sizeof(int) * (size_t)n);
    const size_t o_start = o_cnt + align256(sizeof(int) * (size_t)ns);
    const size_t o_fill = o_start + align256(sizeof(int) * (size_t)ns);
    const size_t o_tot = o_fill + align256(sizeof(int) * (size_t)nn);
    const size_t need = o_tot + align256(sizeof(int) * (size_t)nb);
    if (need > (size_t)BCG_NOALLOC_SCRATCH_BYTES)
        return fail(c, BCG_ERR_CAPACITY,
                    "deposition of %d sources on %ld vertices needs %zu bytes of scratch, more than the %d bytes "
                    "(BCG_NOALLOC_SCRATCH_BYTES) this entry point has",
                    n, nn, need, (int)BCG_NOALLOC_SCRATCH_BYTES);
    unsigned char* base = nullptr;
    int rc = noalloc_scratch_of(c, &base);
    if (rc) return rc;
    int* key = (int*)(base + o_key);
    double* frac = (double*)(base + o_frac);
    int* sorted = (int*)(base + o_sorted);
    int* cnt = (int*)(base + o_cnt);
    int* start = (int*)(base + o_start);
    int* fillc = (int*)(base + o_fill);
    int* tot = (int*)(base + o_tot);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    // counters start at zero (a fill kernel: no memset node if the caller captures)
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(ns, 256)), dim3(256), 0, st, cnt, ns, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nn, 256)), dim3(256), 0, st, fillc, nn, 0);
    if (n > 0)
        hipLaunchKernelGGL(k_dep_key<PointSources>, dim3(nblk(n, 256)), dim3(256), 0, st, PointSources{d_cm}, n,
                           c->prm.origin[0], c->prm.origin[1], c->prm.origin[2], 1.0 / c->prm.h, v.nx, v.ny, v.nz, key,
                           frac, cnt, (double*)nullptr);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, st, cnt, (int)ns, tot);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, st, tot, nb, (int*)nullptr);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, cnt, (int)ns, tot, start);
    if (n > 0) {
        hipLaunchKernelGGL(k_dep_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, key, n, start, fillc, sorted);
        hipLaunchKernelGGL(k_dep_sort, dim3(nblk(nn, 256)), dim3(256), 0, st, start, nn, sorted);
    }
    hipLaunchKernelGGL(k_dep_gather, grid2(v), dim3(64, 4), 0, st, start, sorted, frac, d_rate, v.nx, v.ny, v.nz,
                       1.0 / (c->prm.h * c->prm.h * c->prm.h), d_q);
    HIPCHK(c, hipGetLastError());
    return BCG_OK;
}

int bcg_rd3d_deposit_cells(bcg_rd3d* c, bcg_ctx* cells, double rate, int32_t* n_inside) {
    if (!c || !cells) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(cells)) return fail(c, rc_, "%s", cells->err.c_str());
    const int n = cells->N;
    int rc = dep_reserve(c, n, cells->prm.maxCells);
    if (rc) return rc;
    HIPCHK(c, join_cells(c, cells));
    if ((rc = dep_run(c, CellSources{cells->hot[cells->cur], rate}, n, n_inside))) return rc;
    HIPCHK(c, release_cells(c, cells));
    return BCG_OK;
}

int bcg_rd3d_update_cells(bcg_rd3d* c, bcg_ctx* cells, double rate, double mu, double K, int32_t vcycles) {
    if (!c || !cells || !(K > 0) || vcycles < 1) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(cells)) return fail(c, rc_, "%s", cells->err.c_str());
    const int n = cells->N;
    int rc = dep_reserve(c, n, cells->prm.maxCells);
    if (rc) return rc;
    HIPCHK(c, join_cells(c, cells));
    if ((rc = dep_run(c, CellSources{cells->hot[cells->cur], rate}, n, nullptr))) return rc;
    // the implicit step from the current field and the growth rates: kernel
    // launches of fixed shape, replayed from a graph
    Level& f = c->L[0];
    double* x = c->x[c->cur];
    const bcg::PhysParams& P = cells->P;
    auto body = [&]() {
        setup_system(c, false, x);
        f.u64 = x;
        for (int i = 0; i < vcycles; i++) vcycle<double>(c, 0);
        f.u64 = nullptr;
        const long nb = (long)P.BoxX * P.BoxY * P.BoxZ;
        hipLaunchKernelGGL(k_growth_bins, dim3(nblk(nb, 256)), dim3(256), 0, c->stream, x, f.nx, f.ny, f.nz,
                           c->prm.origin[0], c->prm.origin[1], c->prm.origin[2], 1.0 / c->prm.h, P.BoxX, P.BoxY,
                           P.BoxZ, P.BoxLength, mu, K, cells->gr1);
    };
    if (c->upd_graph && (c->upd_cells != cells || c->upd_vcycles != vcycles || c->upd_cur != c->cur ||
                         c->upd_mu != mu || c->upd_K != K))
        drop_graphs(c);
    if (!c->upd_graph) {
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        body();
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) return fail(c, BCG_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&c->upd_graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            c->upd_graph = nullptr;
            return fail(c, BCG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e2));
        }
        c->upd_cells = cells;
        c->upd_vcycles = vcycles;
        c->upd_cur = c->cur;
        c->upd_mu = mu;
        c->upd_K = K;
    }
    HIPCHK(c, hipGraphLaunch(c->upd_graph, c->stream));
    c->a32_valid = false;  // the replayed set-up writes FP64 a0 only
    HIPCHK(c, release_cells(c, cells));
    return BCG_OK;
}

int bcg_rd3d_growth_to_cells(bcg_rd3d* c, bcg_ctx* cells, double mu, double K) {
    if (!c || !cells || !(K > 0)) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(cells)) return fail(c, rc_, "%s", cells->err.c_str());
    const bcg::PhysParams& P = cells->P;
    const Level& v = c->L[0];
    HIPCHK(c, join_cells(c, cells));
    const long nb = (long)P.BoxX * P.BoxY * P.BoxZ;
    hipLaunchKernelGGL(k_growth_bins, dim3(nblk(nb, 256)), dim3(256), 0, c->stream, c->x[c->cur], v.nx, v.ny, v.nz,
                       c->prm.origin[0], c->prm.origin[1], c->prm.origin[2], 1.0 / c->prm.h, P.BoxX, P.BoxY, P.BoxZ,
                       P.BoxLength, mu, K, cells->gr1);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, release_cells(c, cells));
    return BCG_OK;
}

}  // extern "C"

namespace {

// Solve A x = b (the current a0, b) from the initial guess x = c->x[c->cur]
// to ||b - A x|| <= tol ||b||; the result is left in c->x[c->cur].
int solve_core(bcg_rd3d* c, int mode, double tol, int max_iter, int* iters, double* relres) {
    Level& f = c->L[0];
    double* xo = c->x[c->cur];
    int rc = 0;
    int it = 0;
    double r0 = 0, rn = 0, bn = 0;
    const ZmGrid zg = zm_grid(f.nx, f.ny, f.nz, c->res_zm);
    const long np = zg.blocks();  // one partial sum per z-marching workgroup
    if (mode == 1) ensure_a32(c);
    // the tolerance is relative to ||b||
    hipLaunchKernelGGL((k_sumsq<double>), dim3(SUMSQ_BLOCKS), dim3(256), 0, c->stream, c->b, c->N, c->part);
    if ((rc = read_norm(c, SUMSQ_BLOCKS, &bn))) return rc;
    if (mode == 0) {
        f.u64 = xo;
        auto resnorm = [&]() {
            hipLaunchKernelGGL(k_zm_resnorm, dim3(np), dim3(ZM_THREADS), 0, c->stream, xo, c->b, f.a064, f.nx, f.ny,
                               f.nz, f.lx, f.ly, f.lz, zg.tiles_x, zg.tiles_y, zg.kchunk, c->part);
        };
        resnorm();
        if ((rc = read_norm(c, np, &rn))) return rc;
        r0 = bn;
        while (it < max_iter && rn > tol * r0) {
            vcycle<double>(c, 0);
            resnorm();
            if ((rc = read_norm(c, np, &rn))) return rc;
            it++;
        }
        f.u64 = nullptr;
    } else {
        // r = b - A x0, r32 = r (x0 stays where it is)
        double* xn = c->x[c->cur ^ 1];
        auto ir_update = [&](const double* xo_, double* xn_, const float* e, double inv_scale, double rscale) {
            hipLaunchKernelGGL(k_zm_ir, dim3(np), dim3(ZM_THREADS), 0, c->stream, xo_, xn_, e, inv_scale, c->b, f.a064,
                               f.f32, rscale, f.nx, f.ny, f.nz, f.lx, f.ly, f.lz, zg.tiles_x, zg.tiles_y, zg.kchunk,
                               c->part);
        };
        ir_update(xo, nullptr, nullptr, 0.0, 1.0);
        if ((rc = read_norm(c, np, &rn))) return rc;
        r0 = bn;
        double scale = 1.0;  // r32 = r * scale
        while (it < max_iter && rn > tol * r0) {
            // inner FP32 solve of A e = r32 from e = 0
            for (int s = 0; s < c->prm.inner; s++) vcycle<float>(c, 0, s == 0);
            // x += e / scale ; r = b - A x ; r32 = r * next scale
            const double rms = rn / std::sqrt((double)c->N);
            const double next = rms > 0 ? std::ldexp(1.0, -std::ilogb(rms)) : 1.0;
            xo = c->x[c->cur];
            xn = c->x[c->cur ^ 1];
            ir_update(xo, xn, f.u32, 1.0 / scale, next);
            if ((rc = read_norm(c, np, &rn))) return rc;
            c->cur ^= 1;
            scale = next;
            it++;
        }
    }
    HIPCHK(c, hipGetLastError());
    if (iters) *iters = it;
    if (relres) *relres = r0 > 0 ? rn / r0 : 0.0;
    if (rn > tol * r0)
        return fail(c, BCG_ERR_STATE, "no convergence: relative residual %.3e after %d iterations", r0 > 0 ? rn / r0 : 0.0,
                    it);
    return 0;
}

int upload(bcg_rd3d* c, double** dev, size_t n, const double* host) {
    if (!*dev) {
        int rc = dalloc(c, dev, n);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(*dev, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return 0;
}

}  // namespace

extern "C" {

int bcg_rd3d_step(bcg_rd3d* c, int32_t mode, double tol, int32_t max_iter, int32_t* iters, double* relres) {
    if (!c || (mode != 0 && mode != 1) || max_iter < 1 || !(tol > 0)) return BCG_ERR_ARG;
    // c_prev <- c ; a0, b from c_prev ; initial guess c_new = c_prev
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    int rc = setup_system(c, mode == 1, c->x[c->cur]);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    rc = solve_core(c, mode, tol, max_iter, iters, relres);
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipEventElapsedTime(&c->last_ms[0], c->ev[0], c->ev[1]);
    hipEventElapsedTime(&c->last_ms[1], c->ev[1], c->ev[2]);
    return rc;
}

int bcg_rd3d_set_state(bcg_rd3d* c, const double* c_old, const double* c_new) {
    if (!c || !c_old || !c_new) return BCG_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(c->cprev, c_old, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->x[c->cur], c_new, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    int rc = setup_system(c, true);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BCG_OK;
}

int bcg_rd3d_set_coeffs(bcg_rd3d* c, double D, double K) {
    if (!c || !(D >= 0) || !(K > 0)) return BCG_ERR_ARG;
    c->prm.D = D;
    c->prm.K = K;
    drop_graphs(c);  // a recorded update holds the old coefficients
    for (Level& v : c->L) {
        v.lx = c->prm.dt * D / (v.hx * v.hx);
        v.ly = c->prm.dt * D / (v.hy * v.hy);
        v.lz = c->prm.dt * D / (v.hz * v.hz);
    }
    return BCG_OK;
}

int bcg_rd3d_adjoint(bcg_rd3d* c, const double* g, int32_t mode, double tol, int32_t max_iter, double* dL_dq,
                     double* dL_dcold, double* dL_dK, double* dL_dD) {
    if (!c || !g || (mode != 0 && mode != 1) || max_iter < 1 || !(tol > 0)) return BCG_ERR_ARG;
    Level& f = c->L[0];
    const size_t N = (size_t)c->N;
    int rc = upload(c, &c->dfield, N, g);
    if (rc) return rc;
    // keep c_new; solve A mu = W^-1 g with zero Dirichlet rows from mu = 0
    HIPCHK(c, hipMemcpyAsync(c->save, c->x[c->cur], N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_adj_rhs, grid2(f), dim3(64, 4), 0, c->stream, c->dfield, c->b, f.nx, f.ny, f.nz);
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->x[c->cur], c->N, 0.0);
    // the forward state (c_new in x, the step's b) comes back on every exit
    auto restore = [&]() -> int {
        HIPCHK(c, hipMemcpyAsync(c->x[c->cur], c->save, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL(k_rhs, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->cprev, c->b, c->prm.c_res,
                           (long)f.nx * f.ny, c->N);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    };
    int it = 0;
    double rr = 0;
    rc = solve_core(c, mode, tol, max_iter, &it, &rr);
    if (rc) {
        const std::string msg = c->err;
        restore();
        c->err = "adjoint solve: " + msg;
        return rc;
    }
    const double* mu = c->x[c->cur];
    const double h = c->prm.h;
    hipLaunchKernelGGL(k_adj_grad, grid2(f), dim3(64, 4), 0, c->stream, mu, c->save, c->cprev, c->q, c->prm.K, c->prm.dt,
                       1.0 / (h * h), f.nx, f.ny, f.nz, c->grad, c->dfield, c->gpart, c->gpart + c->npart);
    HIPCHK(c, hipGetLastError());
    if (dL_dq) HIPCHK(c, hipMemcpyAsync(dL_dq, c->grad, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (dL_dcold) HIPCHK(c, hipMemcpyAsync(dL_dcold, c->dfield, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    double gk = 0, gd = 0;
    sum_partials(c, c->gpart, c->npart);
    HIPCHK(c, hipMemcpyAsync(c->hnorm, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    gk = *c->hnorm;
    sum_partials(c, c->gpart + c->npart, c->npart);
    HIPCHK(c, hipMemcpyAsync(c->hnorm, c->dnorm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    gd = *c->hnorm;
    if (dL_dK) *dL_dK = gk;
    if (dL_dD) *dL_dD = gd;
    return restore();
}

int bcg_rd3d_solve(bcg_rd3d* c, int32_t mode, const double* rhs, double tol, int32_t max_iter, double* x,
                   int32_t* iters, double* relres) {
    if (!c || !rhs || !x || (mode != 0 && mode != 1) || max_iter < 1 || !(tol > 0)) return BCG_ERR_ARG;
    Level& f = c->L[0];
    const size_t N = (size_t)c->N;
    int rc = upload(c, &c->dfield, N, rhs);
    if (rc) return rc;
    // keep the field; b <- rhs; solve from x = 0
    HIPCHK(c, hipMemcpyAsync(c->save, c->x[c->cur], N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->b, c->dfield, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->x[c->cur], c->N, 0.0);
    int it = 0;
    double rr = 0;
    rc = solve_core(c, mode, tol, max_iter, &it, &rr);
    const std::string msg = c->err;
    // the last iterate goes out whether or not tol was reached
    HIPCHK(c, hipMemcpyAsync(x, c->x[c->cur], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    // the step's state comes back: the field and its right-hand side
    HIPCHK(c, hipMemcpyAsync(c->x[c->cur], c->save, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_rhs, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->cprev, c->b, c->prm.c_res,
                       (long)f.nx * f.ny, c->N);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (iters) *iters = it;
    if (relres) *relres = rr;
    if (rc) c->err = msg;
    return rc;
}

int bcg_rd3d_gather(bcg_rd3d* c, const double* cm, int32_t n, const double* field, double* out) {
    if (!c || n < 0 || (n > 0 && (!cm || !field || !out))) return BCG_ERR_ARG;
    if (n == 0) return BCG_OK;
    Level& f = c->L[0];
    int rc = upload(c, &c->dfield, (size_t)c->N, field);
    if (rc) return rc;
    if (n > c->dcap) {
        for (void* q : {(void*)c->dkey, (void*)c->dsorted, (void*)c->dfrac, (void*)c->dcm, (void*)c->drate})
            if (q) hipFree(q);
        c->dkey = c->dsorted = nullptr;
        c->dfrac = c->dcm = c->drate = nullptr;
        auto A = [&](auto** ptr, size_t m) {
            if (!rc) rc = dalloc(c, ptr, m);
        };
        A(&c->dkey, n);
        A(&c->dsorted, n);
        A(&c->dfrac, 3 * (size_t)n);
        A(&c->dcm, 3 * (size_t)n);
        A(&c->drate, n);
        if (rc) return rc;
        c->dcap = n;
    }
    HIPCHK(c, hipMemcpyAsync(c->dcm, cm, 3 * sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_cic_gather, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->dcm, n, c->prm.origin[0],
                       c->prm.origin[1], c->prm.origin[2], 1.0 / c->prm.h, f.nx, f.ny, f.nz, c->dfield,
                       1.0 / (c->prm.h * c->prm.h * c->prm.h), c->drate);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, c->drate, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BCG_OK;
}

int bcg_rd3d_vcycle(bcg_rd3d* c, int32_t mode, float* ms) {
    if (!c || (mode != 0 && mode != 1)) return BCG_ERR_ARG;
    Level& f = c->L[0];
    int rc = setup_system(c, mode == 1, c->x[c->cur]);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    if (mode == 0) {
        f.u64 = c->x[c->cur];
        vcycle<double>(c, 0);
        f.u64 = nullptr;
    } else {
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->b, f.f32, c->N);
        vcycle<float>(c, 0, true);
    }
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ms) hipEventElapsedTime(ms, c->ev[0], c->ev[1]);
    return BCG_OK;
}

int bcg_rd3d_last_times(bcg_rd3d* c, float* ms2) {
    if (!c || !ms2) return BCG_ERR_ARG;
    ms2[0] = c->last_ms[0];
    ms2[1] = c->last_ms[1];
    return BCG_OK;
}

}  // extern "C"
