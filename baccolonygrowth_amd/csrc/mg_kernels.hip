// mg_kernels.hip - 3-D implicit reaction-diffusion multigrid (gfx950).
//
// Operator on a vertex grid (x fastest, then y, then z):
//   (A u)_p = a0_p u_p - lx (uW + uE) - ly (uS + uN) - lz (uB + uT)
//             + 2 (lx + ly + lz) u_p,           a0 = 1 + dt*sigma
// Neumann (mirror) faces at x = 0, nx-1, y = 0, ny-1, z = nz-1; the z = 0
// rows are Dirichlet (identity). l* = dt*D/h*^2 per direction and level.
//
// Kernels are templated on the storage/arithmetic type T (float for the
// inner solve of the mixed-precision refinement, double for the FP64 path):
//   rbgs       red-black Gauss-Seidel sweep of one colour; each thread owns
//              one point of that colour in an x row, so all lanes are busy
//              and a wave covers a contiguous 512 B (fp32) row segment
//   residual   r = f - A u
//   restrict   full weighting (27 points, mirrored at Neumann faces)
//   prolong    trilinear interpolation added to the fine iterate
//   norms      per-block FP64 partial sums, reduced in a fixed order
// plus the FP64 outer-loop kernels of the mixed-precision refinement.
#include <hip/hip_runtime.h>

#include "mg_kernels.h"
#include "rbgs_fused.h"
#include "zmarch.h"

namespace bcg {
namespace mg {

__device__ __forceinline__ int mirror(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i); }

// Tile (bx, by, bz) of a 1-D launch over gx x gy x gz tiles of 64 x 4 points
// (x, y) by one plane (z). The hardware deals workgroups round-robin over the
// 8 XCDs, each with its own L2; the remap gives XCD x the consecutive tiles
// [x * per, (x + 1) * per) in x-fastest, then y, then z order - a slab of
// planes - so that the cache lines a tile shares with its x neighbours (rows
// are not line-aligned) and the planes it shares with its z neighbours are
// read through the same L2, shortly after each other.
__device__ __forceinline__ void slab_tile(int gx, int gy, int& bx, int& by, int& bz) {
    constexpr int X = 8;
    const int nb = gridDim.x, per = nb / X, rem = nb % X;
    const int xcd = blockIdx.x % X;
    const int b = xcd * per + (xcd < rem ? xcd : rem) + blockIdx.x / X;
    bx = b % gx;
    by = (b / gx) % gy;
    bz = b / (gx * gy);
}

template <class T>
struct Lvl {
    T* u;
    const T* f;
    const T* a0;
    int nx, ny, nz;
    T lx, ly, lz;
};

template <class T>
__device__ __forceinline__ T neigh_sum(const T* __restrict__ u, int i, int j, int k, int nx, int ny, int nz, T lx, T ly,
                                       T lz, long& p) {
    const long sxy = (long)nx * ny;
    p = (long)k * sxy + (long)j * nx + i;
    const T w = u[p + (i > 0 ? -1 : 1)];
    const T e = u[p + (i < nx - 1 ? 1 : -1)];
    const T s = u[p + (j > 0 ? -nx : nx)];
    const T n = u[p + (j < ny - 1 ? nx : -nx)];
    const T b = u[p - sxy];  // k >= 1 here
    const T t = u[p + (k < nz - 1 ? sxy : -sxy)];
    return lx * (w + e) + ly * (s + n) + lz * (b + t);
}

template <class T>
__global__ void __launch_bounds__(256) k_rbgs(T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ a0,
                                              int nx, int ny, int nz, T lx, T ly, T lz, int color) {
    int bx, by, bz;
    slab_tile(((nx + 1) / 2 + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int j = by * blockDim.y + threadIdx.y;
    const int k = bz;
    if (j >= ny) return;
    const int i = 2 * (bx * blockDim.x + threadIdx.x) + ((j + k + color) & 1);
    if (i >= nx) return;
    const long sxy = (long)nx * ny;
    if (k == 0) {
        const long p = (long)j * nx + i;
        u[p] = f[p];
        return;
    }
    long p;
    const T s = neigh_sum<T>(u, i, j, k, nx, ny, nz, lx, ly, lz, p);
    const T d = a0[p] + T(2) * (lx + ly + lz);
    u[p] = (f[p] + s) / d;
    (void)sxy;
}

// r = f - A u at (i, j, k).
template <class T>
__device__ __forceinline__ T residual_point(const T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ a0,
                                            int i, int j, int k, int nx, int ny, int nz, T lx, T ly, T lz) {
    if (k == 0) {
        const long p = (long)j * nx + i;
        return f[p] - u[p];
    }
    long p;
    const T s = neigh_sum<T>(u, i, j, k, nx, ny, nz, lx, ly, lz, p);
    const T d = a0[p] + T(2) * (lx + ly + lz);
    return f[p] - (d * u[p] - s);
}

template <class T>
__global__ void __launch_bounds__(256) k_residual(const T* __restrict__ u, const T* __restrict__ f,
                                                  const T* __restrict__ a0, T* __restrict__ r, int nx, int ny, int nz,
                                                  T lx, T ly, T lz) {
    int bx, by, bz;
    slab_tile((nx + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int k = bz;
    if (i >= nx || j >= ny) return;
    r[((long)k * ny + j) * nx + i] = residual_point<T>(u, f, a0, i, j, k, nx, ny, nz, lx, ly, lz);
}

// Full weighting at coarse point (I,J,K) (K >= 1); cx/cy/cz = 1 where that
// direction is coarsened, Neumann faces mirrored. rv(i, j, k): the fine
// residual at a fine point.
template <class T, class RV>
__device__ __forceinline__ T restrict_point_of(RV rv, int nx, int ny, int nz, int I, int J, int K, int cx, int cy,
                                               int cz) {
    if (K == 0) return T(0);  // Dirichlet rows: zero correction
    const int i0 = cx ? 2 * I : I, j0 = cy ? 2 * J : J, k0 = cz ? 2 * K : K;
    T acc = T(0);
    for (int c = -cz; c <= cz; c++) {
        const int kk = mirror(k0 + c, nz);
        const T wz = (c == 0) ? T(cz ? 0.5 : 1.0) : T(0.25);
        T accy = T(0);
        for (int b = -cy; b <= cy; b++) {
            const int jj = mirror(j0 + b, ny);
            const T wy = (b == 0) ? T(cy ? 0.5 : 1.0) : T(0.25);
            T accx;
            if (cx)
                accx = T(0.25) * rv(mirror(i0 - 1, nx), jj, kk) + T(0.5) * rv(i0, jj, kk) +
                       T(0.25) * rv(mirror(i0 + 1, nx), jj, kk);
            else accx = rv(i0, jj, kk);
            accy += wy * accx;
        }
        acc += wz * accy;
    }
    return acc;
}

template <class T>
__device__ __forceinline__ T restrict_point(const T* __restrict__ r, int nx, int ny, int nz, int I, int J, int K, int cx,
                                            int cy, int cz) {
    const long sxy = (long)nx * ny;
    auto rv = [&](int i, int j, int k) { return r[(long)k * sxy + (long)j * nx + i]; };
    return restrict_point_of<T>(rv, nx, ny, nz, I, J, K, cx, cy, cz);
}

template <class T>
__global__ void __launch_bounds__(256) k_restrict(const T* __restrict__ r, T* __restrict__ fc, int nx, int ny, int nz,
                                                  int ncx, int ncy, int ncz, int cx, int cy, int cz) {
    int bx, by, bz;
    slab_tile((ncx + 63) / 64, (ncy + 3) / 4, bx, by, bz);
    const int I = bx * blockDim.x + threadIdx.x;
    const int J = by * blockDim.y + threadIdx.y;
    const int K = bz;
    if (I >= ncx || J >= ncy) return;
    fc[((long)K * ncy + J) * ncx + I] = restrict_point<T>(r, nx, ny, nz, I, J, K, cx, cy, cz);
}

// restrict_point for coarse planes [K0, K0 + kchunk) of one coarse column
// (I, J), marching up: the x-y weighted sum of a fine plane (restrict_point's
// accy) is formed once and serves the two coarse planes it belongs to. Same
// operations in the same order as restrict_point.
template <class T>
__global__ void __launch_bounds__(256) k_zm_restrict(const T* __restrict__ r, T* __restrict__ fc, int nx, int ny, int nz,
                                                     int ncx, int ncy, int ncz, int cx, int cy, int cz, int kchunk) {
    int bx, by, bz;
    slab_tile((ncx + 63) / 64, (ncy + 3) / 4, bx, by, bz);
    const int I = bx * 64 + (threadIdx.x & 63), J = by * 4 + (threadIdx.x >> 6);
    if (I >= ncx || J >= ncy) return;
    const int K0 = bz * kchunk, K1 = min(K0 + kchunk, ncz);
    const int i0 = cx ? 2 * I : I, j0 = cy ? 2 * J : J;
    const int im = mirror(i0 - 1, nx), ip = mirror(i0 + 1, nx);
    long rows[3];
    for (int b = -cy; b <= cy; b++) rows[b + 1] = (long)mirror(j0 + b, ny) * nx;
    const long sxy = (long)nx * ny;
    auto plane = [&](int kk) {
        const T* base = r + kk * sxy;
        T accy = T(0);
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            if (b < -cy || b > cy) continue;
            const T wy = (b == 0) ? T(cy ? 0.5 : 1.0) : T(0.25);
            const T* row = base + rows[b + 1];
            T accx;
            if (cx) accx = T(0.25) * row[im] + T(0.5) * row[i0] + T(0.25) * row[ip];
            else accx = row[i0];
            accy += wy * accx;
        }
        return accy;
    };
    T* out = fc + (long)J * ncx + I;
    const long scxy = (long)ncx * ncy;
    int kc = -1;  // fine plane whose sum is cached
    T pc = T(0);
    for (int K = K0; K < K1; K++) {
        if (K == 0) {
            out[0] = T(0);  // Dirichlet rows: zero correction
            continue;
        }
        const int k0 = cz ? 2 * K : K;
        T acc = T(0);
        for (int c = -cz; c <= cz; c++) {
            const int kk = mirror(k0 + c, nz);
            const T wz = (c == 0) ? T(cz ? 0.5 : 1.0) : T(0.25);
            const T pk = kk == kc ? pc : plane(kk);
            acc += wz * pk;
            if (c == cz) {
                kc = kk;
                pc = pk;
            }
        }
        out[K * scxy] = acc;
    }
}

// k_zm_restrict with every axis coarsened: the fine planes 2K-1, 2K, 2K+1
// of consecutive coarse planes are walked in order and the next two fine
// planes' nine values are requested a coarse plane ahead (two register
// sets), so that a column's loads overlap its sums. Same operations in the
// same order as restrict_point.
template <class T>
__global__ void __launch_bounds__(256) k_zm_restrict3(const T* __restrict__ r, T* __restrict__ fc, int nx, int ny,
                                                      int nz, int ncx, int ncy, int ncz, int kchunk) {
    int bx, by, bz;
    slab_tile((ncx + 63) / 64, (ncy + 3) / 4, bx, by, bz);
    const int I = bx * 64 + (threadIdx.x & 63), J = by * 4 + (threadIdx.x >> 6);
    if (I >= ncx || J >= ncy) return;
    int K = bz * kchunk;
    const int K1 = min(K + kchunk, ncz);
    T* out = fc + (long)J * ncx + I;
    const long scxy = (long)ncx * ncy;
    if (K == 0) {
        out[0] = T(0);  // Dirichlet rows: zero correction
        K = 1;
    }
    if (K >= K1) return;
    const int i0 = 2 * I, j0 = 2 * J;
    const int im = mirror(i0 - 1, nx), ip = mirror(i0 + 1, nx);
    const long r0 = (long)mirror(j0 - 1, ny) * nx, r1 = (long)j0 * nx, r2 = (long)mirror(j0 + 1, ny) * nx;
    const long sxy = (long)nx * ny;
    auto load9 = [&](int f, T (&v)[9]) {
        const T* base = r + (long)mirror(f, nz) * sxy;
        v[0] = base[r0 + im], v[1] = base[r0 + i0], v[2] = base[r0 + ip];
        v[3] = base[r1 + im], v[4] = base[r1 + i0], v[5] = base[r1 + ip];
        v[6] = base[r2 + im], v[7] = base[r2 + i0], v[8] = base[r2 + ip];
    };
    auto sum9 = [&](const T (&v)[9]) {
        T accy = T(0);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const T wy = (b == 1) ? T(0.5) : T(0.25);
            const T accx = T(0.25) * v[3 * b] + T(0.5) * v[3 * b + 1] + T(0.25) * v[3 * b + 2];
            accy += wy * accx;
        }
        return accy;
    };
    T A[9], B[9];
    load9(2 * K - 1, A);
    const T p_first = sum9(A);
    load9(2 * K, B);
    load9(2 * K + 1, A);
    T p_odd = p_first;
    for (; K < K1; K++) {
        const T p_even = sum9(B);
        if (K + 1 < K1) load9(2 * K + 2, B);
        const T p_next = sum9(A);
        if (K + 1 < K1) load9(2 * K + 3, A);
        T acc = T(0);
        acc += T(0.25) * p_odd;
        acc += T(0.5) * p_even;
        acc += T(0.25) * p_next;
        out[K * scxy] = acc;
        p_odd = p_next;
    }
}

// Trilinear interpolation of the coarse correction at fine point (i,j,k).
template <class T>
__device__ __forceinline__ T prolong_point(const T* __restrict__ ec, int ncx, int ncy, int i, int j, int k, int cx, int cy,
                                           int cz) {
    int I0, I1, J0, J1, K0, K1;
    T wi, wj, wk;  // weight of the second point
    if (cx) {
        I0 = i >> 1;
        I1 = (i + 1) >> 1;
        wi = (i & 1) ? T(0.5) : T(0);
    } else {
        I0 = I1 = i;
        wi = T(0);
    }
    if (cy) {
        J0 = j >> 1;
        J1 = (j + 1) >> 1;
        wj = (j & 1) ? T(0.5) : T(0);
    } else {
        J0 = J1 = j;
        wj = T(0);
    }
    if (cz) {
        K0 = k >> 1;
        K1 = (k + 1) >> 1;
        wk = (k & 1) ? T(0.5) : T(0);
    } else {
        K0 = K1 = k;
        wk = T(0);
    }
    const long scxy = (long)ncx * ncy;
    auto E = [&](int K, int J, int I) { return ec[(long)K * scxy + (long)J * ncx + I]; };
    const T e00 = (T(1) - wi) * E(K0, J0, I0) + wi * E(K0, J0, I1);
    const T e01 = (T(1) - wi) * E(K0, J1, I0) + wi * E(K0, J1, I1);
    const T e10 = (T(1) - wi) * E(K1, J0, I0) + wi * E(K1, J0, I1);
    const T e11 = (T(1) - wi) * E(K1, J1, I0) + wi * E(K1, J1, I1);
    const T e0 = (T(1) - wj) * e00 + wj * e01;
    const T e1 = (T(1) - wj) * e10 + wj * e11;
    return (T(1) - wk) * e0 + wk * e1;
}

// u_fine += P e_coarse.
template <class T>
__global__ void __launch_bounds__(256) k_prolong(T* __restrict__ u, const T* __restrict__ ec, int nx, int ny, int nz,
                                                 int ncx, int ncy, int cx, int cy, int cz) {
    int bx, by, bz;
    slab_tile((nx + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int k = bz;
    if (i >= nx || j >= ny) return;
    u[((long)k * ny + j) * nx + i] += prolong_point<T>(ec, ncx, ncy, i, j, k, cx, cy, cz);
}

// The red sweep that starts a smoothing pass from u = 0: the black
// neighbours enter as zeros (the sweep that follows sets every black point
// without reading it, so u need not be cleared first).
template <class T>
__global__ void __launch_bounds__(256) k_rbgs_zero(T* __restrict__ u, const T* __restrict__ f, const T* __restrict__ a0,
                                                   int nx, int ny, int nz, T lx, T ly, T lz) {
    int bx, by, bz;
    slab_tile(((nx + 1) / 2 + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int j = by * blockDim.y + threadIdx.y;
    const int k = bz;
    if (j >= ny) return;
    const int i = 2 * (bx * blockDim.x + threadIdx.x) + ((j + k) & 1);
    if (i >= nx) return;
    const long p = ((long)k * ny + j) * nx + i;
    if (k == 0) {
        u[p] = f[p];
        return;
    }
    const T s = T(0);
    const T d = a0[p] + T(2) * (lx + ly + lz);
    u[p] = (f[p] + s) / d;
}

// k_prolong followed by the red sweep, in one launch: a red point reads each
// black neighbour as u + P e (what k_prolong would have stored there). Its
// own interpolated value is overwritten and the black points are set by the
// black sweep that follows without being read, so nothing else is stored.
template <class T>
__global__ void __launch_bounds__(256) k_rbgs_prolong(T* __restrict__ u, const T* __restrict__ f,
                                                      const T* __restrict__ a0, const T* __restrict__ ec, int nx,
                                                      int ny, int nz, T lx, T ly, T lz, int ncx, int ncy, int cx,
                                                      int cy, int cz) {
    int bx, by, bz;
    slab_tile(((nx + 1) / 2 + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int j = by * blockDim.y + threadIdx.y;
    const int k = bz;
    if (j >= ny) return;
    const int i = 2 * (bx * blockDim.x + threadIdx.x) + ((j + k) & 1);
    if (i >= nx) return;
    const long p = ((long)k * ny + j) * nx + i;
    if (k == 0) {
        u[p] = f[p];
        return;
    }
    auto v = [&](int ii, int jj, int kk) {
        return u[((long)kk * ny + jj) * nx + ii] + prolong_point<T>(ec, ncx, ncy, ii, jj, kk, cx, cy, cz);
    };
    const T w = v(i > 0 ? i - 1 : i + 1, j, k);
    const T e = v(i < nx - 1 ? i + 1 : i - 1, j, k);
    const T sn = v(i, j > 0 ? j - 1 : j + 1, k);
    const T nn = v(i, j < ny - 1 ? j + 1 : j - 1, k);
    const T b = v(i, j, k - 1);
    const T t = v(i, j, k < nz - 1 ? k + 1 : k - 1);
    const T s = lx * (w + e) + ly * (sn + nn) + lz * (b + t);
    const T d = a0[p] + T(2) * (lx + ly + lz);
    u[p] = (f[p] + s) / d;
}

// u_fine += P e_coarse, a thread per fine (x, y) column marching up planes
// [k0, k0 + kchunk): the x-y interpolated coarse value of a coarse plane,
// the e0 / e1 of prolong_point, is formed once and serves the two fine
// planes it touches. Same operations as prolong_point.
template <class T>
__global__ void __launch_bounds__(256) k_zm_prolong(T* __restrict__ u, const T* __restrict__ ec, int nx, int ny, int nz,
                                                    int ncx, int ncy, int cx, int cy, int cz, int kchunk) {
    int bx, by, bz;
    slab_tile((nx + 63) / 64, (ny + 3) / 4, bx, by, bz);
    const int i = bx * 64 + (threadIdx.x & 63);
    const int j = by * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const int k0 = bz * kchunk, k1 = min(k0 + kchunk, nz);
    int I0, I1, J0, J1;
    T wi, wj;
    if (cx) {
        I0 = i >> 1;
        I1 = (i + 1) >> 1;
        wi = (i & 1) ? T(0.5) : T(0);
    } else {
        I0 = I1 = i;
        wi = T(0);
    }
    if (cy) {
        J0 = j >> 1;
        J1 = (j + 1) >> 1;
        wj = (j & 1) ? T(0.5) : T(0);
    } else {
        J0 = J1 = j;
        wj = T(0);
    }
    const long scxy = (long)ncx * ncy;
    const long o00 = (long)J0 * ncx + I0, o01 = (long)J0 * ncx + I1;
    const long o10 = (long)J1 * ncx + I0, o11 = (long)J1 * ncx + I1;
    auto plane = [&](int K) {
        const T* p = ec + K * scxy;
        const T r0 = (T(1) - wi) * p[o00] + wi * p[o01];
        const T r1 = (T(1) - wi) * p[o10] + wi * p[o11];
        return (T(1) - wj) * r0 + wj * r1;
    };
    // two cached coarse planes; K only grows with k
    int ka = -1, kb = -1;
    T ea = T(0), eb = T(0);
    auto val = [&](int K) {
        if (K == ka) return ea;
        if (K == kb) return eb;
        const T v = plane(K);
        if (ka < kb) {
            ka = K;
            ea = v;
        } else {
            kb = K;
            eb = v;
        }
        return v;
    };
    const long sxy = (long)nx * ny;
    T* up = u + (long)j * nx + i;
    // four planes at a time, their loads all issued before the first is
    // used: with one load in flight per thread the pass is latency-bound
    constexpr int G = 4;
    for (int kb = k0; kb < k1; kb += G) {
        T old[G];
#pragma unroll
        for (int g = 0; g < G; g++) old[g] = kb + g < k1 ? up[(kb + g) * sxy] : T(0);
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int k = kb + g;
            if (k >= k1) break;
            int K0, K1;
            T wk;
            if (cz) {
                K0 = k >> 1;
                K1 = (k + 1) >> 1;
                wk = (k & 1) ? T(0.5) : T(0);
            } else {
                K0 = K1 = k;
                wk = T(0);
            }
            const T e0 = val(K0);
            const T e1 = val(K1);
            up[k * sxy] = old[g] + ((T(1) - wk) * e0 + wk * e1);
        }
    }
}

template <class T>
__global__ void k_fill(T* __restrict__ a, long n, T v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <class Ts, class Td>
__global__ void k_cast(const Ts* __restrict__ a, Td* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (Td)a[i];
}

// Block sums of x^2 (FP64), one partial per block.
template <class T>
__global__ void __launch_bounds__(256) k_sumsq(const T* __restrict__ x, long n, double* __restrict__ part) {
    __shared__ double lds[4];
    double s = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = (double)x[i];
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double lds[4];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// The system of one implicit step from the previous field c, in one pass over
// the fine grid: a0 = 1 + dt * q / (K + c) (Monod uptake linearised about c;
// FP32 copy when a0f is given) and b as in k_rhs; with `x` given, c is x and
// is copied into cprev on the way.
__global__ void k_setup_fine(const double* __restrict__ x, double* __restrict__ cprev, const double* __restrict__ q,
                             double K, double dt, double c_res, long nxy, double* __restrict__ a0,
                             float* __restrict__ a0f, double* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c;
    if (x) {
        c = x[i];
        cprev[i] = c;
    } else {
        c = cprev[i];
    }
    const double a = 1.0 + dt * q[i] / (K + c);
    a0[i] = a;
    if (a0f) a0f[i] = (float)a;
    b[i] = i < nxy ? c_res : c;
}

// b = c with the Dirichlet rows set to c_res.
__global__ void k_rhs(const double* __restrict__ c, double* __restrict__ b, double c_res, long nxy, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = i < nxy ? c_res : c[i];
}

// Coarse-level a0 from the fine one: a0_c - 1 = restricted (a0_f - 1).
template <class T, int cx, int cy, int cz>
__global__ void __launch_bounds__(256) k_restrict_a0(const T* __restrict__ a0f, T* __restrict__ a0c, int nx, int ny,
                                                     int nz, int ncx, int ncy, int ncz) {
    int bx, by, bz;
    slab_tile((ncx + 63) / 64, (ncy + 3) / 4, bx, by, bz);
    const int I = bx * blockDim.x + threadIdx.x;
    const int J = by * blockDim.y + threadIdx.y;
    const int K = bz;
    if (I >= ncx || J >= ncy) return;
    const int i0 = cx ? 2 * I : I, j0 = cy ? 2 * J : J, k0 = cz ? 2 * K : K;
    const long sxy = (long)nx * ny;
    T acc = T(0), wsum = T(0);
#pragma unroll
    for (int c = -cz; c <= cz; c++) {
        const int kk = k0 + c;
        if (kk < 0 || kk >= nz) continue;
        const T wz = (c == 0) ? T(cz ? 0.5 : 1.0) : T(0.25);
#pragma unroll
        for (int b = -cy; b <= cy; b++) {
            const int jj = mirror(j0 + b, ny);
            const T wy = (b == 0) ? T(cy ? 0.5 : 1.0) : T(0.25);
#pragma unroll
            for (int a = -cx; a <= cx; a++) {
                const int ii = mirror(i0 + a, nx);
                const T wx = (a == 0) ? T(cx ? 0.5 : 1.0) : T(0.25);
                acc += wz * wy * wx * (a0f[(long)kk * sxy + (long)jj * nx + ii] - T(1));
                wsum += wz * wy * wx;
            }
        }
    }
    a0c[((long)K * ncy + J) * ncx + I] = T(1) + acc / wsum;
}


// One red-black sweep (both colours) from uin into uout, z-marching; the
// algorithm and its phases are in rbgs_fused.h. Workgroups are numbered
// tile-x fastest, then tile-y, then z range, and the hardware workgroup
// order (round-robin over the 8 XCDs) is remapped so that each XCD runs a
// contiguous run of tiles, which share their rings through its L2. AHEAD is
// the number of steps a plane load is in flight.
template <class T, int AHEAD, class D>
__global__ void __launch_bounds__(D::THREADS) k_rbgs2(const T* __restrict__ uin, T* __restrict__ uout,
                                                      const T* __restrict__ f, const T* __restrict__ a0, int nx, int ny,
                                                      int nz, T lx, T ly, T lz, int tiles_x, int tiles_y, int kchunk) {
    __shared__ RbShared<T, D> sh;
    constexpr int X = 8;
    const int nb = gridDim.x, per = nb / X, rem = nb % X;
    const int xcd = blockIdx.x % X;
    const int b = xcd * per + (xcd < rem ? xcd : rem) + blockIdx.x / X;
    const int bx = b % tiles_x, by = (b / tiles_x) % tiles_y, bz = b / (tiles_x * tiles_y);
    const RbArgs<T> a{uin, uout, f, a0, nx, ny, nz, lx, ly, lz};
    const int tid = threadIdx.x;
    RbThread<T, D> t;
    RbPlane<T> pa, pb;
    rb_init(a, bx, by, bz, kchunk, tid, t);
    rb_prologue(a, sh, t, tid);
    if (AHEAD == 2 && rb_plane_needed(a, t, t.kfirst + 3)) rb_fetch_plane(a, t, t.kfirst + 3, pa);
    __syncthreads();
    rb_prologue_init(sh, t);
    __syncthreads();
    auto step = [&](int k, RbPlane<T>& now, RbPlane<T>& later) {
        rb_fetch_next(a, t, k, AHEAD, later);
        rb_compute(a, sh, t, k);
        rb_finish(a, sh, t, k, tid, now);
        __syncthreads();
    };
    if (AHEAD == 2) {
        for (int k = t.kfirst; k < t.k1; k += 2) {
            step(k, pa, pb);
            if (k + 1 < t.k1) step(k + 1, pb, pa);
        }
    } else {
        for (int k = t.kfirst; k < t.k1; k++) step(k, pa, pa);
    }
}

// z-marching stencil passes (zmarch.h): one workgroup per ZM_TX x ZM_TY
// tile and z range, XCD-remapped like k_rbgs2; `part` (when given) gets one
// sum of squares per workgroup, in workgroup order.
template <class T, class Op>
__device__ __forceinline__ void zm_run(const Op& op, int nx, int ny, int nz, int tiles_x, int tiles_y, int kchunk,
                                       double* __restrict__ part) {
    __shared__ ZmShared<T> sh;
    __shared__ double red[ZM_THREADS / 64];
    constexpr int X = 8;
    const int nb = gridDim.x, per = nb / X, rem = nb % X;
    const int xcd = blockIdx.x % X;
    const int b = xcd * per + (xcd < rem ? xcd : rem) + blockIdx.x / X;
    const int bx = b % tiles_x, by = (b / tiles_x) % tiles_y, bz = b / (tiles_x * tiles_y);
    const int tid = threadIdx.x;
    ZmThread<T> t;
    typename Op::Pre pa, pb;
    zm_init<T>(nx, ny, nz, bx, by, bz, kchunk, tid, t);
    zm_prologue(op, sh, t, pa);
    __syncthreads();
    double acc = 0.0;
    auto step = [&](int k, const typename Op::Pre& cur, typename Op::Pre& next) {
        zm_fetch(op, t, next, nz, k);
        acc += zm_point(op, sh, t, cur, nz, k);
        zm_finish(sh, t, nz, k);
        __syncthreads();
    };
    for (int k = t.k0; k < t.k1; k += 2) {
        step(k, pa, pb);
        if (k + 1 < t.k1) step(k + 1, pb, pa);
    }
    if (part) {
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < ZM_THREADS / 64; w++) s += red[w];
            part[b] = s;
        }
    }
}

// zm_run two planes a step: a 4-slot LDS ring holds planes k, k+1 (read) and
// k+2, k+3 (being stored), so a workgroup passes one barrier per two planes
// and a thread has two planes' loads in flight. The same point arithmetic
// (Op::point) and, for `part`, the same per-workgroup sums in the same order
// of planes.
template <class T, class Op>
__device__ __forceinline__ void zm_run2(const Op& op, int nx, int ny, int nz, int tiles_x, int tiles_y, int kchunk,
                                        double* __restrict__ part) {
    __shared__ T U[4][ZM_PLANE];
    __shared__ double red[ZM_THREADS / 64];
    constexpr int X = 8;
    const int nb = gridDim.x, per = nb / X, rem = nb % X;
    const int xcd = blockIdx.x % X;
    const int b = xcd * per + (xcd < rem ? xcd : rem) + blockIdx.x / X;
    const int bx = b % tiles_x, by = (b / tiles_x) % tiles_y, bz = b / (tiles_x * tiles_y);
    const int tid = threadIdx.x;
    ZmThread<T> t;
    zm_init<T>(nx, ny, nz, bx, by, bz, kchunk, tid, t);
    const int k0 = t.k0, k1 = t.k1;
    // prologue: planes k0, k0+1 into their slots, the own column at k0-1 .. k0+1
    T below = k0 >= 1 ? op.u(k0 - 1, t.pb) : T(0);
    T center = op.u(k0, t.pb);
    T c1 = k0 + 1 <= nz - 1 ? op.u(k0 + 1, t.pb) : T(0);
    U[k0 & 3][t.li] = center;
    if (t.lr >= 0) U[k0 & 3][t.lr] = op.u(k0, t.rb);
    if (k0 + 1 <= nz - 1) {
        U[(k0 + 1) & 3][t.li] = c1;
        if (t.lr >= 0) U[(k0 + 1) & 3][t.lr] = op.u(k0 + 1, t.rb);
    }
    typename Op::Pre p0, p1;
    op.fetch(k0, t.pb, p0);
    if (k0 + 1 < k1) op.fetch(k0 + 1, t.pb, p1);
    __syncthreads();
    double acc = 0.0;
    for (int k = k0; k < k1; k += 2) {
        // loads of planes k+2, k+3 and of the next step's inputs
        T na = T(0), nb2 = T(0), ra = T(0), rb = T(0);
        if (k + 2 <= nz - 1) {
            na = op.u(k + 2, t.pb);
            if (t.lr >= 0) ra = op.u(k + 2, t.rb);
        }
        if (k + 3 <= nz - 1) {
            nb2 = op.u(k + 3, t.pb);
            if (t.lr >= 0) rb = op.u(k + 3, t.rb);
        }
        typename Op::Pre q0, q1;
        if (k + 2 < k1) op.fetch(k + 2, t.pb, q0);
        if (k + 3 < k1) op.fetch(k + 3, t.pb, q1);
        if (t.act) {
            const int i = t.li;
            {
                const T* P = U[k & 3];
                const T above = k + 1 <= nz - 1 ? c1 : below;  // mirror across the top face
                const T sm = op.lx * (P[i - 1] + P[i + 1]) + op.ly * (P[i - ZM_PX] + P[i + ZM_PX]) + op.lz * (below + above);
                acc += op.point(k, t.pb, center, sm, p0);
            }
            if (k + 1 < k1) {
                const T* P = U[(k + 1) & 3];
                const T above = k + 2 <= nz - 1 ? na : center;
                const T sm = op.lx * (P[i - 1] + P[i + 1]) + op.ly * (P[i - ZM_PX] + P[i + ZM_PX]) + op.lz * (center + above);
                acc += op.point(k + 1, t.pb, c1, sm, p1);
            }
        }
        if (k + 2 <= nz - 1) {
            U[(k + 2) & 3][t.li] = na;
            if (t.lr >= 0) U[(k + 2) & 3][t.lr] = ra;
        }
        if (k + 3 <= nz - 1) {
            U[(k + 3) & 3][t.li] = nb2;
            if (t.lr >= 0) U[(k + 3) & 3][t.lr] = rb;
        }
        below = c1;
        center = na;
        c1 = nb2;
        p0 = q0;
        p1 = q1;
        __syncthreads();
    }
    if (part) {
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < ZM_THREADS / 64; w++) s += red[w];
            part[b] = s;
        }
    }
}

// k_zm_residual on zm_run2 (FP32, whose points carry half the bytes of FP64).
template <class T>
__global__ void __launch_bounds__(ZM_THREADS) k_zm_residual2(const T* __restrict__ u, const T* __restrict__ f,
                                                             const T* __restrict__ a0, T* __restrict__ r, int nx,
                                                             int ny, int nz, T lx, T ly, T lz, int tiles_x,
                                                             int tiles_y, int kchunk) {
    const ZmResidual<T> op{u, f, a0, r, (long)nx * ny, lx, ly, lz};
    zm_run2<T>(op, nx, ny, nz, tiles_x, tiles_y, kchunk, nullptr);
}

template <class T>
__global__ void __launch_bounds__(ZM_THREADS) k_zm_residual(const T* __restrict__ u, const T* __restrict__ f,
                                                            const T* __restrict__ a0, T* __restrict__ r, int nx, int ny,
                                                            int nz, T lx, T ly, T lz, int tiles_x, int tiles_y,
                                                            int kchunk) {
    const ZmResidual<T> op{u, f, a0, r, (long)nx * ny, lx, ly, lz};
    zm_run<T>(op, nx, ny, nz, tiles_x, tiles_y, kchunk, nullptr);
}

__global__ void __launch_bounds__(ZM_THREADS) k_zm_resnorm(const double* __restrict__ u, const double* __restrict__ b,
                                                           const double* __restrict__ a0, int nx, int ny, int nz,
                                                           double lx, double ly, double lz, int tiles_x, int tiles_y,
                                                           int kchunk, double* __restrict__ part) {
    const ZmResNorm op{u, b, a0, (long)nx * ny, lx, ly, lz};
    zm_run<double>(op, nx, ny, nz, tiles_x, tiles_y, kchunk, part);
}

__global__ void __launch_bounds__(ZM_THREADS) k_zm_ir(const double* __restrict__ xo, double* __restrict__ xn,
                                                      const float* __restrict__ e, double inv_scale,
                                                      const double* __restrict__ b, const double* __restrict__ a0,
                                                      float* __restrict__ r32, double rscale, int nx, int ny, int nz,
                                                      double lx, double ly, double lz, int tiles_x, int tiles_y,
                                                      int kchunk, double* __restrict__ part) {
    const ZmIrUpdate op{xo, xn, e, inv_scale, b, a0, r32, rscale, (long)nx * ny, lx, ly, lz};
    zm_run<double>(op, nx, ny, nz, tiles_x, tiles_y, kchunk, part);
}

// ------------------------------------------------------------- adjoint
// Control-volume weight of a vertex: 1/2 per Neumann face it lies on
// (x = 0, nx-1; y = 0, ny-1; z = nz-1). W A is symmetric on the free rows.
__device__ __forceinline__ double cv_weight(int i, int j, int k, int nx, int ny, int nz) {
    double w = 1.0;
    if (i == 0 || i == nx - 1) w *= 0.5;
    if (j == 0 || j == ny - 1) w *= 0.5;
    if (k == nz - 1) w *= 0.5;
    return w;
}

// rhs of the adjoint solve: W^-1 g on the free rows, 0 on the Dirichlet rows.
__global__ void __launch_bounds__(256) k_adj_rhs(const double* __restrict__ g, double* __restrict__ b, int nx, int ny,
                                                 int nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const long p = ((long)k * ny + j) * nx + i;
    b[p] = (k == 0) ? 0.0 : g[p] / cv_weight(i, j, k, nx, ny, nz);
}

// Parameter gradients from mu (A mu = W^-1 g), lambda = W mu on free rows:
//   dL/da0   = -lambda c1
//   dL/dq    = dL/da0 * dt/(K + c0)
//   dL/dc0   = lambda + dL/da0 * (-dt q/(K + c0)^2)
//   dL/dK    = sum dL/da0 * (-dt q/(K + c0)^2)          (block partials)
//   dL/dD    = sum lambda (Lap_h c1) dt/h^2             (block partials)
__global__ void __launch_bounds__(256) k_adj_grad(const double* __restrict__ mu, const double* __restrict__ c1,
                                                  const double* __restrict__ c0, const double* __restrict__ q, double K,
                                                  double dt, double inv_h2, int nx, int ny, int nz,
                                                  double* __restrict__ dq, double* __restrict__ dc0,
                                                  double* __restrict__ partK, double* __restrict__ partD) {
    __shared__ double ldsK[4], ldsD[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    double gK = 0, gD = 0;
    if (i < nx && j < ny) {
        const long sxy = (long)nx * ny;
        const long p = (long)k * sxy + (long)j * nx + i;
        if (k == 0) {
            if (dq) dq[p] = 0.0;
            if (dc0) dc0[p] = 0.0;
        } else {
            const double lam = cv_weight(i, j, k, nx, ny, nz) * mu[p];
            const double u = c1[p];
            const double den = K + c0[p];
            const double ga0 = -lam * u;
            const double da0_dc = -dt * q[p] / (den * den);
            if (dq) dq[p] = ga0 * dt / den;
            if (dc0) dc0[p] = lam + ga0 * da0_dc;
            gK = ga0 * da0_dc;
            const double lap = (c1[p + (i > 0 ? -1 : 1)] + c1[p + (i < nx - 1 ? 1 : -1)]) +
                               (c1[p + (j > 0 ? -nx : nx)] + c1[p + (j < ny - 1 ? nx : -nx)]) +
                               (c1[p - sxy] + c1[p + (k < nz - 1 ? sxy : -sxy)]) - 6.0 * u;
            gD = lam * lap * dt * inv_h2;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        gK += __shfl_xor(gK, o, 64);
        gD += __shfl_xor(gD, o, 64);
    }
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if ((tid & 63) == 0) {
        ldsK[tid >> 6] = gK;
        ldsD[tid >> 6] = gD;
    }
    __syncthreads();
    if (tid == 0) {
        const long bid = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partK[bid] = (ldsK[0] + ldsK[1]) + (ldsK[2] + ldsK[3]);
        partD[bid] = (ldsD[0] + ldsD[1]) + (ldsD[2] + ldsD[3]);
    }
}

// Transpose of the cloud-in-cell deposition: out[c] = sum_v field[v] w_cv / h^3.
__global__ void k_cic_gather(const double* __restrict__ cm, int n, double ox, double oy, double oz, double inv_h, int nx,
                             int ny, int nz, const double* __restrict__ field, double inv_vol, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double gx = (cm[3 * c] - ox) * inv_h, gy = (cm[3 * c + 1] - oy) * inv_h, gz = (cm[3 * c + 2] - oz) * inv_h;
    const double fx = floor(gx), fy = floor(gy), fz = floor(gz);
    double acc = 0.0;
    if (fx >= 0 && fx < nx - 1 && fy >= 0 && fy < ny - 1 && fz >= 0 && fz < nz - 1) {
        const double ax = gx - fx, ay = gy - fy, az = gz - fz;
        const long b = ((long)fz * ny + (long)fy) * nx + (long)fx;
        for (int dz = 0; dz <= 1; dz++)
            for (int dy = 0; dy <= 1; dy++)
                for (int dx = 0; dx <= 1; dx++) {
                    const double w = (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay) * (dz ? az : 1.0 - az);
                    acc += field[b + (long)dz * nx * ny + (long)dy * nx + dx] * w;
                }
    }
    out[c] = acc * inv_vol;
}


// ------------------------------------------------------ bottom solver
// The small coarse levels run as one workgroup: a whole sub-V-cycle (and
// the coarsest-level sweeps) in a single launch, FP64, in LDS when the
// levels fit. Levels hold at most a few thousand points: 32-bit indices.
// The point updates repeat the arithmetic of k_rbgs / k_residual /
// k_restrict / k_prolong.
// One colour: the threads take the points of that colour only (slot m of a
// row is i = 2m + ((j + k + color) & 1); the last slot of a row of odd
// length may fall outside it).
__device__ __forceinline__ void bot_rbgs(const BotLevel& v, int color, int nthr, int tid) {
    // a thread keeps its x slot and steps over rows (j, k) R at a time
    const int nx = v.nx, ny = v.ny, nz = v.nz;
    const int hx = (nx + 1) >> 1, R = nthr / hx;
    const int ix = tid % hx, r0 = tid / hx;
    if (r0 >= R) return;
    const int sxy = nx * ny;
    const double d2 = 2.0 * (v.lx + v.ly + v.lz);
    int k = r0 / ny, j = r0 - k * ny;
    for (int r = r0; r < ny * nz; r += R) {
        const int i = 2 * ix + ((j + k + color) & 1);
        if (i < nx) {
            const int p = k * sxy + j * nx + i;
            if (k == 0) {
                v.u[p] = v.f[p];
            } else {
                long q;
                const double s = neigh_sum<double>(v.u, i, j, k, v.nx, v.ny, v.nz, v.lx, v.ly, v.lz, q);
                v.u[p] = (v.f[p] + s) / (v.a0[p] + d2);
            }
        }
        j += R;
        while (j >= ny) {
            j -= ny;
            k++;
        }
    }
}

// A thread's point of each colour on a level whose colour has at most one
// point per thread: index, neighbour offsets (bot_rbgs' mirrored stencil)
// and the diagonal, formed once for all the level's sweeps.
struct BotSlot {
    int p[2], o[2][6];
    double d[2];
    bool on[2], dir[2];
};

__device__ __forceinline__ void bot_slot(const BotLevel& v, int tid, BotSlot& S) {
    const int nx = v.nx, ny = v.ny, nz = v.nz, hx = (nx + 1) >> 1, sxy = nx * ny;
    const int nslot = hx * ny * nz;
    const double d2 = 2.0 * (v.lx + v.ly + v.lz);
#pragma unroll
    for (int color = 0; color < 2; color++) {
        S.on[color] = false;
        if (tid >= nslot) continue;
        const int row = tid / hx;
        const int k = row / ny, j = row - k * ny;
        const int i = 2 * (tid - row * hx) + ((j + k + color) & 1);
        if (i >= nx) continue;
        const int p = k * sxy + j * nx + i;
        S.on[color] = true;
        S.dir[color] = k == 0;
        S.p[color] = p;
        int* o = S.o[color];
        o[0] = i > 0 ? -1 : 1;
        o[1] = i < nx - 1 ? 1 : -1;
        o[2] = j > 0 ? -nx : nx;
        o[3] = j < ny - 1 ? nx : -nx;
        o[4] = -sxy;
        o[5] = k < nz - 1 ? sxy : -sxy;
        S.d[color] = k == 0 ? 1.0 : v.a0[p] + d2;
    }
}

// bot_rbgs from a BotSlot (the same operations).
__device__ __forceinline__ void bot_rbgs_slot(const BotLevel& v, const BotSlot& S, int color) {
    if (!S.on[color]) return;
    const int p = S.p[color];
    if (S.dir[color]) {
        v.u[p] = v.f[p];
        return;
    }
    const int* o = S.o[color];
    const double s = v.lx * (v.u[p + o[0]] + v.u[p + o[1]]) + v.ly * (v.u[p + o[2]] + v.u[p + o[3]]) +
                     v.lz * (v.u[p + o[4]] + v.u[p + o[5]]);
    v.u[p] = (v.f[p] + s) / S.d[color];
}

// `sweeps` red-black sweeps of a bottom level, a barrier after each colour.
__device__ __forceinline__ void bot_sweeps(const BotLevel& v, int sweeps, int nthr, int tid) {
    if (sweeps <= 0) return;
    const int nslot = ((v.nx + 1) >> 1) * v.ny * v.nz;
    if (nslot <= nthr) {
        BotSlot S;
        bot_slot(v, tid, S);
        for (int s = 0; s < sweeps; s++)
#pragma unroll
            for (int color = 0; color < 2; color++) {
                bot_rbgs_slot(v, S, color);
                __syncthreads();
            }
    } else {
        for (int s = 0; s < sweeps; s++)
            for (int color = 0; color < 2; color++) {
                bot_rbgs(v, color, nthr, tid);
                __syncthreads();
            }
    }
}

// Levels lds_from .. n-1 live in dynamic LDS for the whole launch (u, f, a0
// per level and one residual buffer of the largest of them); the levels
// above are worked on in place in memory. Only u of the first level goes
// back to memory.
__global__ void __launch_bounds__(1024) k_bottom(BotLevels B, int nu1, int nu2, int coarse_sweeps, int lds_from) {
    extern __shared__ double smem[];
    __shared__ BotLevel S[BOT_MAX_LEVELS];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int last = B.n - 1;
    const int in_lds = lds_from == 0;  // level 0 in LDS
    if (tid == 0) {
        int off = 0;
        for (int l = 0; l < B.n; l++) {
            S[l] = B.L[l];
            const int n = (int)S[l].nx * S[l].ny * S[l].nz;
            if (l >= lds_from) {
                S[l].u = smem + off;
                S[l].f = smem + off + n;
                S[l].a0 = smem + off + 2 * n;
                off += 3 * n;
            }
        }
        for (int l = lds_from; l < B.n; l++) S[l].r = smem + off;
    }
    __syncthreads();
    for (int l = lds_from; l < B.n; l++) {
        const int n = (int)S[l].nx * S[l].ny * S[l].nz;
        for (int p = tid; p < n; p += nthr) S[l].a0[p] = B.L[l].a0[p];
    }
    // level 0 into its working arrays: from the FP32 arrays, from memory into
    // LDS, and/or from u = 0 (in place otherwise)
    const int n0 = (int)S[0].nx * S[0].ny * S[0].nz;
    if (in_lds || B.f32 || B.zero)
        for (int p = tid; p < n0; p += nthr) {
            const double f = B.f32 ? (double)B.f32[p] : B.L[0].f[p];
            const double u = B.zero ? 0.0 : (B.u32 ? (double)B.u32[p] : B.L[0].u[p]);
            S[0].f[p] = f;
            S[0].u[p] = u;
        }
    __syncthreads();
    for (int l = 0; l < last; l++) {
        const BotLevel v = S[l];  // in registers for the level's phases
        const BotLevel w = S[l + 1];
        bot_sweeps(v, nu1, nthr, tid);
        // residual into r
        const int n = (int)v.nx * v.ny * v.nz, sxy = (int)v.nx * v.ny;
        const double d2 = 2.0 * (v.lx + v.ly + v.lz);
        for (int p = tid; p < n; p += nthr) {
            const int k = (int)(p / sxy);
            const int j = (int)((p - k * sxy) / v.nx);
            const int i = (int)(p - k * sxy - (int)j * v.nx);
            if (k == 0) {
                v.r[p] = v.f[p] - v.u[p];
                continue;
            }
            long q;
            const double sn = neigh_sum<double>(v.u, i, j, k, v.nx, v.ny, v.nz, v.lx, v.ly, v.lz, q);
            v.r[p] = v.f[p] - ((v.a0[p] + d2) * v.u[p] - sn);
        }
        __syncthreads();
        // restrict r into f of the next level, clear its u
        const int nc = (int)w.nx * w.ny * w.nz, scxy = (int)w.nx * w.ny;
        for (int pc = tid; pc < nc; pc += nthr) {
            const int K = (int)(pc / scxy);
            const int J = (int)((pc - K * scxy) / w.nx);
            const int I = (int)(pc - K * scxy - (int)J * w.nx);
            w.f[pc] = restrict_point<double>(v.r, v.nx, v.ny, v.nz, I, J, K, v.cx, v.cy, v.cz);
            w.u[pc] = 0.0;
        }
        __syncthreads();
    }
    {
        const BotLevel v = S[last];
        const int n = (int)v.nx * v.ny * v.nz;
        for (int p = tid; p < n; p += nthr) v.u[p] = 0.0;
        __syncthreads();
        bot_sweeps(v, coarse_sweeps, nthr, tid);
    }
    for (int l = last - 1; l >= 0; l--) {
        const BotLevel v = S[l];  // in registers for the level's phases
        const BotLevel w = S[l + 1];
        const int n = (int)v.nx * v.ny * v.nz, sxy = (int)v.nx * v.ny;
        for (int p = tid; p < n; p += nthr) {
            const int k = (int)(p / sxy);
            const int j = (int)((p - k * sxy) / v.nx);
            const int i = (int)(p - k * sxy - (int)j * v.nx);
            v.u[p] += prolong_point<double>(w.u, w.nx, w.ny, i, j, k, v.cx, v.cy, v.cz);
        }
        __syncthreads();
        bot_sweeps(v, nu2, nthr, tid);
    }
    if (B.u32)
        for (int p = tid; p < n0; p += nthr) B.u32[p] = (float)S[0].u[p];
    else if (in_lds)
        for (int p = tid; p < n0; p += nthr) B.L[0].u[p] = S[0].u[p];
}

// Deterministic reduction of many partial sums: fixed chunks per block.
__global__ void __launch_bounds__(256) k_sum_chunks(const double* __restrict__ part, long n, double* __restrict__ out) {
    __shared__ double lds[4];
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    double s = 0;
    for (long i = lo + threadIdx.x; i < hi; i += 256) s += part[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// ---------------------------------------------------- explicit instances
template __global__ void k_rbgs<float>(float*, const float*, const float*, int, int, int, float, float, float, int);
template __global__ void k_rbgs<double>(double*, const double*, const double*, int, int, int, double, double, double,
                                        int);
template __global__ void k_rbgs2<float, 2, RbNarrow>(const float*, float*, const float*, const float*, int, int, int,
                                                     float, float, float, int, int, int);
template __global__ void k_rbgs2<float, 2, RbWide>(const float*, float*, const float*, const float*, int, int, int,
                                                   float, float, float, int, int, int);
template __global__ void k_rbgs2<double, 2, RbWide>(const double*, double*, const double*, const double*, int, int, int,
                                                    double, double, double, int, int, int);
template __global__ void k_rbgs2<double, 2, RbNarrow>(const double*, double*, const double*, const double*, int, int,
                                                      int, double, double, double, int, int, int);
template __global__ void k_zm_residual2<float>(const float*, const float*, const float*, float*, int, int, int, float,
                                               float, float, int, int, int);
template __global__ void k_zm_residual<float>(const float*, const float*, const float*, float*, int, int, int, float,
                                              float, float, int, int, int);
template __global__ void k_zm_residual<double>(const double*, const double*, const double*, double*, int, int, int,
                                               double, double, double, int, int, int);
template __global__ void k_residual<float>(const float*, const float*, const float*, float*, int, int, int, float,
                                           float, float);
template __global__ void k_residual<double>(const double*, const double*, const double*, double*, int, int, int,
                                            double, double, double);
template __global__ void k_restrict<float>(const float*, float*, int, int, int, int, int, int, int, int, int);
template __global__ void k_restrict<double>(const double*, double*, int, int, int, int, int, int, int, int, int);
template __global__ void k_prolong<float>(float*, const float*, int, int, int, int, int, int, int, int);
template __global__ void k_rbgs_zero<float>(float*, const float*, const float*, int, int, int, float, float, float);
template __global__ void k_rbgs_zero<double>(double*, const double*, const double*, int, int, int, double, double,
                                             double);
template __global__ void k_rbgs_prolong<float>(float*, const float*, const float*, const float*, int, int, int, float,
                                               float, float, int, int, int, int, int);
template __global__ void k_rbgs_prolong<double>(double*, const double*, const double*, const double*, int, int, int,
                                                double, double, double, int, int, int, int, int);
template __global__ void k_prolong<double>(double*, const double*, int, int, int, int, int, int, int, int);
template __global__ void k_zm_restrict<float>(const float*, float*, int, int, int, int, int, int, int, int, int, int);
template __global__ void k_zm_restrict3<float>(const float*, float*, int, int, int, int, int, int, int);
template __global__ void k_zm_restrict3<double>(const double*, double*, int, int, int, int, int, int, int);
template __global__ void k_zm_restrict<double>(const double*, double*, int, int, int, int, int, int, int, int, int,
                                               int);
template __global__ void k_zm_prolong<float>(float*, const float*, int, int, int, int, int, int, int, int, int);
template __global__ void k_zm_prolong<double>(double*, const double*, int, int, int, int, int, int, int, int, int);
template __global__ void k_fill<float>(float*, long, float);
template __global__ void k_fill<double>(double*, long, double);
template __global__ void k_cast<float, double>(const float*, double*, long);
template __global__ void k_cast<double, float>(const double*, float*, long);
template __global__ void k_sumsq<float>(const float*, long, double*);
template __global__ void k_sumsq<double>(const double*, long, double*);
#define BCG_RESTRICT_A0(CX, CY, CZ) \
    template __global__ void k_restrict_a0<double, CX, CY, CZ>(const double*, double*, int, int, int, int, int, int);
BCG_RESTRICT_A0(0, 0, 0)
BCG_RESTRICT_A0(0, 0, 1)
BCG_RESTRICT_A0(0, 1, 0)
BCG_RESTRICT_A0(0, 1, 1)
BCG_RESTRICT_A0(1, 0, 0)
BCG_RESTRICT_A0(1, 0, 1)
BCG_RESTRICT_A0(1, 1, 0)
BCG_RESTRICT_A0(1, 1, 1)
#undef BCG_RESTRICT_A0

}  // namespace mg
}  // namespace bcg
