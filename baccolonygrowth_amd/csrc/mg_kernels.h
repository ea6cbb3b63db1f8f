// mg_kernels.h - declarations of the 3-D reaction-diffusion multigrid kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace bcg {
namespace mg {

// One level of the single-workgroup bottom solver (FP64 arrays).
struct BotLevel {
    double *u, *f, *a0, *r;
    int nx, ny, nz;
    int cx, cy, cz;  // coarsening to the next level
    double lx, ly, lz;
};
constexpr int BOT_MAX_LEVELS = 8;
struct BotLevels {
    BotLevel L[BOT_MAX_LEVELS];
    int n;
    // level 0 from FP32 arrays (f32, u32; u32 also receives the result)
    // instead of L[0].f, L[0].u, and/or from u = 0 (zero)
    const float* f32;
    float* u32;
    int zero;
};
__global__ void k_bottom(BotLevels B, int nu1, int nu2, int coarse_sweeps, int lds_from);
constexpr long BOT_LDS_BYTES = 60 * 1024;  // levels kept in LDS when they fit
__global__ void k_sum_chunks(const double* part, long n, double* out);

template <class T>
__global__ void k_rbgs(T* u, const T* f, const T* a0, int nx, int ny, int nz, T lx, T ly, T lz, int color);
template <class T, int AHEAD, class D>
__global__ void k_rbgs2(const T* uin, T* uout, const T* f, const T* a0, int nx, int ny, int nz, T lx, T ly, T lz,
                        int tiles_x, int tiles_y, int kchunk);
template <class T>
__global__ void k_zm_residual(const T* u, const T* f, const T* a0, T* r, int nx, int ny, int nz, T lx, T ly, T lz,
                              int tiles_x, int tiles_y, int kchunk);
__global__ void k_zm_resnorm(const double* u, const double* b, const double* a0, int nx, int ny, int nz, double lx,
                             double ly, double lz, int tiles_x, int tiles_y, int kchunk, double* part);
__global__ void k_zm_ir(const double* xo, double* xn, const float* e, double inv_scale, const double* b,
                        const double* a0, float* r32, double rscale, int nx, int ny, int nz, double lx, double ly,
                        double lz, int tiles_x, int tiles_y, int kchunk, double* part);
template <class T>
__global__ void k_zm_residual2(const T* u, const T* f, const T* a0, T* r, int nx, int ny, int nz, T lx, T ly, T lz,
                               int tiles_x, int tiles_y, int kchunk);
template <class T>
__global__ void k_residual(const T* u, const T* f, const T* a0, T* r, int nx, int ny, int nz, T lx, T ly, T lz);
template <class T>
__global__ void k_restrict(const T* r, T* fc, int nx, int ny, int nz, int ncx, int ncy, int ncz, int cx, int cy,
                           int cz);
template <class T>
__global__ void k_prolong(T* u, const T* ec, int nx, int ny, int nz, int ncx, int ncy, int cx, int cy, int cz);
// small levels: the red sweep from u = 0; the prolongation + the red sweep
// after it
template <class T>
__global__ void k_rbgs_zero(T* u, const T* f, const T* a0, int nx, int ny, int nz, T lx, T ly, T lz);
template <class T>
__global__ void k_rbgs_prolong(T* u, const T* f, const T* a0, const T* ec, int nx, int ny, int nz, T lx, T ly, T lz,
                               int ncx, int ncy, int cx, int cy, int cz);
template <class T>
__global__ void k_zm_restrict(const T* r, T* fc, int nx, int ny, int nz, int ncx, int ncy, int ncz, int cx, int cy, int cz,
                              int kchunk);
template <class T>
__global__ void k_zm_prolong(T* u, const T* ec, int nx, int ny, int nz, int ncx, int ncy, int cx, int cy, int cz,
                             int kchunk);
template <class T>
__global__ void k_zm_restrict3(const T* r, T* fc, int nx, int ny, int nz, int ncx, int ncy, int ncz, int kchunk);
template <class T>
__global__ void k_fill(T* a, long n, T v);
template <class Ts, class Td>
__global__ void k_cast(const Ts* a, Td* b, long n);
template <class T>
__global__ void k_sumsq(const T* x, long n, double* part);
__global__ void k_sum_parts(const double* part, int n, double* out);
__global__ void k_rhs(const double* c, double* b, double c_res, long nxy, long n);
__global__ void k_setup_fine(const double* x, double* cprev, const double* q, double K, double dt, double c_res,
                             long nxy, double* a0, float* a0f, double* b, long n);
// Coarsening factors (0 or 1 per axis) as template arguments: the 27-point
// loop unrolls.
template <class T, int cx, int cy, int cz>
__global__ void k_restrict_a0(const T* a0f, T* a0c, int nx, int ny, int nz, int ncx, int ncy, int ncz);

__global__ void k_adj_rhs(const double* g, double* b, int nx, int ny, int nz);
__global__ void k_adj_grad(const double* mu, const double* c1, const double* c0, const double* q, double K, double dt,
                           double inv_h2, int nx, int ny, int nz, double* dq, double* dc0, double* partK,
                           double* partD);
__global__ void k_cic_gather(const double* cm, int n, double ox, double oy, double oz, double inv_h, int nx, int ny,
                             int nz, const double* field, double inv_vol, double* out);

}  // namespace mg
}  // namespace bcg
