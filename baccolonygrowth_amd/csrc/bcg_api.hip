// bcg_api.hip - host side of the C ABI declared in include/bcg.h.
//
// Owns the device state of one colony and drives the kernels of
// mech_kernels.hip on one HIP stream. Cell state in HBM:
//   hot[2][maxCells][16] doubles  p q v w L R Ldiv GR   (double buffered,
//                                 one 128-byte line per cell for the gathers)
//   cold SoA                       time_p time_q age_p age_q Type Ancestor
//                                  DynFric StaFric
//   bins                           key, count, start, fill, sorted (counting sort)
//   lists (ELL)                    nl[maxNeighbours][stride] by slot, nl_count[slot],
//                                  order[slot] -> ID, slot_of[ID] -> slot
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bcg.h"
#include "glibc_rand.h"
#include "mech_kernels.h"

using namespace bcg;

#include "ctx_internal.h"

namespace {

int fail(bcg_ctx* c, int code, const char* fmt, ...) {
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
int dalloc(bcg_ctx* c, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(c, BCG_ERR_HIP, "hipMalloc(%zu): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

PhysParams phys(const bcg_params& p) {
    PhysParams P;
    P.k_cc = p.k_cc;
    P.k_wc = p.k_wc;
    P.gamma_t = p.gamma_t;
    P.cell_mu = p.cell_mu;
    P.wall_mu = p.wall_mu;
    P.viscosity = p.viscosity;
    P.tension = p.tension;
    P.DH = p.DH;
    P.cellRadius = p.cellRadius;
    P.BoxLength = p.BoxLength;
    P.Rc = p.Rc;
    P.L_divide = p.L_divide;
    P.varL = p.varL;
    P.varAngle = p.varAngle;
    P.BoxX = p.BoxX;
    P.BoxY = p.BoxY;
    P.BoxZ = p.BoxZ;
    P.ref = p.refinementGridHeight;
    P.nhx = p.BoxX * p.refinementGridHeight;
    P.nhy = p.BoxY * p.refinementGridHeight;
    P.planar = p.planar;
    P.crossFeeding = p.crossFeeding;
    return P;
}

// Exclusive scan of n ints (in -> out); *grand (device) receives the total.
int scan(bcg_ctx* c, const int* in, int n, int* out, int* grand) {
    const int nb = nblk(n, TSCAN_ITEMS);
    if (nb > 1024 * 64) return fail(c, BCG_ERR_ARG, "scan of %d items too large", n);
    hipLaunchKernelGGL(k_tscan_sums, dim3(nb), dim3(TSCAN_T), 0, c->stream, in, n, c->scan_tot);
    hipLaunchKernelGGL(k_tscan_apply, dim3(nb), dim3(TSCAN_T), 0, c->stream, in, n, c->scan_tot, out, grand);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Counting sort of cells 0..n-1 of hot[buf] into bins.
int build_bins(bcg_ctx* c, int buf, int n) {
    HIPCHK(c, hipMemsetAsync(c->bin_cnt, 0, sizeof(int) * (c->nbin + 1), c->stream));
    if (n > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->hot[buf], n, c->P, c->key,
                           c->bin_cnt, c->bin_slot, c->cmL, &c->stats->err);
    int rc = scan(c, c->bin_cnt, c->nbin + 1, c->bin_start, nullptr);
    if (rc) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_place, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->key, c->bin_slot, n,
                           c->bin_start, c->scattered);
        hipLaunchKernelGGL(k_bin_rank, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->scattered, n, c->key,
                           c->bin_start, c->cmL, c->sorted, c->cand);
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

int gather_all(bcg_ctx* c) {
    if (c->N > 0) {
        const int nb = nblk(c->N, 256);
        hipLaunchKernelGGL(k_gather_all, dim3(nb), dim3(256), 0, c->stream, c->hot[c->cur], c->N, c->P, c->bin_start,
                           c->sorted, c->cand, c->nl, c->nl_count, c->stride, c->cap, c->order, c->slot_of,
                           c->part_max, &c->stats->err);
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, c->stream, c->part_max, nb, &c->stats->maxlist);
    }
    HIPCHK(c, hipGetLastError());
    c->lists_valid = true;
    return 0;
}

int reset_stats(bcg_ctx* c) {
    HIPCHK(c, hipMemsetAsync(c->stats, 0, sizeof(BcgStats), c->stream));
    return 0;
}

int read_stats(bcg_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->hstats, c->stats, sizeof(BcgStats), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->last = *c->hstats;
    return 0;
}

int check_err(bcg_ctx* c) {
    if (c->last.err & ERRBIT_DOMAIN)
        return fail(c, BCG_ERR_OUT_OF_DOMAIN, "a cell centre lies outside the %dx%dx%d bin grid", c->prm.BoxX,
                    c->prm.BoxY, c->prm.BoxZ);
    if (c->last.err & ERRBIT_NLIST)
        return fail(c, BCG_ERR_NEIGHBOUR_OVERFLOW, "a neighbour list exceeds maxNeighbours=%d", c->cap);
    return 0;
}

int set_identity_slots(bcg_ctx* c) {
    std::vector<int> idn(c->prm.maxCells);
    for (int i = 0; i < c->prm.maxCells; i++) idn[i] = i;
    HIPCHK(c, hipMemcpyAsync(c->order, idn.data(), sizeof(int) * idn.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->slot_of, idn.data(), sizeof(int) * idn.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

int bcg_bin_current(bcg_ctx* c) {
    if (int rc_ = ctx_check_failed(c)) return rc_;
    HIPCHK(c, ctx_join(c));
    int rc = reset_stats(c);
    if (!rc) rc = build_bins(c, c->cur, c->N);
    if (!rc) rc = read_stats(c);
    if (!rc) rc = check_err(c);
    return rc;
}

extern "C" {

int bcg_default_params(bcg_params* p) {
    if (!p) return BCG_ERR_ARG;
    memset(p, 0, sizeof *p);
    // ABM/Constants.cpp defaults overlaid with run_sim/inputFile_20mM_SmallSimulation.txt
    p->cellRadius = 0.5;
    p->L_divide = 3.0;
    p->k_cc = 200000.0;
    p->k_wc = 200000.0;
    p->varL = 0.25;
    p->varAngle = 0.001;
    p->var_pos = 0.0;
    p->viscosity = 0.1;
    p->wall_rough = 0.1;
    p->gamma_t = 5000.0;
    p->wall_mu = 0.4;
    p->cell_mu = 0.1;
    p->density_threshold = 0.55;
    p->tension = 500.0;
    p->DH = -0.01;
    p->Rc = 3.0;
    p->BoxX = 256;
    p->BoxY = 1;
    p->BoxZ = 50;
    p->refinementGridHeight = 1;
    p->maxCells = 5000;
    p->maxNeighbours = 64;
    p->crossFeeding = 0;
    p->planar = 1;
    p->BoxLength = p->L_divide + 2 * p->cellRadius;
    return BCG_OK;
}

int bcg_create(const bcg_params* p, bcg_ctx** out) {
    if (!p || !out) return BCG_ERR_ARG;
    *out = nullptr;
    if (p->maxCells <= 0 || p->BoxX <= 0 || p->BoxY <= 0 || p->BoxZ <= 0 || p->maxNeighbours <= 0 ||
        p->refinementGridHeight <= 0 || !(p->BoxLength > 0))
        return BCG_ERR_ARG;
    if (p->var_pos != 0.0) return BCG_ERR_UNSUPPORTED;  // F_cw draw values would enter the forces
    if ((long)p->BoxX * p->BoxY * p->BoxZ > (1L << 30)) return BCG_ERR_ARG;
    // the list build's queue entries hold bin-order positions in QK_BITS bits
    if (p->maxCells > QK_MASK + 1) return BCG_ERR_ARG;
    bcg_ctx* c = new bcg_ctx();
    c->prm = *p;
    c->P = phys(*p);
    c->cap = p->maxNeighbours;
    c->stride = (p->maxCells + 63) / 64 * 64;
    c->nbin = p->BoxX * p->BoxY * p->BoxZ;
    const size_t M = (size_t)p->maxCells;
    int rc = 0;
    auto A = [&](auto** ptr, size_t n) {
        if (!rc) rc = dalloc(c, ptr, n);
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BCG_ERR_HIP;
    }
    A(&c->hot[0], M * HOT);
    A(&c->hot[1], M * HOT);
    A(&c->time_p, M);
    A(&c->time_q, M);
    A(&c->fric, 6 * M);
    A(&c->fric_id, M);
    A(&c->age_p, M);
    A(&c->age_q, M);
    A(&c->type, M);
    A(&c->ancestor, M);
    A(&c->key, M);
    A(&c->bin_cnt, c->nbin + 1);
    A(&c->bin_start, c->nbin + 1);
    A(&c->bin_slot, M);
    A(&c->sorted, M);
    A(&c->scattered, M);
    A(&c->cmL, 2 * (size_t)M);   // FP32 end points p, q per cell
    A(&c->cand, 2 * ((size_t)M + WALK_GROUP));  // + the walk's read-ahead
    A(&c->scan_tot, 1024 * 64);
    A(&c->nl, (size_t)c->cap * c->stride);
    A(&c->nl_count, M);
    A(&c->order, M);
    A(&c->slot_of, M);
    A(&c->divbits, (size_t)nblk(M, 32));
    A(&c->div_ids, M);
    A(&c->div_index_of, M);
    A(&c->ids_tmp, M);
    A(&c->draws, 9 * M);  // rand() values of the divisions' draws
    A(&c->mother_pre, M * HOT);
    const size_t nh = (size_t)c->P.nhx * c->P.nhy;
    A(&c->height, nh);
    A(&c->normal, 3 * nh);
    A(&c->wall, (size_t)p->BoxX * p->BoxY);
    A(&c->gr1, c->nbin);
    A(&c->gr2, c->nbin);
    A(&c->Fbuf, 3 * M);
    A(&c->Tbuf, 3 * M);
    A(&c->part_max, (size_t)nblk(M, 256));
    A(&c->part_draws, (size_t)nblk(M, std::min(MOVE_BLOCK, 64)));  // per block or per wave
    A(&c->Sbuf, 5 * M);
    A(&c->stats, 1);
    if (!rc && hipHostMalloc((void**)&c->hstats, sizeof(BcgStats), hipHostMallocDefault) != hipSuccess)
        rc = BCG_ERR_HIP;
    for (int i = 0; i < 7 && !rc; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = BCG_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->after, hipEventDisableTiming) != hipSuccess) rc = BCG_ERR_HIP;
    if (rc) {
        bcg_destroy(c);
        return rc;
    }
    // default fields: flat surface at cellRadius, normal -z, flat wall, no growth
    std::vector<double> h(nh, p->cellRadius), n3(3 * nh, 0.0);
    for (size_t i = 0; i < nh; i++) n3[3 * i + 2] = -1.0;
    hipMemcpy(c->height, h.data(), nh * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(c->normal, n3.data(), 3 * nh * sizeof(double), hipMemcpyHostToDevice);
    hipMemset(c->wall, 0, sizeof(double) * p->BoxX * p->BoxY);
    hipMemset(c->gr1, 0, sizeof(double) * c->nbin);
    hipMemset(c->gr2, 0, sizeof(double) * c->nbin);
    std::vector<int> neg(M, -1);
    hipMemcpy(c->div_index_of, neg.data(), M * sizeof(int), hipMemcpyHostToDevice);
    hipMemset(c->nl_count, 0, M * sizeof(int));
    hipMemset(c->divbits, 0, nblk(M, 32) * sizeof(unsigned));
    hipMemset(c->hot[0], 0, M * HOT * sizeof(double));
    hipMemset(c->hot[1], 0, M * HOT * sizeof(double));
    c->rng.seed(1);
    if (set_identity_slots(c)) {
        bcg_destroy(c);
        return BCG_ERR_HIP;
    }
    *out = c;
    return BCG_OK;
}

int bcg_destroy(bcg_ctx* c) {
    if (!c) return BCG_OK;
    if (c->stream) {
        ctx_join(c);
        hipStreamSynchronize(c->stream);
    }
    void* ptrs[] = {c->hot[0],   c->hot[1],    c->time_p,  c->time_q,     c->fric,      c->fric_id,  c->age_p,
                    c->age_q,    c->type,      c->ancestor, c->key,       c->bin_cnt,   c->bin_start, c->bin_slot,
                    c->sorted,   c->scattered, c->scan_tot, c->cmL, c->cand,  c->nl,      c->nl_count,   c->order,     c->slot_of,  c->divbits,
                    c->div_ids,  c->div_index_of, c->ids_tmp, c->draws,   c->mother_pre, c->height,  c->normal,
                    c->wall,     c->gr1,       c->gr2,     c->Fbuf,       c->Tbuf,      c->Sbuf,      c->stats,
                    c->part_max, c->part_draws};
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (c->hstats) hipHostFree(c->hstats);
    if (c->hids) hipHostFree(c->hids);
    if (c->hdraws) hipHostFree(c->hdraws);
    for (int i = 0; i < 7; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->after) hipEventDestroy(c->after);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return BCG_OK;
}

int bcg_last_error(const bcg_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return BCG_ERR_ARG;
    snprintf(buf, len, "%s", c->err.c_str());
    return BCG_OK;
}

int bcg_set_cells(bcg_ctx* c, const bcg_cell* cells, int32_t n) {
    if (!c || (n > 0 && !cells) || n < 0) return BCG_ERR_ARG;
    if (n > c->prm.maxCells) return fail(c, BCG_ERR_CAPACITY, "%d cells > maxCells %d", n, c->prm.maxCells);
    HIPCHK(c, ctx_join(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the copies below are synchronous on the null stream
    std::vector<double> hot((size_t)n * HOT), tp(n), tq(n), fr(6 * (size_t)n);
    std::vector<int> ap(n), aq(n), ty(n), an(n), ident(n);
    for (int i = 0; i < n; i++) {
        const bcg_cell& s = cells[i];
        double* h = &hot[(size_t)i * HOT];
        for (int k = 0; k < 3; k++) {
            h[k] = s.p[k];
            h[3 + k] = s.q[k];
            h[6 + k] = s.Velocity[k];
            h[9 + k] = s.AngularVelocity[k];
            fr[6 * (size_t)i + k] = s.DynFric[k];
            fr[6 * (size_t)i + 3 + k] = s.StaFric[k];
        }
        ident[i] = i;
        h[12] = s.Length;
        h[13] = s.Radius;
        h[14] = s.Ldiv;
        h[15] = s.GrowthRate;
        tp[i] = s.time_p;
        tq[i] = s.time_q;
        ap[i] = s.age_p;
        aq[i] = s.age_q;
        ty[i] = s.Type;
        an[i] = s.Ancestor;
    }
    c->cur = 0;
    c->N = n;
    c->failed = false;
    auto up = [&](void* d, const void* h, size_t b) { return hipMemcpy(d, h, b, hipMemcpyHostToDevice); };
    if (n > 0) {
        HIPCHK(c, up(c->hot[0], hot.data(), hot.size() * sizeof(double)));
        HIPCHK(c, up(c->time_p, tp.data(), n * sizeof(double)));
        HIPCHK(c, up(c->time_q, tq.data(), n * sizeof(double)));
        HIPCHK(c, up(c->fric, fr.data(), fr.size() * sizeof(double)));
        HIPCHK(c, up(c->fric_id, ident.data(), n * sizeof(int)));
        HIPCHK(c, up(c->age_p, ap.data(), n * sizeof(int)));
        HIPCHK(c, up(c->age_q, aq.data(), n * sizeof(int)));
        HIPCHK(c, up(c->type, ty.data(), n * sizeof(int)));
        HIPCHK(c, up(c->ancestor, an.data(), n * sizeof(int)));
    }
    HIPCHK(c, hipMemset(c->nl_count, 0, sizeof(int) * c->prm.maxCells));
    HIPCHK(c, hipMemset(c->divbits, 0, sizeof(unsigned) * nblk(c->prm.maxCells, 32)));
    c->lists_valid = true;  // empty lists, as InitializeNeighbourList leaves them
    int rc = set_identity_slots(c);
    if (rc) return rc;
    // place every cell in its bin (LoadCells, ABM/InputOutput.cpp:422-426): fails on a cell outside the grid
    if ((rc = reset_stats(c))) return rc;
    if ((rc = build_bins(c, c->cur, c->N))) return rc;
    if ((rc = read_stats(c))) return rc;
    return check_err(c);
}

}  // extern "C"

namespace {

// Cells 0..N-1 of hot buffer `buf` with the cold fields, in the reference's
// record (the context's stream is idle when this is called).
int download_cells(bcg_ctx* c, int buf, int N, bcg_cell* cells) {
    if (N == 0) return BCG_OK;
    std::vector<double> hot((size_t)N * HOT), tp(N), tq(N), fr(6 * (size_t)N);
    std::vector<int> ap(N), aq(N), ty(N), an(N), fid(N);
    auto dn = [&](void* h, const void* d, size_t b) { return hipMemcpy(h, d, b, hipMemcpyDeviceToHost); };
    HIPCHK(c, dn(hot.data(), c->hot[buf], hot.size() * sizeof(double)));
    HIPCHK(c, dn(tp.data(), c->time_p, N * sizeof(double)));
    HIPCHK(c, dn(tq.data(), c->time_q, N * sizeof(double)));
    HIPCHK(c, dn(fr.data(), c->fric, fr.size() * sizeof(double)));
    HIPCHK(c, dn(fid.data(), c->fric_id, N * sizeof(int)));
    HIPCHK(c, dn(ap.data(), c->age_p, N * sizeof(int)));
    HIPCHK(c, dn(aq.data(), c->age_q, N * sizeof(int)));
    HIPCHK(c, dn(ty.data(), c->type, N * sizeof(int)));
    HIPCHK(c, dn(an.data(), c->ancestor, N * sizeof(int)));
    // frictions are held by slot: slot t belongs to cell fid[t]
    for (int t = 0; t < N; t++) {
        const int i = fid[t];
        if (i < 0 || i >= N) return fail(c, BCG_ERR_HIP, "friction slot %d names cell %d", t, i);
        for (int k = 0; k < 3; k++) {
            cells[i].DynFric[k] = fr[6 * (size_t)t + k];
            cells[i].StaFric[k] = fr[6 * (size_t)t + 3 + k];
        }
    }
    for (int i = 0; i < N; i++) {
        bcg_cell& s = cells[i];
        const double* h = &hot[(size_t)i * HOT];
        for (int k = 0; k < 3; k++) {
            s.p[k] = h[k];
            s.q[k] = h[3 + k];
            s.Velocity[k] = h[6 + k];
            s.AngularVelocity[k] = h[9 + k];
        }
        s.Length = h[12];
        s.Radius = h[13];
        s.Ldiv = h[14];
        s.GrowthRate = h[15];
        s.time_p = tp[i];
        s.time_q = tq[i];
        s.age_p = ap[i];
        s.age_q = aq[i];
        s.Type = ty[i];
        s.Ancestor = an[i];
    }
    return BCG_OK;
}

}  // namespace

extern "C" {

int bcg_get_cells(bcg_ctx* c, bcg_cell* cells, int32_t cap, int32_t* n) {
    if (!c || !n) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(c)) return rc_;
    *n = c->N;
    if (cap < c->N || (c->N > 0 && !cells)) return fail(c, BCG_ERR_ARG, "buffer of %d cells < %d", cap, c->N);
    HIPCHK(c, ctx_join(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return download_cells(c, c->cur, c->N, cells);
}

int bcg_num_cells(bcg_ctx* c, int32_t* n) {
    if (!c || !n) return BCG_ERR_ARG;
    *n = c->N;
    return BCG_OK;
}

int bcg_set_fields(bcg_ctx* c, const double* height, const double* normal, const double* wall, const double* gr1,
                   const double* gr2) {
    if (!c) return BCG_ERR_ARG;
    HIPCHK(c, ctx_join(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the copies below are synchronous on the null stream
    const size_t nh = (size_t)c->P.nhx * c->P.nhy;
    if (height) HIPCHK(c, hipMemcpy(c->height, height, nh * sizeof(double), hipMemcpyHostToDevice));
    if (normal) HIPCHK(c, hipMemcpy(c->normal, normal, 3 * nh * sizeof(double), hipMemcpyHostToDevice));
    if (wall)
        HIPCHK(c, hipMemcpy(c->wall, wall, sizeof(double) * c->prm.BoxX * c->prm.BoxY, hipMemcpyHostToDevice));
    if (gr1) HIPCHK(c, hipMemcpy(c->gr1, gr1, sizeof(double) * c->nbin, hipMemcpyHostToDevice));
    if (gr2) HIPCHK(c, hipMemcpy(c->gr2, gr2, sizeof(double) * c->nbin, hipMemcpyHostToDevice));
    return BCG_OK;
}

int bcg_get_fields(bcg_ctx* c, double* height, double* normal, double* wall, double* gr1, double* gr2) {
    if (!c) return BCG_ERR_ARG;
    HIPCHK(c, ctx_join(c));
    const size_t nh = (size_t)c->P.nhx * c->P.nhy;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (height) HIPCHK(c, hipMemcpy(height, c->height, nh * sizeof(double), hipMemcpyDeviceToHost));
    if (normal) HIPCHK(c, hipMemcpy(normal, c->normal, 3 * nh * sizeof(double), hipMemcpyDeviceToHost));
    if (wall)
        HIPCHK(c, hipMemcpy(wall, c->wall, sizeof(double) * c->prm.BoxX * c->prm.BoxY, hipMemcpyDeviceToHost));
    if (gr1) HIPCHK(c, hipMemcpy(gr1, c->gr1, sizeof(double) * c->nbin, hipMemcpyDeviceToHost));
    if (gr2) HIPCHK(c, hipMemcpy(gr2, c->gr2, sizeof(double) * c->nbin, hipMemcpyDeviceToHost));
    return BCG_OK;
}

int bcg_srand(bcg_ctx* c, uint32_t seed) {
    if (!c) return BCG_ERR_ARG;
    c->rng.seed(seed);
    return BCG_OK;
}

int bcg_rand(bcg_ctx* c, int32_t* out) {
    if (!c || !out) return BCG_ERR_ARG;
    *out = c->rng.next();
    return BCG_OK;
}

int bcg_rand_stream(uint32_t seed, uint64_t skip, int32_t n, int32_t* out) {
    if (n < 0 || (n > 0 && !out)) return BCG_ERR_ARG;
    GlibcRand g;
    g.seed(seed);
    g.discard(skip);
    // one draw at a time, in bulk, one at a time again (the same stream)
    const int a = n / 3, b = 2 * n / 3;
    for (int i = 0; i < a; i++) out[i] = g.next();
    g.fill(out + a, (size_t)(b - a));
    for (int i = b; i < n; i++) out[i] = g.next();
    return BCG_OK;
}

int bcg_rebuild_neighbours(bcg_ctx* c, const int32_t* ids, int32_t n) {
    if (!c) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(c)) return rc_;
    HIPCHK(c, ctx_join(c));
    int rc;
    if ((rc = reset_stats(c))) return rc;
    if ((rc = build_bins(c, c->cur, c->N))) return rc;
    if (!ids || n < 0) {
        if ((rc = gather_all(c))) return rc;
    } else if (n > 0) {
        for (int i = 0; i < n; i++)
            if (ids[i] < 0 || ids[i] >= c->N) return fail(c, BCG_ERR_ARG, "cell id %d out of range", ids[i]);
        if (n > c->prm.maxCells) return BCG_ERR_ARG;
        // the listed cells are gathered sequentially by the reference; the
        // grid does not change in between, so they are independent. A cell
        // listed twice is gathered once.
        std::vector<int> u(ids, ids + n);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        HIPCHK(c, hipMemcpyAsync(c->ids_tmp, u.data(), u.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_gather_ids, dim3(nblk((long)u.size(), 256)), dim3(256), 0, c->stream, c->hot[c->cur],
                           c->ids_tmp, (int)u.size(), c->P, c->bin_start, c->sorted, c->cand, c->nl, c->nl_count,
                           c->stride,
                           c->cap, c->slot_of, &c->stats->maxlist, &c->stats->err);
        HIPCHK(c, hipGetLastError());
    }
    if ((rc = read_stats(c))) return rc;
    return check_err(c);
}

int bcg_get_neighbours(bcg_ctx* c, int32_t* counts, int32_t* lists, int32_t stride) {
    if (!c || !counts || (!lists && stride > 0) || stride < 0) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(c)) return rc_;
    HIPCHK(c, ctx_join(c));
    const int N = c->N;
    if (N == 0) return BCG_OK;
    std::vector<int> cnt(N), slot(N), nl((size_t)c->cap * N);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(cnt.data(), c->nl_count, N * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(slot.data(), c->slot_of, N * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy2D(nl.data(), N * sizeof(int), c->nl, c->stride * sizeof(int), N * sizeof(int), c->cap,
                          hipMemcpyDeviceToHost));
    for (int id = 0; id < N; id++) {
        const int s = slot[id];
        const int m = cnt[s];
        counts[id] = m;
        std::vector<int> L(m);
        for (int k = 0; k < m; k++) L[k] = nl[(size_t)k * N + s];
        std::sort(L.begin(), L.end());
        for (int k = 0; k < m && k < stride; k++) lists[(size_t)id * stride + k] = L[k];
    }
    return BCG_OK;
}

int bcg_sum_forces(bcg_ctx* c, double* F, double* T) {
    if (!c || !F || !T) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(c)) return rc_;
    HIPCHK(c, ctx_join(c));
    int rc;
    if ((rc = reset_stats(c))) return rc;
    if (c->N > 0)
        hipLaunchKernelGGL(k_sum_forces, dim3(nblk(c->N, MOVE_BLOCK)), dim3(MOVE_BLOCK), 0, c->stream, c->hot[c->cur],
                           c->N, c->P, c->order, c->nl, c->nl_count, c->stride, c->height, c->normal, c->wall,
                           c->Fbuf, c->Tbuf, &c->stats->err);
    HIPCHK(c, hipGetLastError());
    if ((rc = read_stats(c))) return rc;
    if ((rc = check_err(c))) return rc;
    if (c->N > 0) {
        HIPCHK(c, hipMemcpy(F, c->Fbuf, 3 * sizeof(double) * c->N, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(T, c->Tbuf, 3 * sizeof(double) * c->N, hipMemcpyDeviceToHost));
    }
    return BCG_OK;
}

int bcg_mean_stress(bcg_ctx* c, double* stress, double* fnet) {
    if (!c || (c->N > 0 && (!stress || !fnet))) return BCG_ERR_ARG;
    if (int rc_ = ctx_check_failed(c)) return rc_;
    HIPCHK(c, ctx_join(c));
    if (!c->lists_valid && c->N > 0) return fail(c, BCG_ERR_STATE, "no neighbour lists: rebuild first");
    int rc;
    if ((rc = reset_stats(c))) return rc;
    if (c->N > 0) {
        const int nw = nblk(c->N, 64);
        HIPCHK(c, hipMemsetAsync(c->part_draws, 0, sizeof(unsigned) * nw, c->stream));
        hipLaunchKernelGGL(k_mean_stress, dim3(nblk(c->N, MOVE_BLOCK)), dim3(MOVE_BLOCK), 0, c->stream,
                           c->hot[c->cur], c->N, c->P, c->order, c->nl, c->nl_count, c->stride, c->height, c->normal,
                           c->wall, c->Sbuf, c->Fbuf, c->part_draws, &c->stats->err);
        hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, c->stream, c->part_draws, nw, &c->stats->wall_draws);
    }
    HIPCHK(c, hipGetLastError());
    if ((rc = read_stats(c))) return rc;
    if ((rc = check_err(c))) return rc;
    // every F_cw call took one draw of the stream (ABM/Forces.cpp:102)
    c->rng.discard(c->last.wall_draws);
    if (c->N > 0) {
        HIPCHK(c, hipMemcpy(stress, c->Sbuf, 5 * sizeof(double) * c->N, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(fnet, c->Fbuf, 3 * sizeof(double) * c->N, hipMemcpyDeviceToHost));
    }
    return BCG_OK;
}

}  // extern "C"

namespace {

// The body of bcg_step; *started is set once the step has written state that
// a failure cannot take back (move/grow's friction accumulators, the other
// cell buffer, then the rng position and the lists).
int step_body(bcg_ctx* c, double dt, double t, int32_t flags, int32_t* n_div, int32_t* lineage, int32_t lineage_cap,
              bcg_cell* moved, bool* started) {
    int rc;
    const int N0 = c->N;
    const int in = c->cur, out = c->cur ^ 1;
    if ((rc = reset_stats(c))) return rc;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    if (flags & 1) {
        if ((rc = build_bins(c, in, N0))) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    if (flags & 1) {
        if ((rc = gather_all(c))) return rc;
    } else if (!c->lists_valid && N0 > 0) {
        return fail(c, BCG_ERR_STATE, "no neighbour lists: rebuild before stepping");
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // the growth rates and the other cell buffer may still be in use by work
    // queued on another stream (a nutrient update): bins and lists above did
    // not touch them and ran alongside it
    HIPCHK(c, ctx_join(c));
    HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
    *started = true;
    // MoveCell + GrowCell for every cell (ABM/Simulation.cpp:330-335)
    if (N0 > 0) {
        const int nb = nblk(N0, MOVE_BLOCK);
        hipLaunchKernelGGL(k_move_grow, dim3(nb), dim3(MOVE_BLOCK), 0, c->stream, c->hot[in], c->hot[out], N0, c->P,
                           dt, c->order, c->nl, c->nl_count, c->stride, c->height, c->normal, c->wall, c->gr1, c->gr2,
                           c->type, c->fric, c->fric_id, c->divbits, c->part_draws, &c->stats->err);
        hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(1024), 0, c->stream, c->part_draws, nb,
                           &c->stats->wall_draws);
    }
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    if (N0 > 0) {
        const int nw = nblk(N0, 32), nb = nblk(nw, 256);
        hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(256), 0, c->stream, c->divbits, nw, c->ids_tmp);
        if ((rc = scan(c, c->ids_tmp, nb, c->ids_tmp, &c->stats->ndiv))) return rc;
        hipLaunchKernelGGL(k_flag_write, dim3(nb), dim3(256), 0, c->stream, c->divbits, nw, c->ids_tmp, c->div_ids);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    if ((rc = read_stats(c))) return rc;
    if ((rc = check_err(c))) return rc;
    // the moved and grown cells before the divisions (what the reference's
    // output step saves, ABM/Simulation.cpp:349)
    if (moved && (rc = download_cells(c, out, N0, moved))) return rc;
    // draws: every F_cw call took one (values unused: var_pos == 0), then
    // 9 per division in ascending mother ID (ABM/Simulation.cpp:380-404)
    c->rng.discard(c->last.wall_draws);
    c->last_flagged = c->last.ndiv;
    int D = std::min(c->last.ndiv, c->prm.maxCells - N0);
    if (D < 0) D = 0;
    if (D > 0) {
        // the 9 draws of each division as rand() values (mapped on the device),
        // from pinned memory (the previous step's copy has completed)
        const size_t nd = (size_t)9 * D;
        if (nd > c->hdraws_cap) {
            if (c->hdraws) HIPCHK(c, hipHostFree(c->hdraws));
            c->hdraws = nullptr;
            c->hdraws_cap = 0;
            const size_t cap = std::max(nd, (size_t)9 * 1024);
            HIPCHK(c, hipHostMalloc((void**)&c->hdraws, sizeof(int32_t) * cap, hipHostMallocDefault));
            c->hdraws_cap = cap;
        }
        c->rng.fill(c->hdraws, nd);
        HIPCHK(c, hipMemcpyAsync(c->draws, c->hdraws, nd * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_divide, dim3(nblk(D, 64)), dim3(64), 0, c->stream, c->hot[out], N0, D, c->div_ids,
                           c->draws, c->P, t, c->mother_pre, c->div_index_of, c->time_p, c->time_q, c->age_p,
                           c->age_q, c->type, c->ancestor, c->fric, c->fric_id, c->order, c->slot_of,
                           c->nl_count);
        HIPCHK(c, hipGetLastError());
        if (!(flags & 2)) {
            // list patch of ABM/Simulation.cpp:395-403 against the final grid
            if ((rc = build_bins(c, out, N0 + D))) return rc;
            Versioned V{c->hot[out], c->mother_pre, c->div_index_of, N0};
            hipLaunchKernelGGL(k_patch, dim3(1), dim3(PATCH_BLOCK), (c->cap + 2) * sizeof(int), c->stream, c->P, V, D,
                               c->div_ids, c->bin_start, c->sorted, c->nl, c->nl_count, c->stride, c->cap,
                               c->slot_of, &c->stats->maxlist, &c->stats->err);
        } else {
            c->lists_valid = false;
        }
        hipLaunchKernelGGL(k_reset_div_index, dim3(nblk(D, 256)), dim3(256), 0, c->stream, c->div_index_of,
                           c->div_ids, D);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    // the mothers' IDs come back with the stats (one host synchronisation)
    const int nl = lineage ? std::min(D, lineage_cap) : 0;
    if (nl > 0) {
        if (nl > c->hids_cap) {
            if (c->hids) HIPCHK(c, hipHostFree(c->hids));
            c->hids = nullptr;
            c->hids_cap = 0;
            const int cap = std::max(nl, 1024);
            HIPCHK(c, hipHostMalloc((void**)&c->hids, sizeof(int) * cap, hipHostMallocDefault));
            c->hids_cap = cap;
        }
        HIPCHK(c, hipMemcpyAsync(c->hids, c->div_ids, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    if ((rc = read_stats(c))) return rc;
    if ((rc = check_err(c))) return rc;
    for (int i = 0; i < 5; i++) hipEventElapsedTime(&c->last_ms[i], c->ev[i == 2 ? 6 : i], c->ev[i + 1]);
    for (int k = 0; k < nl; k++) {
        lineage[2 * k] = c->hids[k];
        lineage[2 * k + 1] = N0 + k;
    }
    c->N = N0 + D;
    c->cur = out;
    if (n_div) *n_div = D;
    return BCG_OK;
}

}  // namespace

extern "C" {

int bcg_step(bcg_ctx* c, double dt, double t, int32_t flags, int32_t* n_div, int32_t* lineage, int32_t lineage_cap) {
    if (!c) return BCG_ERR_ARG;
    if (n_div) *n_div = 0;
    if (int rc = ctx_check_failed(c)) return rc;
    bool started = false;
    const int rc = step_body(c, dt, t, flags, n_div, lineage, lineage_cap, nullptr, &started);
    if (rc && started) {
        // the friction accumulators and possibly the rng, the other buffer and
        // the lists have moved on: nothing but bcg_set_cells may follow
        c->failed = true;
        c->lists_valid = false;
    }
    return rc;
}

int bcg_step_out(bcg_ctx* c, double dt, double t, int32_t flags, bcg_cell* moved, int32_t moved_cap, int32_t* n_div,
                 int32_t* lineage, int32_t lineage_cap) {
    if (!c) return BCG_ERR_ARG;
    if (n_div) *n_div = 0;
    if (int rc = ctx_check_failed(c)) return rc;
    if (moved && moved_cap < c->N) return fail(c, BCG_ERR_ARG, "buffer of %d cells < %d", moved_cap, c->N);
    bool started = false;
    const int rc = step_body(c, dt, t, flags, n_div, lineage, lineage_cap, moved, &started);
    if (rc && started) {
        c->failed = true;
        c->lists_valid = false;
    }
    return rc;
}

int bcg_last_step_stats(bcg_ctx* c, int64_t* wall_draws, int32_t* n_flagged, int32_t* max_list) {
    if (!c) return BCG_ERR_ARG;
    if (wall_draws) *wall_draws = (int64_t)c->last.wall_draws;
    if (n_flagged) *n_flagged = c->last_flagged;
    if (max_list) *max_list = c->last.maxlist;
    return BCG_OK;
}

int bcg_last_step_times(bcg_ctx* c, float* ms5) {
    if (!c || !ms5) return BCG_ERR_ARG;
    for (int i = 0; i < 5; i++) ms5[i] = c->last_ms[i];
    return BCG_OK;
}

}  // extern "C"
