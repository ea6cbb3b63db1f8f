// mech_kernels.h - kernel declarations shared by mech_kernels.hip and the
// host side (bcg_api.hip).
#pragma once

#include "mech_device.h"

namespace bcg {

constexpr int ERRBIT_DOMAIN = 1;  // a centre of mass outside the bin grid
constexpr int ERRBIT_NLIST = 2;   // a reduced list longer than its capacity

constexpr int MOVE_BLOCK = 64;
constexpr int PATCH_BLOCK = 256;
constexpr int SCAN_BLOCK_ITEMS = 2048;  // items per block of the scan kernels
// The list build reads candidate records in groups of this many from any
// position below the cell count: the candidate array holds as many records
// beyond its last cell (never used for a decision).
constexpr int WALK_GROUP = 32;
constexpr int WALK_STEP = 8;  // candidates a step of the list build's walk (divides WALK_GROUP)
// Queue entries of the list build: the candidate's position in bin order (28
// bits, so at most 2^28 cells) and its bin column relative to the owner's,
// (dx + 1) * 3 + (dy + 1).
constexpr int QK_BITS = 28;
constexpr int QK_MASK = (1 << QK_BITS) - 1;

__global__ void k_scan_reduce(const int* in, int n, int* tot);
__global__ void k_scan_totals(int* tot, int nb, int* grand);
__global__ void k_scan_apply(const int* in, int n, const int* tot, int* out);
// the two-launch scan (k_tscan_sums, then k_tscan_apply): tiles of TSCAN_ITEMS
constexpr int TSCAN_T = 256, TSCAN_ITEMS = 4096;
__global__ void k_tscan_sums(const int* in, int n, int* tot);
__global__ void k_tscan_apply(const int* in, int n, const int* tot, int* out, int* grand);

__global__ void k_bin_count(const double* hot, int n, PhysParams P, int* key, int* cnt, int* slot, float4* cmL,
                            int* err);
__global__ void k_bin_place(const int* key, const int* slot, int n, const int* start, int* scattered);
__global__ void k_bin_rank(const int* scattered, int n, const int* key, const int* start, const float4* cmL,
                           int* sorted, float4* cand);

__global__ void k_gather_all(const double* hot, int n, PhysParams P, const int* start, const int* sorted,
                             const float4* cand, int* nl, int* nl_count, int stride, int cap, int* order,
                             int* slot_of, int* block_max, int* err);
__global__ void k_reduce_max(const int* part, int n, int* out);
__global__ void k_reduce_sum(const unsigned* part, int n, unsigned long long* out);
__global__ void k_gather_ids(const double* hot, const int* ids, int nid, PhysParams P, const int* start,
                             const int* sorted, const float4* cand, int* nl, int* nl_count, int stride, int cap,
                             const int* slot_of, int* maxlist, int* err);

__global__ void k_move_grow(const double* hot_old, double* hot_new, int n, PhysParams P, double dt, const int* order,
                            const int* nl, const int* nl_count, int stride, const double* H, const double* Nrm,
                            const double* wall, const double* gr1, const double* gr2, const int* type,
                            double* fric, int* fric_id, unsigned* divbits, unsigned* block_draws, int* err);
__global__ void k_sum_forces(const double* hot, int n, PhysParams P, const int* order, const int* nl,
                             const int* nl_count, int stride, const double* H, const double* Nrm, const double* wall,
                             double* Fo, double* To, int* err);
__global__ void k_mean_stress(const double* hot, int n, PhysParams P, const int* order, const int* nl,
                              const int* nl_count, int stride, const double* H, const double* Nrm, const double* wall,
                              double* S5, double* F3, unsigned* wave_draws, int* err);

__global__ void k_flag_count(const unsigned* bits, int nw, int* tot);
__global__ void k_flag_write(unsigned* bits, int nw, const int* off, int* ids);

__global__ void k_divide(double* hot, int n0, int D, const int* div_ids, const int32_t* u, PhysParams P, double t,
                         double* mother_pre, int* div_index_of, double* time_p, double* time_q, int* age_p,
                         int* age_q, int* type, int* ancestor, double* fric, int* fric_id, int* order,
                         int* slot_of, int* nl_count);

// Position of cell j as the reference's grid and cell array hold it just
// after division k: a mother of a later division is still undivided, a
// daughter of a later division does not exist yet.
struct Versioned {
    const double* hot;         // final states (all divisions applied)
    const double* mother_pre;  // pre-division mother records, by division
    const int* div_index_of;   // division index of a mother, -1 otherwise
    int n0;                    // cells before the divisions
};

__global__ void k_patch(PhysParams P, Versioned V, int D, const int* div_ids, const int* start, const int* sorted,
                        int* nl, int* nl_count, int stride, int cap, const int* slot_of, int* maxlist, int* err);
__global__ void k_reset_div_index(int* div_index_of, const int* div_ids, int D);

}  // namespace bcg
