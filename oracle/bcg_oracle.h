/*
 * bcg_oracle.h - CPU restatement of the reference mechanics (test
 * infrastructure: the checker for the HIP path, and the timed CPU "port"
 * baseline). Plain C99 (one OpenMP form of the step), FP64, no FMA contraction.
 *
 * Parity status of each function (how it is pinned):
 *   or_min_distance, or_F_cc, or_F_cw, or_sum_forces, or_sim_* (planar):
 *       bit-exact against the reference's own code compiled from source
 *       (oracle/_ref/libbcgref.so) on golden vectors and restart states,
 *       and against closed forms (tests/test_oracle.py).
 *   or_glibc_* : against the host C library's rand() and its published
 *       srand(1) known answers.
 *   planar == 0 (3-D generalisation, no reference counterpart): pinned by
 *       invariants - a planar state stepped in 3-D mode equals planar mode
 *       when nothing leaves the plane, rigid translation in x/y invariance,
 *       brute-force neighbour sets.
 */
#ifndef BCG_ORACLE_H_
#define BCG_ORACLE_H_

#include <stdint.h>
#include "../include/bcg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

void or_min_distance(const bcg_cell* a, const bcg_cell* b, double* out7);
/* reduceNeighbours' test of pairs (a[i], b[i]): 1 when b[i] stays in a[i]'s list */
void or_keep_batch(const bcg_cell* a, const bcg_cell* b, int n, double Rc, int* out);
void or_F_cc(const bcg_params* p, const bcg_cell* a, const bcg_cell* b, double* out7);
/* wall contact at both vertices; consumes one rand() draw like the reference */
void or_F_cw(const bcg_params* p, const bcg_cell* c, double wall_z, double* out18);

or_sim* or_sim_create(const bcg_params* p);
void or_sim_destroy(or_sim* s);
int or_sim_set_cells(or_sim* s, const bcg_cell* cells, int n);
void or_sim_set_fields(or_sim* s, const double* height, const double* normal, const double* wall,
                       const double* gr1, const double* gr2);
int or_sim_num_cells(or_sim* s);
void or_sim_get_cells(or_sim* s, bcg_cell* out);
int or_sim_rebuild(or_sim* s);
int or_sim_rebuild_ids(or_sim* s, const int* ids, int n);
void or_sim_get_neighbours(or_sim* s, int* counts, int* lists, int stride);
void or_sim_get_boxes(or_sim* s, int* counts, int* ids, int stride);
void or_sim_sum_forces(or_sim* s, double* F, double* T);
void or_sim_mean_stress(or_sim* s, double* S5, double* F3);
/* returns divisions performed (>= 0) or a BCG_ERR_* code */
int or_sim_step(or_sim* s, double dt, double t, int rebuild, int* lineage, int lineage_cap);
int64_t or_sim_wall_draws(or_sim* s);
/* The same step with its per-cell work on OpenMP threads (timed CPU
 * baseline); bit-identical to or_sim_step. var_pos must be 0. */
int or_sim_step_par(or_sim* s, double dt, double t, int rebuild, int* lineage, int lineage_cap);
void or_set_threads(int n);
int or_get_threads(void);

/* The oracle's own rand() stream (glibc's algorithm, below): every draw of
 * F_cw and divide comes from it; starts as after srand(1). */
void or_srand(uint32_t seed);
int32_t or_rand(void);

/* glibc TYPE_3 rand() restated from its published description */
typedef struct or_glibc_rng { int32_t r[34]; int32_t i; } or_glibc_rng;
void or_glibc_srand(or_glibc_rng* g, uint32_t seed);
int32_t or_glibc_rand(or_glibc_rng* g);

#ifdef __cplusplus
}
#endif
#endif
