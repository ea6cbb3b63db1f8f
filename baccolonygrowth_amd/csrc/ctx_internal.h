// ctx_internal.h - the mechanics context (struct bcg_ctx) shared by the
// sources of the library: bcg_api.hip owns it, env_api.hip reads its cells
// and bins and writes the fields it reads.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bcg.h"
#include "glibc_rand.h"
#include "mech_kernels.h"

struct BcgStats {
    unsigned long long wall_draws;
    int err;
    int maxlist;
    int ndiv;
    int pad;
};

struct bcg_ctx {
    bcg_params prm;
    bcg::PhysParams P;
    hipStream_t stream = nullptr;
    int N = 0;
    int cur = 0;
    int stride = 0;  // list stride (>= maxCells)
    int cap = 0;
    int nbin = 0;
    bool lists_valid = false;
    // set when a step fails after its divisions began (the rng, the cell
    // buffers and the lists no longer agree); cleared by bcg_set_cells
    bool failed = false;
    double* hot[2] = {nullptr, nullptr};
    double *time_p = nullptr, *time_q = nullptr;
    // DynFric, StaFric of the cell in each slot of the last move (6 doubles a
    // slot) and that cell's ID
    double* fric = nullptr;
    int* fric_id = nullptr;
    int *age_p = nullptr, *age_q = nullptr, *type = nullptr, *ancestor = nullptr;
    int *key = nullptr, *bin_cnt = nullptr, *bin_start = nullptr, *bin_slot = nullptr, *sorted = nullptr;
    int* scattered = nullptr;  // bin order before the in-bin ranking
    int* scan_tot = nullptr;
    float4 *cmL = nullptr, *cand = nullptr;  // FP32 end points by ID, by bin-order slot
    int *nl = nullptr, *nl_count = nullptr, *order = nullptr, *slot_of = nullptr;
    unsigned* divbits = nullptr;  // division flags, a bit per cell ID (zero between steps)
    int *div_ids = nullptr, *div_index_of = nullptr, *ids_tmp = nullptr;
    int32_t* draws = nullptr;   // rand() values of the divisions' draws, 9 a division
    int32_t* hdraws = nullptr;  // pinned host staging of them
    size_t hdraws_cap = 0;
    double* mother_pre = nullptr;
    double *height = nullptr, *normal = nullptr, *wall = nullptr, *gr1 = nullptr, *gr2 = nullptr;
    double *Fbuf = nullptr, *Tbuf = nullptr;
    double* Sbuf = nullptr;  // mean_stress output, 5 per cell
    int* part_max = nullptr;        // per block of the list build: its longest list
    unsigned* part_draws = nullptr; // per block of move/grow: its F_cw draws
    BcgStats* stats = nullptr;  // device
    BcgStats* hstats = nullptr; // pinned host
    int* hids = nullptr;         // pinned host: the step's dividing mothers (lineage)
    int hids_cap = 0;
    bcg::GlibcRand rng;
    hipEvent_t ev[7];
    float last_ms[6] = {0, 0, 0, 0, 0, 0};
    BcgStats last{};
    int last_flagged = 0;
    std::string err;
    // work queued on another stream against the context (the 3-D nutrient
    // update reads the current cells and writes gr1): `after` is recorded
    // there and the context's stream waits for it before its next kernel
    // that reads the growth rates or writes cells
    hipEvent_t after = nullptr;
    bool after_pending = false;
};

// A context whose last step failed part-way accepts only bcg_set_cells.
inline int ctx_check_failed(bcg_ctx* c) {
    if (!c->failed) return 0;
    c->err = "a step failed after its divisions began: upload cells (bcg_set_cells) before further use";
    return BCG_ERR_STATE;
}

inline hipError_t ctx_join(bcg_ctx* c) {
    if (!c->after_pending) return hipSuccess;
    c->after_pending = false;
    return hipStreamWaitEvent(c->stream, c->after, 0);
}


// Bin the context's current cells (counting sort, ascending IDs per bin);
// BCG_ERR_OUT_OF_DOMAIN when a centre lies outside the bin grid.
int bcg_bin_current(bcg_ctx* c);
