// env_api.hip - host side of the planar colony environment (bcg_env_*), the
// time-loop clock (bcg_clock_*) and one pass of the time loop (bcg_sim_step).
//
// Device state of the environment, all in the reference layouts (x-major,
// z fastest, BoxY = 1):
//   colony[2]            Environment, oldEnvironment         [BX][BZ] LocalEnv
//   agar[2][ML]          FieldAgar, oldFieldAgar             [BX][BZA] LocalAga
//   wall[2][ML]          FieldWall, oldFieldWall             [BX] LocalAga
//   aga/res/err[8]       the nested multigrid's arrays       3 species x [nz+1][nx+1]
//   rough*, *shift, ...  density fields of GetDensity / ShiftDensity / BoxAverage
// Height and Normal live in the bound mechanics context (it reads them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bcg.h"
#include "ctx_internal.h"
#include "env_kernels.h"

using namespace bcg;
using namespace bcg::env;

struct bcg_env {
    bcg_env_params prm;
    bcg_ctx* mech = nullptr;
    int BX = 0, BZ = 0, BZA = 0, ML = 0, REF = 0, NHY = 0;
    bcg_env_cell* colony[2] = {nullptr, nullptr};
    bcg_aga* agar[2][MAX_NEST] = {};
    bcg_aga* wall[2][MAX_NEST] = {};
    double* aga[MG_LEVELS] = {};
    double* res[MG_LEVELS] = {};
    double* err[MG_LEVELS] = {};
    double* rt[MG_LEVELS] = {};
    int mgx[MG_LEVELS] = {}, mgz[MG_LEVELS] = {};
    double* scol[2] = {nullptr, nullptr};   // working copies of the nutrient update
    double* swall[2] = {nullptr, nullptr};
    double* srow[2] = {nullptr, nullptr};
    double *insT = nullptr, *dshT = nullptr, *phantom = nullptr;
    size_t lds_device = 0;  // dynamic LDS the nutrient kernel can take on this device
    size_t lds_budget = 0;  // ... and may take (bcg_env_set_lds_limit)
    size_t lds_device_wg = 0;  // the same for the one-workgroup sweeps kernel
    int exec = 0;              // BCG_ENV_EXEC_* (bcg_env_set_execution)
    double *rough = nullptr, *rough1 = nullptr, *rough2 = nullptr, *inside = nullptr;
    double *rwall = nullptr, *rwall1 = nullptr, *rwall2 = nullptr;
    double *rshift = nullptr, *rshift1 = nullptr, *rshift2 = nullptr;
    double *rwshift = nullptr, *rwshift1 = nullptr, *rwshift2 = nullptr;
    double *dens = nullptr, *dens1 = nullptr, *dens2 = nullptr, *wdens = nullptr, *wdens1 = nullptr, *wdens2 = nullptr;
    double *dshift = nullptr, *d1shift = nullptr, *d2shift = nullptr;
    double *wshift = nullptr, *w1shift = nullptr, *w2shift = nullptr;
    double* rheight = nullptr;
    double *filter3d = nullptr, *filter1d = nullptr;
    int* bounds = nullptr;   // device
    int* errflag = nullptr;  // device
    EnvArgs* dargs = nullptr;  // device copy of the update's arguments (phase kernels)
    EnvArgs* hargs = nullptr;  // pinned staging of it
    int hbounds[5] = {0, 0, 0, 0, 0};
    bool have_fields = false;
    std::string err_text;
};

namespace {

int efail(bcg_env* e, int code, const char* fmt, ...) {
    if (e) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        e->err_text = buf;
    }
    return code;
}

#define EHIP(e, x)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return efail(e, BCG_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int ealloc(bcg_env* e, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t r = hipMalloc((void**)p, n * sizeof(T));
    if (r != hipSuccess) return efail(e, BCG_ERR_HIP, "hipMalloc(%zu): %s", n * sizeof(T), hipGetErrorString(r));
    return 0;
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

bcg_env_cell default_env_cell(const bcg_env_params& p) {  // LocalEnv() (ABM/Nutrients.h:13-38)
    bcg_env_cell c;
    memset(&c, 0, sizeof c);
    c.Oxygen = p.maxO2;
    c.pH = 7.0;
    return c;
}

bcg_aga default_aga(const bcg_env_params& p) {  // LocalAga() (ABM/Nutrients.h:66-72)
    bcg_aga g;
    g.CarbonAgar = p.maxCarbon;
    g.WasteAgar = 0.0;
    g.OxygenAgar = p.maxO2;
    g.AcetateAgar = 0.0;
    return g;
}

FieldArgs field_args(bcg_env* e) {
    FieldArgs a;
    bcg_ctx* m = e->mech;
    a.BX = e->BX;
    a.BZ = e->BZ;
    a.REF = e->REF;
    a.NHY = e->NHY;
    a.N = m->N;
    a.FD = e->prm.FilterLen / 2;
    a.BoxLength = e->prm.BoxLength;
    a.cellRadius = e->prm.cellRadius;
    a.density_threshold = e->prm.density_threshold;
    a.hot = m->hot[m->cur];
    a.type = m->type;
    a.start = m->bin_start;
    a.sorted = m->sorted;
    a.rough = e->rough;
    a.rough1 = e->rough1;
    a.rough2 = e->rough2;
    a.inside = e->inside;
    a.rwall = e->rwall;
    a.rwall1 = e->rwall1;
    a.rwall2 = e->rwall2;
    a.rshift = e->rshift;
    a.rshift1 = e->rshift1;
    a.rshift2 = e->rshift2;
    a.rwshift = e->rwshift;
    a.rwshift1 = e->rwshift1;
    a.rwshift2 = e->rwshift2;
    a.height = m->height;
    a.rheight = e->rheight;
    a.normal = m->normal;
    a.bounds = e->bounds;
    a.errflag = e->errflag;
    return a;
}

int read_flags(bcg_env* e, int* flags) {
    int h[6];
    EHIP(e, hipMemcpyAsync(h, e->bounds, 5 * sizeof(int), hipMemcpyDeviceToHost, e->mech->stream));
    EHIP(e, hipMemcpyAsync(h + 5, e->errflag, sizeof(int), hipMemcpyDeviceToHost, e->mech->stream));
    EHIP(e, hipStreamSynchronize(e->mech->stream));
    memcpy(e->hbounds, h, sizeof e->hbounds);
    *flags = h[5];
    return 0;
}

// x extent of multigrid level l as the update's loops see it
// (Multigrid.cpp's BoxX * 2^agarLevels / 2^l / 2, capped at BoxX).
int level_nx(const bcg_env* e, int l) {
    const long v = (long)e->BX * (1L << e->ML) / (1L << l) / 2;
    return (int)std::min((long)e->BX, v);
}

// The update as phase launches on the mechanics stream (the order of
// k_env_nutrients, one launch per barrier there). Levels below `lc` are
// wide (more column walks a colour than one workgroup runs at once): their
// relaxation passes, restriction and interpolation are device-wide
// launches; levels from lc down run in one workgroup per V-cycle.
struct Phased {
    bcg_env* e;
    const EnvArgs& a;
    const EnvArgs* da;  // device copy of a (the grid-wide phase kernels)
    hipStream_t st;
    int lc;
    bool wg_sweeps;
    size_t wg_lds;
    size_t coarse_lds;  // dynamic LDS of the tail's coarse solve (0: the one-wave or global path)
    int coarse_mode;    // 2: a wave per species, 1: the workgroup, 0: global
    dim3 pts(long n) const {
        return dim3((unsigned)std::max(1L, std::min((long)PH_MAX_BLOCKS, (n + PH_THREADS - 1) / PH_THREADS)));
    }
    long level_points(int l) const { return (long)(level_nx(e, l) + 1) * (a.mgz[l] + 1); }
    void outer(bool last) const {
        const int AL = a.ML, BX = a.BX;
        const dim3 T(PH_THREADS);
        long all = 0;
        for (int l = 0; l < MG_LEVELS; l++) all = std::max(all, (long)(a.mgx[l] + 1) * (a.mgz[l] + 1));
        hipLaunchKernelGGL(k_nut_copy_to_mg, pts(std::max(BX, a.BZA + 1)), T, 0, st, da, 0);
        hipLaunchKernelGGL(k_nut_reset_zero, pts(all), T, 0, st, da);
        hipLaunchKernelGGL(k_nut_update_sol, pts(all), T, 0, st, da);
        hipLaunchKernelGGL(k_nut_update_res, pts(all), T, 0, st, da);
        hipLaunchKernelGGL(k_nut_rterm, pts(3 * level_points(0)), T, 0, st, da, 0);
        auto relax3 = [&](int l) {
            const int nx = level_nx(e, l);
            const long walks = 3L * (nx / 2 + 1);
            const dim3 g((unsigned)((walks + PH_RELAX_THREADS - 1) / PH_RELAX_THREADS));
            for (int j = 0; j < 3; j++)
                for (int par = 0; par < 2; par++)
                    hipLaunchKernelGGL(k_nut_relax, g, dim3(PH_RELAX_THREADS), 0, st, da, l, par);
        };
        for (int cyc = 0; cyc < 3; cyc++) {
            for (int l = 0; l < lc; l++) {
                relax3(l);
                hipLaunchKernelGGL(k_nut_restrict, pts(level_points(l + 1)), T, 0, st, da, l + 1);
                hipLaunchKernelGGL(k_nut_rterm, pts(3 * level_points(l + 1)), T, 0, st, da, l + 1);
            }
            hipLaunchKernelGGL(k_nut_vcycle_tail, dim3(1), dim3(ENV_THREADS), coarse_lds, st, a, lc, coarse_mode);
            for (int l = lc - 1; l >= 0; l--) {
                hipLaunchKernelGGL(k_nut_interp, pts(level_points(l)), T, 0, st, da, l);
                relax3(l);
            }
        }
        hipLaunchKernelGGL(k_nut_update_sol, pts(all), T, 0, st, da);
        for (int l = 0; l < AL - 1; l++)
            hipLaunchKernelGGL(k_nut_rewrite, pts(std::max(level_nx(e, l) + 1, 2 * a.BZA)), T, 0, st, da, l);
        hipLaunchKernelGGL(k_nut_update_res, pts(all), T, 0, st, da);
        hipLaunchKernelGGL(k_nut_reset_zero, pts(all), T, 0, st, da);
        for (int l = 1; l <= AL; l++) hipLaunchKernelGGL(k_nut_nest, pts(level_points(l)), T, 0, st, da, l);
        hipLaunchKernelGGL(k_nut_copy_from_mg, pts((long)AL * 2 * BX), T, 0, st, da);
        if (wg_sweeps) {
            hipLaunchKernelGGL(k_nut_sweeps_wg, dim3(1), dim3(ENV_THREADS), wg_lds, st, a, last ? 1 : 0);
        } else {
            const long items = (long)(e->hbounds[1] - e->hbounds[0] + 1) * (e->hbounds[4] + 1) + (long)AL * BX;
            for (int j = 0; j < 200; j++)
                hipLaunchKernelGGL(k_nut_sweep, pts(items), T, 0, st, da, (last && j >= 198) ? 1 : 0,
                                   (j & 1) ? -1 : 1, j & 1);
        }
    }
};

int nutrients_phased(bcg_env* e, const EnvArgs& a) {
    hipStream_t st = e->mech->stream;
    // first narrow level: it and every coarser one fit one workgroup a colour
    // and are small enough for the workgroup's transfers
    int lc = MG_LEVELS - 1;
    while (lc > 0 && 3L * (level_nx(e, lc - 1) / 2 + 1) <= ENV_THREADS &&
           (long)(level_nx(e, lc - 1) + 1) * (e->mgz[lc - 1] + 1) <= TAIL_MAX_POINTS)
        lc--;
    // sweeps in one workgroup when one round of it covers a sweep's boxes and
    // rows (or when asked); device-wide launches otherwise
    const long items = (long)(e->hbounds[1] - e->hbounds[0] + 1) * (e->hbounds[4] + 1) + (long)e->ML * e->BX;
    bool wg = items <= ENV_THREADS;
    if (e->exec & BCG_ENV_EXEC_WG_SWEEPS) wg = true;
    if (e->exec & BCG_ENV_EXEC_GRID_SWEEPS) wg = false;
    EnvArgs aa = a;
    size_t lds = 0;
    if (wg) {
        const size_t need = (size_t)2 * 4 * aa.lds_nf * sizeof(double);
        aa.lds_colony = need <= std::min(e->lds_budget, e->lds_device_wg);
        lds = aa.lds_colony ? need : 0;
    } else {
        aa.lds_colony = 0;
    }
    // the coarsest level in LDS for the tail's coarse solve: one wave per
    // species when its columns fit a wave, else the whole workgroup; the
    // one-wave global path when neither fits
    size_t coarse_lds = 0;
    int coarse_mode = 0;
    {
        const int l = MG_LEVELS - 1;
        const long npts = (long)(e->mgx[l] + 1) * (e->mgz[l] + 1);
        const int nx = level_nx(e, l);
        const size_t need = (size_t)3 * 3 * npts * sizeof(double);
        if (need <= (size_t)COARSE_LDS_BYTES) {
            coarse_lds = need;
            coarse_mode = nx + 1 <= 64 ? 2 : 1;
        }
    }
    *e->hargs = aa;
    EHIP(e, hipMemcpyAsync(e->dargs, e->hargs, sizeof(EnvArgs), hipMemcpyHostToDevice, st));
    const Phased P{e, aa, e->dargs, st, lc, wg, lds, coarse_lds, coarse_mode};
    hipLaunchKernelGGL(k_nut_begin, P.pts(std::max((long)e->BX * e->BZ, (long)e->ML * e->BX)), dim3(PH_THREADS), 0, st,
                       P.da);
    const int n = aa.n_outer;
    if (n >= 2 && !(e->exec & BCG_ENV_EXEC_NO_GRAPH)) {
        // outer iterations 0 .. n-2 are the same launches: record one, replay
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        EHIP(e, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        P.outer(false);
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (ec != hipSuccess) return efail(e, BCG_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) return efail(e, BCG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ei));
        hipError_t el = hipSuccess;
        for (int o = 0; o < n - 1 && el == hipSuccess; o++) el = hipGraphLaunch(x, st);
        // the executable graph may go once its launches are queued
        const hipError_t es = hipStreamSynchronize(st);
        hipGraphExecDestroy(x);
        if (el != hipSuccess) return efail(e, BCG_ERR_HIP, "graph launch: %s", hipGetErrorString(el));
        EHIP(e, es);
    } else {
        for (int o = 0; o < n - 1; o++) P.outer(false);
    }
    if (n >= 1) P.outer(true);
    hipLaunchKernelGGL(k_nut_end, P.pts(std::max((long)e->BX * std::max(e->BZ, e->BZA), (long)e->ML * e->BX)),
                       dim3(PH_THREADS), 0, st, P.da);
    EHIP(e, hipGetLastError());
    // the staging buffer is reused by the next update: its copy must have run
    EHIP(e, hipStreamSynchronize(st));
    return 0;
}

}  // namespace

extern "C" {

int bcg_env_set_execution(bcg_env* e, int32_t flags) {
    if (!e || (flags & ~(BCG_ENV_EXEC_ONE_WORKGROUP | BCG_ENV_EXEC_NO_GRAPH | BCG_ENV_EXEC_WG_SWEEPS |
                         BCG_ENV_EXEC_GRID_SWEEPS)) ||
        ((flags & BCG_ENV_EXEC_WG_SWEEPS) && (flags & BCG_ENV_EXEC_GRID_SWEEPS)))
        return BCG_ERR_ARG;
    e->exec = flags;
    return BCG_OK;
}

int bcg_env_default_params(bcg_env_params* p) {
    if (!p) return BCG_ERR_ARG;
    memset(p, 0, sizeof *p);
    // demo input (ABM/run_sim/inputFile_20mM_SmallSimulation.txt) over the
    // defaults of ABM/Constants.cpp
    p->BoxX = 256;
    p->BoxY = 1;
    p->BoxZ = 50;
    p->BoxZAgar = 32;
    p->maxLevelsMG = 2;
    p->refinementGridHeight = 1;
    p->FilterLen = 5;
    p->crossFeeding = 0;
    p->NutrientGSI = 0;
    p->InterfaceCondition = 1;
    p->cellRadius = 0.5;
    p->BoxLength = 3.0 + 2 * 0.5;  // L_divide + 2 Radius (ABM/CellMD3D.cpp:52)
    p->density_threshold = 0.55;
    p->UpdateTime = 0.01;
    p->Cdt = 0.1;
    p->KC_aer = 0.02;
    p->KC_ana = 0.02;
    p->KW = 0.02;
    p->KO = 0.0001;
    p->KA = 5;
    p->KI = 4.89;
    p->C_rate = 0.0017;
    p->W_rate = 0.0017;
    p->Carbon_to_Waste = 1.0;
    p->Maintenance_rate = 0.0;
    p->maxGrowthRate1 = 1.0;
    p->maxGrowthRate2 = 1.0;
    p->DiffColony = 0.15;
    p->DiffAgar = 1.0;
    p->maxCarbon = 20.0;
    p->maxO2 = 0.26;
    p->maxO2_Agar = 0.26;
    p->qCarAer = 0.0041;
    p->qCarAna = 0.0105;
    p->qAceAer = 0.0124;
    p->qO2Car = 0.0083;
    p->qO2Ace = 0.0083;
    p->pAceAer = 0.0011;
    p->pAceAna = 0.0060;
    p->DiffColonyCarbon = 0.15;
    p->DiffAgarCarbon = 1.0;
    p->DiffColonyO2 = 0.506;
    p->DiffAgarO2 = 3.378;
    p->DiffColonyAce = 0.223;
    p->DiffAgarAce = 1.486;
    p->maxGrowthRateCarAer = 0.9;
    p->maxGrowthRateCarAna = 0.6;
    p->maxGrowthRateAceAer = 0.4;
    p->Kw = 1e-8;
    p->Ka = 0.01778;
    p->Kb = 0.00015849;
    p->Cb = 100;
    p->c1 = 55000;
    p->qCarAernt = 0.00037;
    p->qCarAnant = 0.00375;
    p->qAceAernt = 0.00113;
    p->qO2Carnt = 0.00075;
    p->qO2Acent = 0.00075;
    p->pAceAnant = 0.00375;
    return BCG_OK;
}

int bcg_env_destroy(bcg_env* e) {
    if (!e) return BCG_OK;
    if (e->mech && e->mech->stream) {
        ctx_join(e->mech);
        hipStreamSynchronize(e->mech->stream);
    }
    std::vector<void*> ptrs = {e->colony[0], e->colony[1], e->rough,   e->rough1,   e->rough2,   e->inside,
                               e->rwall,     e->rwall1,    e->rwall2,  e->rshift,   e->rshift1,  e->rshift2,
                               e->rwshift,   e->rwshift1,  e->rwshift2, e->dens,    e->dens1,    e->dens2,
                               e->wdens,     e->wdens1,    e->wdens2,  e->dshift,   e->d1shift,  e->d2shift,
                               e->wshift,    e->w1shift,   e->w2shift, e->rheight,  e->filter3d, e->filter1d,
                               e->bounds,    e->errflag};
    for (int b = 0; b < 2; b++)
        for (int l = 0; l < MAX_NEST; l++) {
            ptrs.push_back(e->agar[b][l]);
            ptrs.push_back(e->wall[b][l]);
        }
    for (int l = 0; l < MG_LEVELS; l++) {
        ptrs.push_back(e->aga[l]);
        ptrs.push_back(e->res[l]);
        ptrs.push_back(e->err[l]);
        ptrs.push_back(e->rt[l]);
    }
    for (int b = 0; b < 2; b++) {
        ptrs.push_back(e->scol[b]);
        ptrs.push_back(e->swall[b]);
        ptrs.push_back(e->srow[b]);
    }
    ptrs.push_back(e->insT);
    ptrs.push_back(e->dshT);
    ptrs.push_back(e->phantom);
    ptrs.push_back(e->dargs);
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (e->hargs) hipHostFree(e->hargs);
    delete e;
    return BCG_OK;
}

int bcg_env_create(const bcg_env_params* p, bcg_ctx* mech, bcg_env** out) {
    if (!p || !mech || !out) return BCG_ERR_ARG;
    *out = nullptr;
    const bcg_params& mp = mech->prm;
    if (p->BoxY != 1 || p->BoxX < 8 || p->BoxZ < 4 || p->BoxZAgar < 2 || p->maxLevelsMG < 1 ||
        p->maxLevelsMG > MAX_NEST || p->refinementGridHeight < 1 || p->refinementGridHeight > MAX_REF ||
        p->FilterLen < 1 || p->FilterLen % 2 == 0 || !(p->BoxLength > 0) || !(p->density_threshold > 0))
        return BCG_ERR_ARG;
    if (mp.BoxX != p->BoxX || mp.BoxY != p->BoxY || mp.BoxZ != p->BoxZ ||
        mp.refinementGridHeight != p->refinementGridHeight || mp.BoxLength != p->BoxLength || !mp.planar)
        return BCG_ERR_ARG;  // the environment shares the mechanics' bins and height grid
    bcg_env* e = new bcg_env();
    e->prm = *p;
    e->mech = mech;
    e->BX = p->BoxX;
    e->BZ = p->BoxZ;
    e->BZA = p->BoxZAgar;
    e->ML = p->maxLevelsMG;
    e->REF = p->refinementGridHeight;
    e->NHY = p->BoxY * p->refinementGridHeight;
    const size_t nb = (size_t)e->BX * e->BZ;
    int rc = 0;
    auto A = [&](auto** ptr, size_t n) {
        if (!rc) rc = ealloc(e, ptr, n);
    };
    A(&e->colony[0], nb);
    A(&e->colony[1], nb);
    for (int b = 0; b < 2; b++)
        for (int l = 0; l < e->ML; l++) {
            A(&e->agar[b][l], (size_t)e->BX * e->BZA);
            A(&e->wall[b][l], (size_t)e->BX);
        }
    // Multigrid(maxLevelsMG, 8, BoxX, BoxY, BoxZAgar) (ABM/Multigrid.cpp:5-33)
    int nx = e->BX, nz = e->BZA;
    for (int l = 0; l < MG_LEVELS; l++) {
        if (l >= e->ML) {
            nx /= 2;
            nz /= 2;
        }
        // nz reaches 0 on the coarsest level when BoxZAgar < 2^(8 - maxLevelsMG)
        // (the demo input: 32 < 64); the reference allocates that level with
        // one z point and its stencils then read the next column.
        if (nx < 1) rc = rc ? rc : efail(e, BCG_ERR_ARG, "grid too small for 8 multigrid levels");
        e->mgx[l] = nx;
        e->mgz[l] = nz;
        const size_t n = (size_t)(nx + 1) * (nz + 1);
        A(&e->aga[l], 3 * n);
        A(&e->res[l], 3 * n);
        A(&e->err[l], 3 * n);
        A(&e->rt[l], 3 * n);
    }
    for (int b = 0; b < 2; b++) {
        A(&e->scol[b], 4 * nb);
        A(&e->swall[b], (size_t)4 * e->ML * e->BX);
        A(&e->srow[b], (size_t)4 * e->ML * e->BX);
    }
    A(&e->insT, nb);
    A(&e->dshT, nb);
    A(&e->phantom, (size_t)3 * PHANTOM_THREADS);
    for (double** q : {&e->rough, &e->rough1, &e->rough2, &e->inside, &e->rshift, &e->rshift1, &e->rshift2, &e->dens,
                       &e->dens1, &e->dens2, &e->dshift, &e->d1shift, &e->d2shift})
        A(q, nb);
    for (double** q : {&e->rwall, &e->rwall1, &e->rwall2, &e->rwshift, &e->rwshift1, &e->rwshift2, &e->wdens,
                       &e->wdens1, &e->wdens2, &e->wshift, &e->w1shift, &e->w2shift})
        A(q, (size_t)e->BX);
    A(&e->rheight, (size_t)e->BX * e->REF * e->NHY);
    A(&e->filter3d, (size_t)p->FilterLen * p->FilterLen);
    A(&e->filter1d, (size_t)p->FilterLen);
    A(&e->bounds, 5);
    A(&e->errflag, 1);
    A(&e->dargs, 1);
    if (!rc && hipHostMalloc((void**)&e->hargs, sizeof(EnvArgs), hipHostMallocDefault) != hipSuccess)
        rc = efail(e, BCG_ERR_HIP, "hipHostMalloc failed");
    if (rc) {
        mech->err = "bcg_env_create: " + e->err_text;  // readable through bcg_last_error
        bcg_env_destroy(e);
        return rc;
    }
    hipStream_t st = mech->stream;
    {
        // LDS per workgroup on this device, less the kernel's static LDS
        int dev = 0, v = 0;
        hipFuncAttributes fa;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_env_nutrients)) == hipSuccess &&
            (size_t)v > fa.sharedSizeBytes)
            e->lds_budget = std::min((size_t)ENV_LDS_BYTES, (size_t)v - fa.sharedSizeBytes);
        if (e->lds_budget > 64 * 1024 &&
            hipFuncSetAttribute(reinterpret_cast<const void*>(&k_env_nutrients),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds_budget) != hipSuccess) {
            (void)hipGetLastError();
            e->lds_budget = 64 * 1024 - fa.sharedSizeBytes;
        }
        e->lds_device = e->lds_budget;
        // the one-workgroup sweeps kernel of the phased update
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_nut_sweeps_wg)) == hipSuccess &&
            (size_t)v > fa.sharedSizeBytes) {
            e->lds_device_wg = std::min((size_t)ENV_LDS_BYTES, (size_t)v - fa.sharedSizeBytes);
            if (e->lds_device_wg > 64 * 1024 &&
                hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nut_sweeps_wg),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds_device_wg) != hipSuccess) {
                (void)hipGetLastError();
                e->lds_device_wg = 64 * 1024 - fa.sharedSizeBytes;
            }
        }
    }
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (!rc && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = efail(e, BCG_ERR_HIP, "upload failed");
    };
    // initial contents: constructors of the reference's arrays
    {
        std::vector<bcg_env_cell> col(nb, default_env_cell(*p));
        up(e->colony[0], col.data(), nb * sizeof(bcg_env_cell));
        up(e->colony[1], col.data(), nb * sizeof(bcg_env_cell));
        std::vector<bcg_aga> ag((size_t)e->BX * e->BZA, default_aga(*p));
        for (int b = 0; b < 2; b++)
            for (int l = 0; l < e->ML; l++) {
                up(e->agar[b][l], ag.data(), ag.size() * sizeof(bcg_aga));
                up(e->wall[b][l], ag.data(), (size_t)e->BX * sizeof(bcg_aga));
            }
        for (int l = 0; l < MG_LEVELS; l++) {
            // LocalAga() in every point, as three species blocks
            const size_t n = (size_t)(e->mgx[l] + 1) * (e->mgz[l] + 1);
            const bcg_aga d = default_aga(*p);
            std::vector<double> m(3 * n);
            for (size_t q = 0; q < n; q++) {
                m[q] = d.CarbonAgar;
                m[n + q] = d.OxygenAgar;
                m[2 * n + q] = d.AcetateAgar;
            }
            up(e->aga[l], m.data(), m.size() * sizeof(double));
            up(e->res[l], m.data(), m.size() * sizeof(double));
            up(e->err[l], m.data(), m.size() * sizeof(double));
            up(e->rt[l], m.data(), m.size() * sizeof(double));
            // the uploads read m before it goes out of scope
            if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = efail(e, BCG_ERR_HIP, "init sync failed");
        }
        std::vector<double> zero(nb, 0.0);
        for (double* q : {e->rough, e->rough1, e->rough2, e->inside, e->rshift, e->rshift1, e->rshift2, e->dens, e->dens1,
                          e->dens2, e->dshift, e->d1shift, e->d2shift})
            up(q, zero.data(), nb * sizeof(double));
        for (double* q : {e->rwall, e->rwall1, e->rwall2, e->rwshift, e->rwshift1, e->rwshift2, e->wdens, e->wdens1,
                          e->wdens2, e->wshift, e->w1shift, e->w2shift})
            up(q, zero.data(), (size_t)e->BX * sizeof(double));
        // RoughHeight.Initialize(cellRadius) (ABM/Simulation.cpp:129)
        std::vector<double> rh((size_t)e->BX * e->REF * e->NHY, p->cellRadius);
        up(e->rheight, rh.data(), rh.size() * sizeof(double));
        // filters (ABM/Simulation.cpp:60-110): 1-D box, x-z Gaussian exp(-di^2-dk^2)/norm
        const int FL = p->FilterLen;
        std::vector<double> f1(FL, 1.0 / (double)(FL));
        double norm = 0;
        for (int ii = 0; ii < FL; ii++)
            for (int kk = 0; kk < FL; kk++) norm = norm + exp(-(ii - FL / 2) * (ii - FL / 2) - (kk - FL / 2) * (kk - FL / 2));
        std::vector<double> f3((size_t)FL * FL);
        for (int ii = 0; ii < FL; ii++)
            for (int kk = 0; kk < FL; kk++)
                f3[(size_t)ii * FL + kk] = exp(-(ii - FL / 2) * (ii - FL / 2) - (kk - FL / 2) * (kk - FL / 2)) / norm;
        up(e->filter1d, f1.data(), FL * sizeof(double));
        up(e->filter3d, f3.data(), f3.size() * sizeof(double));
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = efail(e, BCG_ERR_HIP, "init sync failed");
    }
    if (rc) {
        mech->err = "bcg_env_create: " + e->err_text;
        bcg_env_destroy(e);
        return rc;
    }
    *out = e;
    return BCG_OK;
}

int bcg_env_set_lds_limit(bcg_env* e, int64_t bytes) {
    if (!e || bytes < 0) return BCG_ERR_ARG;
    e->lds_budget = std::min(e->lds_device, (size_t)bytes);
    return BCG_OK;
}

int bcg_env_last_error(const bcg_env* e, char* buf, int len) {
    if (!e || !buf || len <= 0) return BCG_ERR_ARG;
    snprintf(buf, len, "%s", e->err_text.c_str());
    return BCG_OK;
}

int bcg_env_set_state(bcg_env* e, const bcg_env_cell* colony, const bcg_env_cell* colony_old, const bcg_aga* agar,
                      const bcg_aga* agar_old, const bcg_aga* wall, const bcg_aga* wall_old) {
    if (!e) return BCG_ERR_ARG;
    hipStream_t st = e->mech->stream;
    const size_t nb = (size_t)e->BX * e->BZ, na = (size_t)e->BX * e->BZA, nw = (size_t)e->BX;
    if (colony) EHIP(e, hipMemcpyAsync(e->colony[0], colony, nb * sizeof(bcg_env_cell), hipMemcpyHostToDevice, st));
    if (colony_old) EHIP(e, hipMemcpyAsync(e->colony[1], colony_old, nb * sizeof(bcg_env_cell), hipMemcpyHostToDevice, st));
    for (int l = 0; l < e->ML; l++) {
        if (agar) EHIP(e, hipMemcpyAsync(e->agar[0][l], agar + l * na, na * sizeof(bcg_aga), hipMemcpyHostToDevice, st));
        if (agar_old) EHIP(e, hipMemcpyAsync(e->agar[1][l], agar_old + l * na, na * sizeof(bcg_aga), hipMemcpyHostToDevice, st));
        if (wall) EHIP(e, hipMemcpyAsync(e->wall[0][l], wall + l * nw, nw * sizeof(bcg_aga), hipMemcpyHostToDevice, st));
        if (wall_old) EHIP(e, hipMemcpyAsync(e->wall[1][l], wall_old + l * nw, nw * sizeof(bcg_aga), hipMemcpyHostToDevice, st));
    }
    EHIP(e, hipStreamSynchronize(st));
    return BCG_OK;
}

int bcg_env_get_state(bcg_env* e, bcg_env_cell* colony, bcg_env_cell* colony_old, bcg_aga* agar, bcg_aga* agar_old,
                      bcg_aga* wall, bcg_aga* wall_old) {
    if (!e) return BCG_ERR_ARG;
    hipStream_t st = e->mech->stream;
    const size_t nb = (size_t)e->BX * e->BZ, na = (size_t)e->BX * e->BZA, nw = (size_t)e->BX;
    if (colony) EHIP(e, hipMemcpyAsync(colony, e->colony[0], nb * sizeof(bcg_env_cell), hipMemcpyDeviceToHost, st));
    if (colony_old) EHIP(e, hipMemcpyAsync(colony_old, e->colony[1], nb * sizeof(bcg_env_cell), hipMemcpyDeviceToHost, st));
    for (int l = 0; l < e->ML; l++) {
        if (agar) EHIP(e, hipMemcpyAsync(agar + l * na, e->agar[0][l], na * sizeof(bcg_aga), hipMemcpyDeviceToHost, st));
        if (agar_old) EHIP(e, hipMemcpyAsync(agar_old + l * na, e->agar[1][l], na * sizeof(bcg_aga), hipMemcpyDeviceToHost, st));
        if (wall) EHIP(e, hipMemcpyAsync(wall + l * nw, e->wall[0][l], nw * sizeof(bcg_aga), hipMemcpyDeviceToHost, st));
        if (wall_old) EHIP(e, hipMemcpyAsync(wall_old + l * nw, e->wall[1][l], nw * sizeof(bcg_aga), hipMemcpyDeviceToHost, st));
    }
    EHIP(e, hipStreamSynchronize(st));
    return BCG_OK;
}

int bcg_env_fields(bcg_env* e, int32_t pass) {
    if (!e || pass < BCG_ENV_INIT || pass > BCG_ENV_HEIGHT) return BCG_ERR_ARG;
    bcg_ctx* m = e->mech;
    hipStream_t st = m->stream;
    int rc = bcg_bin_current(m);
    if (rc) return efail(e, rc, "binning the cells failed (code %d)", rc);
    EHIP(e, hipMemsetAsync(e->errflag, 0, sizeof(int), st));
    const FieldArgs a = field_args(e);
    const int T = 128;
    const int nbox = e->BX * e->BZ;
    hipLaunchKernelGGL(k_env_init_bounds, dim3(1), dim3(1), 0, st, a);
    if (pass == BCG_ENV_HEIGHT) {
        hipLaunchKernelGGL(k_env_height, dim3(nblk(e->BX, T)), dim3(T), 0, st, a);
        hipLaunchKernelGGL(k_env_pad_bounds, dim3(1), dim3(1), 0, st, a);
    } else {
        hipLaunchKernelGGL(k_env_density, dim3(nblk(e->BX, T)), dim3(T), 0, st, a);
        hipLaunchKernelGGL(k_env_halo, dim3(nblk(nbox, T)), dim3(T), 0, st, a);
        hipLaunchKernelGGL(k_env_pad_bounds, dim3(1), dim3(1), 0, st, a);
        hipLaunchKernelGGL(k_env_shift, dim3(nblk(nbox, T)), dim3(T), 0, st, a);
        // six BoxAverage calls (ABM/Simulation.cpp:190-195, 275-280); the
        // unshifted ones read the never-written (zero) rough wall rows
        const int nav = e->BX * (e->BZ + 1);
        const double* in[6] = {e->rough, e->rough1, e->rough2, e->rshift, e->rshift1, e->rshift2};
        const double* inw[6] = {e->rwall, e->rwall1, e->rwall2, e->rwshift, e->rwshift1, e->rwshift2};
        double* outd[6] = {e->dens, e->dens1, e->dens2, e->dshift, e->d1shift, e->d2shift};
        double* outw[6] = {e->wdens, e->wdens1, e->wdens2, e->wshift, e->w1shift, e->w2shift};
        for (int s = 0; s < 6; s++)
            hipLaunchKernelGGL(k_env_box_average, dim3(nblk(nav, T)), dim3(T), 0, st, a, in[s], inw[s], outd[s], outw[s],
                               e->filter3d);
    }
    const int nwin = (e->BX + 12) * e->REF;  // the height window never exceeds (BoxX + 11) ref
    if (pass != BCG_ENV_INIT) {
        hipLaunchKernelGGL(k_env_height_average, dim3(nblk(nwin, T)), dim3(T), 0, st, a);
        hipLaunchKernelGGL(k_env_smooth, dim3(nblk(nwin, T)), dim3(T), 0, st, a, e->filter1d);
    }
    hipLaunchKernelGGL(k_env_normal, dim3(nblk(nwin, T)), dim3(T), 0, st, a);
    EHIP(e, hipGetLastError());
    int flags = 0;
    if ((rc = read_flags(e, &flags))) return rc;
    if (pass != BCG_ENV_HEIGHT) e->have_fields = true;
    if (flags & ERR_RANGE)
        return efail(e, BCG_ERR_OUT_OF_DOMAIN, "a filter or stencil window [%d, %d] x [0, %d] leaves the %d x %d grid",
                     e->hbounds[0], e->hbounds[1], e->hbounds[4], e->BX, e->BZ);
    return BCG_OK;
}

int bcg_env_nutrients(bcg_env* e, int32_t* count) {
    if (!e) return BCG_ERR_ARG;
    const bcg_env_params& p = e->prm;
    if (p.crossFeeding || p.NutrientGSI || p.InterfaceCondition != 1)
        return efail(e, BCG_ERR_UNSUPPORTED,
                     "only CrossFeeding 0, NutrientGSI 0, InterfaceCondition 1 are defined in the reference's "
                     "multigrid nutrient path");
    if (!e->have_fields) return efail(e, BCG_ERR_STATE, "no density fields yet (call bcg_env_fields)");
    const int minX = e->hbounds[0], maxX = e->hbounds[1], maxH = e->hbounds[4];
    if (minX < 1 || maxX > e->BX - 2 || maxH > e->BZ - 2)
        return efail(e, BCG_ERR_OUT_OF_DOMAIN, "colony window [%d, %d] x [0, %d] too close to the %d x %d grid edge",
                     minX, maxX, maxH, e->BX, e->BZ);
    bcg_ctx* m = e->mech;
    hipStream_t st = m->stream;
    if (int rc_ = ctx_check_failed(m)) return efail(e, rc_, "%s", m->err.c_str());
    EHIP(e, ctx_join(m));  // writes the context's growth rates
    EnvArgs a;
    memset(&a, 0, sizeof a);
    a.BX = e->BX;
    a.BZ = e->BZ;
    a.BZA = e->BZA;
    a.ML = e->ML;
    a.REF = e->REF;
    a.NHY = e->NHY;
    a.BoxLength = p.BoxLength;
    a.cellRadius = p.cellRadius;
    a.density_threshold = p.density_threshold;
    a.Cdt = p.Cdt;
    a.KC_aer = p.KC_aer;
    a.KC_ana = p.KC_ana;
    a.KW = p.KW;
    a.KO = p.KO;
    a.KA = p.KA;
    a.C_rate = p.C_rate;
    a.W_rate = p.W_rate;
    a.Carbon_to_Waste = p.Carbon_to_Waste;
    a.Maintenance_rate = p.Maintenance_rate;
    a.maxGrowthRate2 = p.maxGrowthRate2;
    a.DiffColony = p.DiffColony;
    a.DiffAgar = p.DiffAgar;
    a.maxCarbon = p.maxCarbon;
    a.maxO2 = p.maxO2;
    a.maxO2_Agar = p.maxO2_Agar;
    a.qCarAer = p.qCarAer;
    a.qCarAna = p.qCarAna;
    a.qAceAer = p.qAceAer;
    a.qO2Car = p.qO2Car;
    a.qO2Ace = p.qO2Ace;
    a.pAceAer = p.pAceAer;
    a.pAceAna = p.pAceAna;
    a.DiffColonyCarbon = p.DiffColonyCarbon;
    a.DiffAgarCarbon = p.DiffAgarCarbon;
    a.DiffColonyO2 = p.DiffColonyO2;
    a.DiffAgarO2 = p.DiffAgarO2;
    a.DiffColonyAce = p.DiffColonyAce;
    a.DiffAgarAce = p.DiffAgarAce;
    a.maxGrowthRateCarAer = p.maxGrowthRateCarAer;
    a.maxGrowthRateCarAna = p.maxGrowthRateCarAna;
    a.maxGrowthRateAceAer = p.maxGrowthRateAceAer;
    a.Kw = p.Kw;
    a.Ka = p.Ka;
    a.Kb = p.Kb;
    a.Cb = p.Cb;
    a.c1 = p.c1;
    a.qCarAernt = p.qCarAernt;
    a.qCarAnant = p.qCarAnant;
    a.qAceAernt = p.qAceAernt;
    a.qO2Carnt = p.qO2Carnt;
    a.qO2Acent = p.qO2Acent;
    a.pAceAnant = p.pAceAnant;
    a.cc = chem_const(a);
    a.crossFeeding = p.crossFeeding;
    // n = int(3600 UpdateTime 740 / dt), dt = 100 (ABM/Multigrid.cpp:993-994)
    const double dt = 100;
    a.n_outer = (int)(3600 * p.UpdateTime * 740 / dt);
    for (int b = 0; b < 2; b++) {
        a.colony[b] = e->colony[b];
        for (int l = 0; l < MAX_NEST; l++) {
            a.agar[b][l] = e->agar[b][l];
            a.wall[b][l] = e->wall[b][l];
        }
    }
    for (int l = 0; l < MG_LEVELS; l++) {
        a.aga[l] = e->aga[l];
        a.res[l] = e->res[l];
        a.err[l] = e->err[l];
        a.rt[l] = e->rt[l];
        a.mgx[l] = e->mgx[l];
        a.mgz[l] = e->mgz[l];
    }
    for (int b = 0; b < 2; b++) {
        a.scol[b] = e->scol[b];
        a.swall[b] = e->swall[b];
        a.srow[b] = e->srow[b];
    }
    a.insT = e->insT;
    a.dshT = e->dshT;
    a.phantom = e->phantom;
    a.inside = e->inside;
    a.dshift = e->dshift;
    a.d1shift = e->d1shift;
    a.d2shift = e->d2shift;
    a.wshift = e->wshift;
    a.bounds = e->bounds;
    a.errflag = e->errflag;
    // the colony window [minX-1, maxX+1] x [0, maxH+1] of both buffers in
    // LDS when it fits
    const int lw = maxX - minX + 3, lh = maxH + 2;
    const size_t lds = (size_t)2 * 4 * lw * lh * sizeof(double);
    a.lds_colony = lds <= e->lds_budget;
    a.lds_w = lw;
    a.lds_nf = lw * lh;
    EHIP(e, hipMemsetAsync(e->errflag, 0, sizeof(int), st));
    if (e->exec & BCG_ENV_EXEC_ONE_WORKGROUP) {
        hipLaunchKernelGGL(k_env_nutrients, dim3(1), dim3(ENV_THREADS), a.lds_colony ? lds : 0, st, a);
    } else {
        int rc = nutrients_phased(e, a);
        if (rc) return rc;
    }
    // the mechanics grows with the caller's Environment (ABM/grow.cpp:26-36)
    const int nb = e->BX * e->BZ;
    hipLaunchKernelGGL(k_env_growth_to_mech, dim3(nblk(nb, 256)), dim3(256), 0, st, e->colony[0], nb, m->gr1, m->gr2);
    EHIP(e, hipGetLastError());
    int flags = 0;
    int rc = read_flags(e, &flags);
    if (rc) return rc;
    if (flags & ERR_COLONY_EDGE) return efail(e, BCG_ERR_OUT_OF_DOMAIN, "a colony box with x <= 2 (Need more boxes)");
    if (count) *count = a.n_outer * 200;
    return BCG_OK;
}

int bcg_env_get_fields(bcg_env* e, double* density, double* density_shift, double* inside, double* wall_shift,
                       int32_t* bounds) {
    if (!e) return BCG_ERR_ARG;
    hipStream_t st = e->mech->stream;
    const size_t nb = (size_t)e->BX * e->BZ;
    if (density) EHIP(e, hipMemcpyAsync(density, e->dens, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    if (density_shift) EHIP(e, hipMemcpyAsync(density_shift, e->dshift, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    if (inside) EHIP(e, hipMemcpyAsync(inside, e->inside, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    if (wall_shift) EHIP(e, hipMemcpyAsync(wall_shift, e->wshift, (size_t)e->BX * sizeof(double), hipMemcpyDeviceToHost, st));
    EHIP(e, hipStreamSynchronize(st));
    if (bounds)
        for (int i = 0; i < 5; i++) bounds[i] = e->hbounds[i];
    return BCG_OK;
}

int bcg_clock_init(bcg_clock* c, double t0, double dt, double OutputTime, double UpdateTime) {
    if (!c || !(dt > 0)) return BCG_ERR_ARG;
    c->t = t0;
    c->dt = dt;
    c->OutputTime = OutputTime;
    c->UpdateTime = UpdateTime;
    c->HeightDensityUpdateTime = UpdateTime / 10.0;
    c->NextOutTime = OutputTime;
    c->NextUpdateTime = UpdateTime;
    c->HeightDensityNextUpdateTime = UpdateTime / 10.0;
    c->OutFlag = 0;  // the initial output happens before the loop
    c->UpdateFlag = 1;
    c->HeightDensityUpdateFlag = 1;
    return BCG_OK;
}

int bcg_clock_advance(bcg_clock* c) {
    if (!c) return BCG_ERR_ARG;
    c->t += c->dt;
    c->OutFlag = (c->NextOutTime <= 0);
    c->NextOutTime = (c->OutFlag ? c->OutputTime : c->NextOutTime) - c->dt;
    c->UpdateFlag = (c->NextUpdateTime <= 0);
    c->NextUpdateTime = (c->UpdateFlag ? c->UpdateTime : c->NextUpdateTime) - c->dt;
    c->HeightDensityUpdateFlag = (c->HeightDensityNextUpdateTime <= 0);
    c->HeightDensityNextUpdateTime =
        (c->HeightDensityUpdateFlag ? c->HeightDensityUpdateTime : c->HeightDensityNextUpdateTime) - c->dt;
    return BCG_OK;
}

int bcg_sim_prepare(bcg_ctx* ctx, bcg_env* env, const bcg_clock* clock) {
    if (!ctx || !clock) return BCG_ERR_ARG;
    if (env && env->mech != ctx) return BCG_ERR_ARG;
    int rc;
    if (clock->UpdateFlag) {
        if (env) {
            if ((rc = bcg_env_fields(env, BCG_ENV_UPDATE))) return rc;
            if ((rc = bcg_env_nutrients(env, nullptr))) return rc;
        }
        // getNeighbours (ABM/Simulation.cpp:290)
        if ((rc = bcg_rebuild_neighbours(ctx, nullptr, -1))) return rc;
    }
    if (clock->HeightDensityUpdateFlag && env)
        if ((rc = bcg_env_fields(env, BCG_ENV_HEIGHT))) return rc;
    return BCG_OK;
}

int bcg_sim_advance(bcg_ctx* ctx, bcg_clock* clock, bcg_cell* moved, int32_t moved_cap, int32_t* n_div,
                    int32_t* lineage, int32_t lineage_cap) {
    if (!ctx || !clock) return BCG_ERR_ARG;
    int rc = bcg_step_out(ctx, clock->dt, clock->t, 0, moved, moved_cap, n_div, lineage, lineage_cap);
    if (rc) return rc;
    return bcg_clock_advance(clock);
}

int bcg_sim_step(bcg_ctx* ctx, bcg_env* env, bcg_clock* clock, int32_t* n_div, int32_t* lineage, int32_t lineage_cap) {
    int rc = bcg_sim_prepare(ctx, env, clock);
    if (rc) return rc;
    return bcg_sim_advance(ctx, clock, nullptr, 0, n_div, lineage, lineage_cap);
}

}  // extern "C"
