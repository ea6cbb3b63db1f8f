// env_kernels.h - device side of the planar colony environment (bcg_env_*):
// density and footprint, height and surface normal, and the nutrient update
// with its nested agar multigrid.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bcg.h"

namespace bcg {
namespace env {

constexpr int MG_LEVELS = 8;      // Multigrid(maxLevelsMG, 8, ...) at ABM/Simulation.cpp:150
constexpr int MAX_NEST = 4;       // nested agar levels (maxLevelsMG) supported
constexpr int MAX_REF = 8;        // refinementGridHeight supported
constexpr int ENV_THREADS = 512;  // the nutrient update runs as one workgroup (up to 256 VGPRs a lane)
constexpr int ENV_LDS_BYTES = 144 * 1024;  // dynamic LDS for the colony window of the sweeps

enum ErrBits {
    ERR_RANGE = 1,       // a stencil or filter would read outside an array
    ERR_CELLTYPE = 2,    // a cell of Type other than 1 or 2 (the reference prints "CellType error!")
    ERR_COLONY_EDGE = 4, // a colony point with x <= 2 (the reference asserts "Need more boxes")
};

// Constants of the colony chemistry derived from the parameters
// (UpdateEnvironment_OxygenAcetate, ABM/Nutrients.cpp:200-506): formed once on
// the host into EnvArgs (IEEE divisions, the same values as on the device).
struct ChemConst {
    double mgrCarAers, mgrCarAnas, mgrAceAers, g1s, g2s, g3s;
};

// Everything the device passes need, by value.
struct EnvArgs {
    // grid
    int BX, BZ, BZA, ML, REF, NHY;  // BoxY = 1; NHY = BoxY * refinementGridHeight
    double BoxLength, cellRadius, density_threshold, Cdt;
    // chemistry
    double KC_aer, KC_ana, KW, KO, KA;
    double C_rate, W_rate, Carbon_to_Waste, Maintenance_rate, maxGrowthRate2;
    double DiffColony, DiffAgar, maxCarbon, maxO2, maxO2_Agar;
    double qCarAer, qCarAna, qAceAer, qO2Car, qO2Ace, pAceAer, pAceAna;
    double DiffColonyCarbon, DiffAgarCarbon, DiffColonyO2, DiffAgarO2, DiffColonyAce, DiffAgarAce;
    double maxGrowthRateCarAer, maxGrowthRateCarAna, maxGrowthRateAceAer;
    double Kw, Ka, Kb, Cb, c1;
    double qCarAernt, qCarAnant, qAceAernt, qO2Carnt, qO2Acent, pAceAnant;
    ChemConst cc;  // chem_const of the above
    int crossFeeding;
    int n_outer;  // outer iterations of the nutrient update
    // state
    bcg_env_cell* colony[2];          // [0] Environment, [1] oldEnvironment
    bcg_aga* agar[2][MAX_NEST];       // FieldAgar, oldFieldAgar
    bcg_aga* wall[2][MAX_NEST];       // FieldWall, oldFieldWall
    // Multigrid::Aga, AgaRes, AgaErr: per level the species C, O2, acetate
    // as three blocks of (mgx+1)(mgz+1) values, each [z][x] (x fastest)
    double* aga[MG_LEVELS];
    double* res[MG_LEVELS];
    double* err[MG_LEVELS];
    // the relaxation's residual term r h^2 / dt / D of every point (same
    // layout as res), formed once each time the residual is
    double* rt[MG_LEVELS];
    int mgx[MG_LEVELS];               // x extent - 1 of each multigrid array
    int mgz[MG_LEVELS];               // z extent - 1 of each multigrid array
    // working copies during an update, x fastest: colony C, W, O2, acetate
    // of both buffers [4][BZ][BX]; interface rows and first agar rows of both
    // buffers [ML][4][BX]; insideColonyDen and DensityShiftP [BZ][BX]
    double* scol[2];
    double* swall[2];
    double* srow[2];
    double* insT;
    double* dshT;
    double* phantom;                  // 3 values per thread: multigrid reads past an array's end
    // the sweeps' colony window in dynamic LDS: width maxX - minX + 3 and
    // lds_nf = width * (maxH + 2) boxes a field (lds_colony = 0: global)
    int lds_colony, lds_w, lds_nf;
    // fields (smoothed, shifted to the grid points)
    const double* inside;             // insideColonyDen [BX][BZ]
    const double* dshift;             // DensityShiftP
    const double* d1shift;
    const double* d2shift;
    const double* wshift;             // WallDensityShiftP [BX]
    const int* bounds;                // minx, maxx, miny, maxy, maxz
    int* errflag;
};

inline ChemConst chem_const(const EnvArgs& a) {
    ChemConst c;
    if (a.maxGrowthRateCarAer != 0.0) {
        c.mgrCarAers = a.qCarAernt / a.qCarAer;
        c.g1s = a.KC_aer * (c.mgrCarAers / a.maxGrowthRateCarAer);
    } else {
        c.mgrCarAers = 0;
        c.g1s = 1e6;
    }
    if (a.maxGrowthRateCarAna != 0.0) {
        c.mgrCarAnas = a.qCarAnant / a.qCarAna;
        c.g2s = a.KC_ana * (c.mgrCarAnas / a.maxGrowthRateCarAna);
    } else {
        c.mgrCarAnas = 0;
        c.g2s = 1e6;
    }
    if (a.maxGrowthRateAceAer != 0.0) {
        c.mgrAceAers = a.qAceAernt / a.qAceAer;
        c.g3s = a.KA * (c.mgrAceAers / a.maxGrowthRateAceAer);
    } else {
        c.mgrAceAers = 0;
        c.g3s = 1e6;
    }
    return c;
}

// density / height passes
struct FieldArgs {
    int BX, BZ, REF, NHY, N, FD;
    double BoxLength, cellRadius, density_threshold;
    const double* hot;   // mechanics cell records
    const int* type;
    const int* start;    // bin starts (bins x-major, z fastest)
    const int* sorted;   // cell IDs by bin, ascending within a bin
    double *rough, *rough1, *rough2, *inside;
    double *rwall, *rwall1, *rwall2;
    double *rshift, *rshift1, *rshift2, *rwshift, *rwshift1, *rwshift2;
    double *height, *rheight, *normal;
    int* bounds;
    int* errflag;
};

__global__ void k_env_density(FieldArgs a);
__global__ void k_env_height(FieldArgs a);
__global__ void k_env_halo(FieldArgs a);
__global__ void k_env_pad_bounds(FieldArgs a);
__global__ void k_env_shift(FieldArgs a);
__global__ void k_env_box_average(FieldArgs a, const double* rough, const double* rwall, double* out, double* outwall,
                                  const double* filter3d);
__global__ void k_env_height_average(FieldArgs a);
__global__ void k_env_smooth(FieldArgs a, const double* filter1d);
__global__ void k_env_normal(FieldArgs a);
__global__ void k_env_init_bounds(FieldArgs a);

__global__ void __launch_bounds__(ENV_THREADS) k_env_nutrients(EnvArgs a);

// The same update in phases (one launch per barrier of k_env_nutrients).
constexpr int PH_THREADS = 256;      // block of the grid-wide phase kernels
constexpr int PH_MAX_BLOCKS = 1024;  // their grids are capped here (grid-stride loops)
constexpr int PH_RELAX_THREADS = 64; // block of the relaxation passes (one column walk a lane)
// per-thread slots of EnvArgs::phantom: every thread of any phase launch
constexpr int PHANTOM_THREADS = PH_THREADS * PH_MAX_BLOCKS;
__global__ void __launch_bounds__(PH_THREADS) k_nut_begin(const EnvArgs* a);
__global__ void __launch_bounds__(PH_THREADS) k_nut_copy_to_mg(const EnvArgs* a, int sw);
__global__ void __launch_bounds__(PH_THREADS) k_nut_reset_zero(const EnvArgs* a);
__global__ void __launch_bounds__(PH_THREADS) k_nut_update_sol(const EnvArgs* a);
__global__ void __launch_bounds__(PH_THREADS) k_nut_update_res(const EnvArgs* a);
__global__ void __launch_bounds__(PH_RELAX_THREADS) k_nut_relax(const EnvArgs* a, int l, int par);
__global__ void __launch_bounds__(PH_THREADS) k_nut_restrict(const EnvArgs* a, int l);
__global__ void __launch_bounds__(PH_THREADS) k_nut_interp(const EnvArgs* a, int l);
__global__ void __launch_bounds__(PH_THREADS) k_nut_rterm(const EnvArgs* a, int l);
// coarse_mode: 2 coarsest level a wave per species, 1 the workgroup (both out
// of dynamic LDS), 0 the one-wave or global path
__global__ void __launch_bounds__(ENV_THREADS) k_nut_vcycle_tail(EnvArgs a, int lc, int coarse_mode);
// the V-cycle tail (one workgroup) starts at the first level whose column
// walks of a colour fit the workgroup and whose points are at most this many:
// on wider levels its restriction and interpolation, a few points a thread,
// wait on memory with only the workgroup's waves to hide it
constexpr long TAIL_MAX_POINTS = 4096;
// the coarsest level in LDS for the tail's coarse solve (E, R and RT), up to this size
constexpr int COARSE_LDS_BYTES = 60 * 1024;
__global__ void __launch_bounds__(PH_THREADS) k_nut_rewrite(const EnvArgs* a, int l);
__global__ void __launch_bounds__(PH_THREADS) k_nut_nest(const EnvArgs* a, int l);
__global__ void __launch_bounds__(PH_THREADS) k_nut_copy_from_mg(const EnvArgs* a);
__global__ void __launch_bounds__(PH_THREADS) k_nut_sweep(const EnvArgs* a, int full, int dir, int sw);
__global__ void __launch_bounds__(ENV_THREADS) k_nut_sweeps_wg(EnvArgs a, int last);
__global__ void __launch_bounds__(PH_THREADS) k_nut_end(const EnvArgs* a);
__global__ void k_env_growth_to_mech(const bcg_env_cell* colony, int n, double* gr1, double* gr2);

}  // namespace env
}  // namespace bcg
