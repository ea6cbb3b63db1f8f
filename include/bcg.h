/*
 * bcg.h - C ABI of the MI355X (gfx950) implementation of the BacColonyGrowth
 * per-timestep hot path: spherocylinder mechanics on a neighbour grid.
 *
 * Every entry point returns an int status (BCG_OK == 0); no C++ exception
 * crosses this boundary. Host buffers belong to the caller, device buffers to
 * the context. One context per host thread; each context orders its work on
 * its own HIP stream.
 *
 * Completion. A call that takes or returns host buffers, or returns a count
 * or status computed on the device, returns after its work has finished.
 * Three calls return before their device work finishes, by design:
 * bcg_rd3d_deposit_cells (when n_inside is NULL), bcg_rd3d_growth_to_cells
 * and bcg_rd3d_update_cells. They are queued on the nutrient context's
 * stream behind the mechanics context's queued work; the mechanics context
 * then waits for them on the device (no host synchronisation): bcg_step
 * before its move/grow kernel, so that step's bins and lists run alongside
 * the nutrient work; every other call on the mechanics context on entry.
 * A call on the nutrient context that reads its field back
 * (bcg_rd3d_get_field) waits for them too.
 *
 * Failure. A bcg_step that fails after its move/grow kernel was queued
 * (list overflow or an out-of-domain daughter during the list patch) leaves
 * the friction accumulators, the rand() position and the lists advanced but
 * the cell count and buffers not swapped: the context is then marked failed
 * and every call that reads its cells returns BCG_ERR_STATE until
 * bcg_set_cells uploads a state again.
 *
 * The reference has no plugin or FFI layer: its seam is a set of free C++
 * functions that take extern globals as parameters. Each entry point below
 * names the reference function (file:line, relative to
 * AgentBasedModel-BacColonyGrowth/) whose behaviour it replaces.
 */
#ifndef BCG_H_
#define BCG_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BCG_OK                      0
#define BCG_ERR_ARG                -1  /* bad argument or size                          */
#define BCG_ERR_CAPACITY           -2  /* more cells than bcg_params.maxCells            */
#define BCG_ERR_OUT_OF_DOMAIN      -3  /* a centre of mass lies outside the bin grid     */
#define BCG_ERR_NEIGHBOUR_OVERFLOW -4  /* a reduced neighbour list exceeds maxNeighbours */
#define BCG_ERR_HIP                -5  /* HIP runtime error (text in bcg_last_error)     */
#define BCG_ERR_UNSUPPORTED        -6  /* parameter combination not implemented          */
#define BCG_ERR_STATE              -7  /* call out of order (e.g. no cells uploaded)     */

/* ---- cell record -------------------------------------------------------
 * Byte-identical to the reference's `struct Cell` (ABM/Cell.h:15-28 with
 * `Segment` from ABM/tools.h:71-87): 208 bytes, little-endian, no padding.
 * This is the record of the reference's Restart files
 * (ABM/InputOutput.cpp:434-448): 16-byte header (t, dt) then Cell[N].        */
typedef struct bcg_cell {
    double p[3];              /* Position.p                        */
    double q[3];              /* Position.q                        */
    double time_p, time_q;    /* Position.time_p / time_q          */
    int32_t age_p, age_q;     /* Position.age_p / age_q            */
    double Length;
    double Ldiv;
    double Radius;
    double GrowthRate;
    double Velocity[3];
    double AngularVelocity[3];
    double DynFric[3];
    double StaFric[3];
    int32_t Type;
    int32_t Ancestor;
} bcg_cell;

/* ---- parameters --------------------------------------------------------
 * Names follow the reference's globals (ABM/Constants.cpp:1-110) and the
 * input-file keys parsed at ABM/InputOutput.cpp:481-661.                    */
typedef struct bcg_params {
    /* physical constants (ABM/Constants.cpp:1-15, :55) */
    double cellRadius;         /* "Radius"            */
    double L_divide;           /* "L_divide"          */
    double k_cc;               /* "k_cc"              */
    double k_wc;               /* "k_wc"              */
    double varL;               /* "var_L"             */
    double varAngle;           /* "var_angle"         */
    double var_pos;            /* "var_pos"  (only 0 is supported, see below) */
    double viscosity;          /* "Viscosity"         */
    double wall_rough;         /* "Wall_Rough"        */
    double gamma_t;            /* "Gamma"             */
    double wall_mu;            /* "Wall_Mu"           */
    double cell_mu;            /* "Cell_Mu"           */
    double density_threshold;  /* "Density_Threshold" */
    double tension;            /* "Surface_Tension"   */
    double DH;                 /* "Delta_H"           */
    double Rc;                 /* "Rc" (default 3.0)  */
    /* neighbour bins = UniformGrid (ABM/CellMD3D.cpp:52-55) */
    double BoxLength;          /* bin edge; the reference uses L_divide + 2*Radius */
    int32_t BoxX, BoxY, BoxZ;  /* "Box_x", "Box_y", "Box_z" */
    int32_t refinementGridHeight;
    int32_t maxCells;          /* "MaxCells": capacity of the cell buffers (<= 2^28) */
    int32_t maxNeighbours;     /* capacity of one reduced neighbour list            */
    int32_t crossFeeding;      /* "CrossFeeding": Type 2 grows with GrowthRate2     */
    int32_t planar;            /* 1: reference x-z model (y forced to 0),
                                  0: full 3-D generalisation                        */
} bcg_params;

/* Opaque context: device-resident cell state, bins, lists, fields, RNG. */
typedef struct bcg_ctx bcg_ctx;

/* Fill `p` with the reference defaults (ABM/Constants.cpp:1-110) overlaid
 * with the demo input (ABM/run_sim/inputFile_20mM_SmallSimulation.txt),
 * i.e. what the reference's ReadParameters (ABM/InputOutput.cpp:450) yields
 * for that file, and BoxLength = L_divide + 2*Radius (ABM/CellMD3D.cpp:52). */
int bcg_default_params(bcg_params* p);

/* Replaces the allocation block of main (ABM/CellMD3D.cpp:46-55,82-104,
 * 138-142): cell double buffers, UniformGrid, neighbour lists, fields.     */
int bcg_create(const bcg_params* p, bcg_ctx** out);
int bcg_destroy(bcg_ctx* ctx);

/* Text of the last error on this context (NUL-terminated, truncated). */
int bcg_last_error(const bcg_ctx* ctx, char* buf, int len);

/* Upload N cells in ID order. Replaces LoadCells (ABM/InputOutput.cpp:394-431)
 * minus the file read: it also places every cell in its bin and clears the
 * neighbour lists (call bcg_rebuild_neighbours before stepping).            */
int bcg_set_cells(bcg_ctx* ctx, const bcg_cell* cells, int32_t n);
/* Download the current cells (after the last step's buffer swap). */
int bcg_get_cells(bcg_ctx* ctx, bcg_cell* cells, int32_t cap, int32_t* n);
int bcg_num_cells(bcg_ctx* ctx, int32_t* n);

/* Fields read by the mechanics, in the reference layouts (x-major):
 *   height  [BoxX*ref][BoxY*ref]     DoubleArray2D Height (ABM/CellMD3D.cpp:82)
 *   normal  [BoxX*ref][BoxY*ref][3]  CoordArray2D Normal  (ABM/CellMD3D.cpp:138)
 *   wall    [BoxX][BoxY]             DoubleArray2D Wall   (ABM/Simulation.cpp:161)
 *   gr1,gr2 [BoxX][BoxY][BoxZ]       Environment(x,y,z).GrowthRate1/2
 *                                    (EnvArray3D, read by grow at ABM/grow.cpp:26-36)
 * Any pointer may be NULL to leave that field unchanged.                    */
int bcg_set_fields(bcg_ctx* ctx, const double* height, const double* normal,
                   const double* wall, const double* gr1, const double* gr2);
/* Read the same fields back (what the environment passes wrote).          */
int bcg_get_fields(bcg_ctx* ctx, double* height, double* normal, double* wall, double* gr1, double* gr2);

/* glibc-compatible rand() stream owned by the context (the reference draws
 * from the process-wide rand(): ABM/CellMD3D.cpp:23, ABM/Forces.cpp:102,
 * ABM/grow.cpp:109-152, ABM/Simulation.cpp:166). */
int bcg_srand(bcg_ctx* ctx, uint32_t seed);
int bcg_rand(bcg_ctx* ctx, int32_t* out);   /* one draw, as rand() */
/* Host-only (no device needed): out[0..n) = the draws number skip .. skip+n-1
 * of the stream after srand(seed), skipping by jump-ahead (O(log skip)).   */
int bcg_rand_stream(uint32_t seed, uint64_t skip, int32_t n, int32_t* out);

/* Neighbour lists. Replaces getNeighbours (ABM/Neighbours.cpp:24-55) with the
 * UniformGrid 27-box gather (ABM/UniformGrid.h:141-170) and the
 * reduceNeighbours filter (ABM/Neighbours.cpp:57-93). ids == NULL (or n < 0)
 * rebuilds every cell; otherwise only the listed cells, sequentially as the
 * reference does.                                                          */
int bcg_rebuild_neighbours(bcg_ctx* ctx, const int32_t* ids, int32_t n);
/* counts[N]; lists[N*stride], each list sorted ascending by neighbour ID.   */
int bcg_get_neighbours(bcg_ctx* ctx, int32_t* counts, int32_t* lists, int32_t stride);

/* sum_forces (ABM/Forces.cpp:332-381) on the current cells against their
 * current lists: F[3N], T[3N]. Diagnostic: consumes no rand() draws.       */
int bcg_sum_forces(bcg_ctx* ctx, double* F, double* T);
/* mean_stress (ABM/Compute.cpp:18-117), as the output step calls it for
 * every cell (ABM/Simulation.cpp:320-326), on the current cells against
 * their current lists: stress[5N] = (xx, yy, zz, rr, tt) per cell, the
 * diagonal of the contact stress tensor over the cell volume and its radial
 * and azimuthal components; fnet[3N] the net contact force. Each F_cw call
 * takes one rand() draw of the context's stream, as in the reference.      */
int bcg_mean_stress(bcg_ctx* ctx, double* stress, double* fnet);

/* One mechanics step = the body of the reference time loop after the field
 * update (ABM/Simulation.cpp:330-413):
 *   MoveCell + GrowCell for every cell (ABM/Cells.cpp:14-57), the division
 *   loop with DivideCell and the neighbour-list patch (:380-404), then the
 *   old/new buffer swap (:409-413).
 * flags bit 0 (BCG_STEP_REBUILD): first rebuild every neighbour list from
 *   the current cells, as the reference does at each field-update step (:292).
 * flags bit 1 (BCG_STEP_NO_PATCH): skip the post-division list patch; the
 *   caller promises BCG_STEP_REBUILD on the next step (the lists are then
 *   rebuilt before anything reads them, so the result is unchanged).
 * lineage (may be NULL) receives (parent, daughter) pairs, lineage_cap pairs
 * at most; *n_div receives the number of divisions performed. The draws of
 * the step (one per F_cw call, nine per division) come from the context's
 * rand() stream in the reference's order.                                 */
#define BCG_STEP_REBUILD  1
#define BCG_STEP_NO_PATCH 2
int bcg_step(bcg_ctx* ctx, double dt, double t, int32_t flags,
             int32_t* n_div, int32_t* lineage, int32_t lineage_cap);

/* bcg_step that also copies, when `moved` is non-NULL, the moved and grown
 * cells before the divisions (N of them, N the count before the step) to
 * moved[moved_cap]: the array the reference's output step saves to its
 * Restart file at that point (SaveCells of new_cells, ABM/Simulation.cpp:349).
 * Adds one host synchronisation and the copy.                              */
int bcg_step_out(bcg_ctx* ctx, double dt, double t, int32_t flags, bcg_cell* moved, int32_t moved_cap,
                 int32_t* n_div, int32_t* lineage, int32_t lineage_cap);

/* Counters of the last bcg_step: wall-contact draws (calls of F_cw), cells
 * that reached Ldiv, and maximum reduced list length seen.                 */
int bcg_last_step_stats(bcg_ctx* ctx, int64_t* wall_draws, int32_t* n_flagged,
                        int32_t* max_list);

/* Per-phase device time of the last bcg_step in milliseconds, from HIP
 * events on the context's stream: [0] bins, [1] lists, [2] the move+grow
 * kernel alone, [3] division compaction, [4] divisions and list patch.      */
int bcg_last_step_times(bcg_ctx* ctx, float* ms5);


/* ======================================================================
 * 3-D nutrient step: implicit reaction-diffusion with Monod uptake.
 *
 * The reference solves the nutrient field only in its planar x-z model
 * (UpdateEnvArrayMG_OxygenAcetate, ABM/Multigrid.cpp:984-1125, with the
 * Crank-Nicolson multigrid of ABM/Multigrid.cpp:5-981 and the Monod colony
 * chemistry of ABM/Nutrients.cpp:200-506). These entry points are its 3-D
 * generalisation: one implicit step on a vertex grid (x fastest)
 *     (1 + dt*q/(K + c_old)) c_new - dt*D*Lap_h c_new = c_old,
 * Neumann on the x, y and top faces, Dirichlet c = c_res on z = 0 (the
 * agar reservoir); q is the uptake-rate density deposited from the cells
 * (the role of GetDensity/ShiftDensity, ABM/Compute.cpp:120-463).
 * ====================================================================== */
typedef struct bcg_rd3d_params {
    int32_t nx, ny, nz;     /* vertices per direction (2^k+1 coarsens fully) */
    double h;               /* vertex spacing, um                            */
    double origin[3];       /* position of vertex (0,0,0), um                */
    double D;               /* diffusivity, um^2/s                           */
    double dt;              /* time step, s                                  */
    double K;               /* Monod constant, mM                            */
    double c_res;           /* Dirichlet value on z = 0, mM                  */
    int32_t nu1, nu2;       /* red-black Gauss-Seidel pre/post sweeps        */
    int32_t coarse_sweeps;  /* sweeps on the coarsest level (FP64)           */
    int32_t max_levels;     /* multigrid levels at most                      */
    int32_t inner;          /* FP32 V-cycles per refinement step (mode 1)    */
    int32_t smoother;       /* BCG_RD3D_SMOOTH_* (default FUSED)             */
} bcg_rd3d_params;

typedef struct bcg_rd3d bcg_rd3d;

#define BCG_RD3D_FP64  0   /* FP64 multigrid V-cycles                         */
#define BCG_RD3D_MIXED 1   /* FP32 V-cycles inside FP64 iterative refinement  */

/* Red-black Gauss-Seidel sweeps on the large levels: both colours in one
 * z-marching launch (FUSED), or one launch per colour (SPLIT). Both give the
 * same iterates; SPLIT is kept as the plain form to check FUSED against.  */
#define BCG_RD3D_SMOOTH_FUSED 0
#define BCG_RD3D_SMOOTH_SPLIT 1
/* FUSED, and the FP64 sweeps of levels of >= 2^24 points on the 62 x 30 tile
   as well (87 KB of LDS: one workgroup a CU); FP32 always uses it there */
#define BCG_RD3D_SMOOTH_FUSED_WIDE 2

int bcg_rd3d_default_params(bcg_rd3d_params* p);
int bcg_rd3d_create(const bcg_rd3d_params* p, bcg_rd3d** out);
int bcg_rd3d_destroy(bcg_rd3d* ctx);
int bcg_rd3d_last_error(const bcg_rd3d* ctx, char* buf, int len);
/* number of multigrid levels and their (nx, ny, nz), cap levels at most */
int bcg_rd3d_levels(bcg_rd3d* ctx, int32_t* n, int32_t* dims, int32_t cap);
/* nutrient field [nz][ny][nx], host buffers */
int bcg_rd3d_set_field(bcg_rd3d* ctx, const double* c);
int bcg_rd3d_get_field(bcg_rd3d* ctx, double* c);
/* uptake-rate density q [nz][ny][nx] (rate per um^3), host buffers */
int bcg_rd3d_set_uptake(bcg_rd3d* ctx, const double* q);
int bcg_rd3d_get_uptake(bcg_rd3d* ctx, double* q);
/* q from cells: cloud-in-cell deposition of rate[i] at centre cm[3i..3i+2]
 * divided by h^3; cells outside the grid are dropped (*n_inside counts the
 * rest). Deterministic: every vertex sums its cells in a fixed order.     */
int bcg_rd3d_deposit(bcg_rd3d* ctx, const double* cm, const double* rate, int32_t n, int32_t* n_inside);
/* One implicit step from the current field (which it replaces), solved to
 * the relative residual ||b - A c|| <= tol * ||b|| starting from c_old.
 * *iters = V-cycles (mode 0) or refinement steps (mode 1); *relres the
 * final ||b - A c|| / ||b||. BCG_ERR_STATE if tol is not reached.          */
int bcg_rd3d_step(bcg_rd3d* ctx, int32_t mode, double tol, int32_t max_iter, int32_t* iters, double* relres);
/* Make (c_old, c_new) the state of the last step without solving: the
 * adjoint of a recorded step is then available (checkpoint + recompute).  */
int bcg_rd3d_set_state(bcg_rd3d* ctx, const double* c_old, const double* c_new);
/* Change the diffusivity D and the Monod constant K. */
int bcg_rd3d_set_coeffs(bcg_rd3d* ctx, double D, double K);
/* Adjoint of the last step: given g = dL/dc_new, solves A^T lambda = g (as
 * A mu = W^-1 g, lambda = W mu, W the control-volume weights that make W A
 * symmetric) by the same multigrid, and returns (each may be NULL)
 *   dL_dq[N]    gradient w.r.t. the uptake density q
 *   dL_dcold[N] gradient w.r.t. c_old (chains steps backwards)
 *   *dL_dK, *dL_dD                                                          */
int bcg_rd3d_adjoint(bcg_rd3d* ctx, const double* g, int32_t mode, double tol, int32_t max_iter,
                     double* dL_dq, double* dL_dcold, double* dL_dK, double* dL_dD);
/* bcg_rd3d_deposit without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Inputs and output are DEVICE pointers: d_cm[3n] centres,
 * d_rate[n] rates, d_q[nz*ny*nx] receives the uptake density (the same
 * values, bit for bit, that bcg_rd3d_deposit leaves in the context). The
 * context is used for its grid only; its own uptake field is not touched.
 * The work is queued on `stream` (a hipStream_t; NULL: the context's
 * stream) and the call returns without waiting for it. Its scratch (keys,
 * fractions, bucket counts and starts) is one static device array per GPU
 * of BCG_NOALLOC_SCRATCH_BYTES (16 MiB, held until the process ends, e.g.
 * 2^18 sources on 129 x 129 x 33 vertices); larger cases return
 * BCG_ERR_CAPACITY. Every call on a GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race and must be
 * serialised by the caller.                                                 */
#define BCG_NOALLOC_SCRATCH_BYTES (16 << 20)
int bcg_rd3d_deposit_noalloc(bcg_rd3d* ctx, const double* d_cm, const double* d_rate, int32_t n, double* d_q,
                             void* stream);
/* The system of the last set-up (bcg_rd3d_step or bcg_rd3d_set_state: the
 * same a0 and D) for another right-hand side: A x = rhs from x = 0, to
 * ||rhs - A x|| <= tol ||rhs|| (Dirichlet rows: x = rhs). x[N] (host)
 * receives the last iterate, also when tol is not reached (BCG_ERR_STATE);
 * the context's field and right-hand side are left as they were. No
 * counterpart in the reference: it serves sensitivity solves, e.g. the
 * difference of two steps whose operators differ, A1 (c1 - c2) =
 * (A2 - A1) c2, solved to the precision of the difference itself.         */
int bcg_rd3d_solve(bcg_rd3d* ctx, int32_t mode, const double* rhs, double tol, int32_t max_iter, double* x,
                   int32_t* iters, double* relres);
/* Transpose of bcg_rd3d_deposit: out[i] = sum_v field[v] w_iv / h^3.     */
int bcg_rd3d_gather(bcg_rd3d* ctx, const double* cm, int32_t n, const double* field, double* out);
/* The 3-D step coupled to a mechanics context on the same device, without
 * host copies. No counterpart in the reference, whose nutrient model is
 * planar; these take the roles of GetDensity (ABM/Compute.cpp:120-253) and
 * of the growth-rate lookup of grow (ABM/grow.cpp:17-30):
 *   bcg_rd3d_deposit_cells: q from the context's current cells, each
 *     depositing rate * Length at its centre (cloud-in-cell, as
 *     bcg_rd3d_deposit); *n_inside may be NULL (non-NULL synchronises);
 *   bcg_rd3d_growth_to_cells: the context's growth rate of every bin
 *     (GrowthRate1, before the reference's factor 1.5850) = mu c / (K + c),
 *     c the current field at the bin centre (trilinear, clamped to the grid).
 * Each runs after the context's queued work. The context then waits for
 * them (stream-ordered, no host synchronisation) before its next kernel that
 * reads growth rates or writes cells: in bcg_step before move/grow, so that
 * step's bins and lists run alongside the nutrient work; every other call on
 * the context waits on entry.                                              */
int bcg_rd3d_deposit_cells(bcg_rd3d* ctx, bcg_ctx* cells, double rate, int32_t* n_inside);
int bcg_rd3d_growth_to_cells(bcg_rd3d* ctx, bcg_ctx* cells, double mu, double K);
/* One coupled nutrient update without host synchronisation:
 * bcg_rd3d_deposit_cells, then the implicit step from the current field by
 * `vcycles` FP64 V-cycles (no convergence test), then
 * bcg_rd3d_growth_to_cells. The fixed-shape part (set-up, V-cycles, growth
 * rates) is recorded as a HIP graph on first use for this (cells, vcycles,
 * mu, K) and replayed; the results equal those of the separate calls.     */
int bcg_rd3d_update_cells(bcg_rd3d* ctx, bcg_ctx* cells, double rate, double mu, double K, int32_t vcycles);
/* One V-cycle of the current step's system (timing); *ms = its duration.  */
int bcg_rd3d_vcycle(bcg_rd3d* ctx, int32_t mode, float* ms);
/* device times of the last step, ms: [0] setup (a0, rhs, coarse a0), [1] solve */
int bcg_rd3d_last_times(bcg_rd3d* ctx, float* ms2);

/* ---- planar colony environment ------------------------------------------
 * The fields of the reference's x-z model around the mechanics: cell
 * density and colony footprint (ABM/Compute.cpp:120-253, 344-605), height
 * and surface normal (ABM/Compute.cpp:256-341, 470-493, 640-738), and the
 * nutrient update of the colony, the colony-agar interface and the nested
 * agar multigrid (ABM/Multigrid.cpp:5-1125, ABM/Nutrients.cpp:200-659,
 * 806-855). BoxY must be 1 (the reference's planar build). The context is
 * bound to a mechanics context: it reads its cells and writes the fields the
 * mechanics reads (Height, Normal, Environment.GrowthRate1/2).              */

/* One colony box of the reference's Environment: `struct LocalEnv`
 * (ABM/Nutrients.h:11-62), 17 doubles, same order.                         */
typedef struct bcg_env_cell {
    double Carbon, Waste, GrowthRate1, GrowthRate2, Oxygen, Acetate;
    double pH, Buffer, Acetic, Hydrox;
    double qCarmnt, qAcemnt, pAcemnt, qOxymnt;
    double GrCarAer, GrCarAna, GrAceAer;
} bcg_env_cell;

/* One agar / wall point: `struct LocalAga` (ABM/Nutrients.h:64-79).        */
typedef struct bcg_aga {
    double CarbonAgar, WasteAgar, OxygenAgar, AcetateAgar;
} bcg_aga;

/* Constants of the environment; names follow ABM/Constants.cpp (the input
 * keys of ABM/InputOutput.cpp:481-661 in comments where they differ).      */
typedef struct bcg_env_params {
    int32_t BoxX, BoxY, BoxZ;          /* colony boxes (BoxY = 1)            */
    int32_t BoxZAgar;                  /* "Box_z_agar"                      */
    int32_t maxLevelsMG;               /* nested agar levels                */
    int32_t refinementGridHeight;
    int32_t FilterLen;
    int32_t crossFeeding;              /* "CrossFeeding"                    */
    int32_t NutrientGSI;               /* only 0 (Jacobi) is supported      */
    int32_t InterfaceCondition;        /* only 1 is defined                 */
    double BoxLength, cellRadius, density_threshold;
    double UpdateTime;                 /* "Update_Time" (hours)             */
    double Cdt;
    double KC_aer, KC_ana, KW, KO, KA, KI;
    double C_rate, W_rate, Carbon_to_Waste, Maintenance_rate;
    double maxGrowthRate1, maxGrowthRate2;            /* "Growth_Rate1/2"   */
    double DiffColony, DiffAgar, maxCarbon;
    double maxO2;                                     /* "Max_Oxygen"       */
    double maxO2_Agar;
    double qCarAer, qCarAna, qAceAer, qO2Car, qO2Ace, pAceAer, pAceAna;
    double DiffColonyCarbon, DiffAgarCarbon, DiffColonyO2, DiffAgarO2, DiffColonyAce, DiffAgarAce;
    double maxGrowthRateCarAer, maxGrowthRateCarAna, maxGrowthRateAceAer;
    double Kw, Ka, Kb, Cb, c1;
    double qCarAernt, qCarAnant, qAceAernt, qO2Carnt, qO2Acent, pAceAnant;
} bcg_env_params;

typedef struct bcg_env bcg_env;

/* Reference defaults overlaid with the demo input file (as bcg_default_params). */
int bcg_env_default_params(bcg_env_params* p);
/* Replaces the field allocation of main (ABM/CellMD3D.cpp:56-104) and the
 * Multigrid construction (ABM/Simulation.cpp:150-151). Fields start as the
 * reference's constructors leave them (LocalEnv(), LocalAga()).            */
int bcg_env_create(const bcg_env_params* p, bcg_ctx* mech, bcg_env** out);
int bcg_env_destroy(bcg_env* env);
int bcg_env_last_error(const bcg_env* env, char* buf, int len);
/* Caps the on-chip memory (LDS) the nutrient update may use for the colony
 * boxes it sweeps (default: what the device allows, up to 144 KiB; 0: keep
 * them in global memory). Both placements give the same results. No
 * counterpart in the reference (a placement choice of this build).        */
int bcg_env_set_lds_limit(bcg_env* env, int64_t bytes);
/* How the nutrient update runs (no counterpart in the reference: placement
 * choices of this build; every choice gives the same results, bit for bit).
 * Default 0: each barrier of the update is a launch boundary; the
 * relaxation passes, restrictions and prolongations of the wide multigrid
 * levels and the grid-wide passes use the whole device; the narrow levels of
 * each V-cycle run in one workgroup; the outer iterations are replayed from
 * one recorded HIP graph.                                                   */
#define BCG_ENV_EXEC_ONE_WORKGROUP 1  /* the whole update in one workgroup  */
#define BCG_ENV_EXEC_NO_GRAPH      2  /* launch every phase directly        */
#define BCG_ENV_EXEC_WG_SWEEPS     4  /* the 200 sweeps of an outer iteration in one workgroup */
#define BCG_ENV_EXEC_GRID_SWEEPS   8  /* one device-wide launch per sweep   */
int bcg_env_set_execution(bcg_env* env, int32_t flags);
/* Nutrient state in the reference layouts (x-major, z fastest):
 *   colony, colony_old  [BoxX][BoxY][BoxZ]                 Environment, oldEnvironment
 *   agar, agar_old      [maxLevelsMG][BoxX][BoxY][BoxZAgar] FieldAgar, oldFieldAgar
 *   wall, wall_old      [maxLevelsMG][BoxX][BoxY]          FieldWall, oldFieldWall
 * Any pointer may be NULL (set: left unchanged; get: not copied).           */
int bcg_env_set_state(bcg_env* env, const bcg_env_cell* colony, const bcg_env_cell* colony_old,
                      const bcg_aga* agar, const bcg_aga* agar_old, const bcg_aga* wall, const bcg_aga* wall_old);
int bcg_env_get_state(bcg_env* env, bcg_env_cell* colony, bcg_env_cell* colony_old, bcg_aga* agar,
                      bcg_aga* agar_old, bcg_aga* wall, bcg_aga* wall_old);

#define BCG_ENV_INIT    0  /* the pass before the time loop (ABM/Simulation.cpp:187-211)   */
#define BCG_ENV_UPDATE  1  /* the UpdateFlag pass (ABM/Simulation.cpp:272-282)            */
#define BCG_ENV_HEIGHT  2  /* the HeightDensityUpdateFlag pass (ABM/Simulation.cpp:308-312) */

/* Fields from the mechanics context's current cells, on the device:
 *   INIT, UPDATE: GetDensity, ShiftDensity and six BoxAverage; then for
 *                 UPDATE Height_Average + Smooth; then GetSurfaceNormal.
 *   HEIGHT:       GetHeight, Height_Average, Smooth, GetSurfaceNormal.
 * Height and Normal go to the mechanics context. BCG_ERR_OUT_OF_DOMAIN when
 * a filter or stencil would read outside the arrays (the reference reads
 * out of bounds there).                                                     */
int bcg_env_fields(bcg_env* env, int32_t pass);
/* UpdateEnvArrayMG_OxygenAcetate (ABM/Multigrid.cpp:984-1125), single
 * threaded semantics, from the last INIT/UPDATE fields; *count = its return
 * value. GrowthRate1/2 of the resulting Environment go to the mechanics.
 * Multigrid elements are addressed by the reference's flat index, so a z
 * past the top of a level reads the next column as there; an index past
 * the end of an array (the coarsest level when BoxZAgar < 2^(8 -
 * maxLevelsMG), e.g. the demo input) reads 0 where the reference reads
 * whatever memory follows.                                                  */
int bcg_env_nutrients(bcg_env* env, int32_t* count);
/* Field read-back for tests and output, reference layouts (NULL: skip):
 *   density, density_shift, inside  [BoxX][BoxY][BoxZ]  Density, DensityShiftP, insideColonyDen
 *   wall_shift [BoxX][BoxY]                              WallDensityShiftP
 *   bounds[5] = minx, maxx, miny, maxy, maxz (after GetDensity/GetHeight)   */
int bcg_env_get_fields(bcg_env* env, double* density, double* density_shift, double* inside, double* wall_shift,
                       int32_t* bounds);

/* ---- time loop cadence ---------------------------------------------------
 * The flags of the reference time loop (ABM/Simulation.cpp:49-58, 415-426):
 * set by decrementing doubles, so the periods are what double arithmetic
 * makes them.                                                               */
typedef struct bcg_clock {
    double t, dt;
    double OutputTime, UpdateTime, HeightDensityUpdateTime;
    double NextOutTime, NextUpdateTime, HeightDensityNextUpdateTime;
    int32_t OutFlag, UpdateFlag, HeightDensityUpdateFlag;
} bcg_clock;

int bcg_clock_init(bcg_clock* c, double t0, double dt, double OutputTime, double UpdateTime);
/* t += dt, then the three flags for the next step (ABM/Simulation.cpp:415-426). */
int bcg_clock_advance(bcg_clock* c);

/* One pass of the time loop (ABM/Simulation.cpp:254-427), split at its
 * output point so that a driver's output calls go where the reference has
 * them:
 *   bcg_sim_prepare: UpdateFlag -> fields(UPDATE), nutrients and the list
 *     rebuild (getNeighbours, :290); HeightDensityUpdateFlag ->
 *     fields(HEIGHT) (:296-312). At OutFlag the driver then writes the cell
 *     file from bcg_mean_stress and bcg_get_cells (:313-327; the F_cw
 *     draws of mean_stress come before the step's, as there).
 *   bcg_sim_advance: move + grow; at OutFlag `moved` receives the moved
 *     cells before the divisions, which the reference saves as its Restart
 *     state (:349; NULL: skipped); divisions (lineage as bcg_step); the
 *     buffer swap and the clock advance (:407-426).
 * bcg_sim_step = bcg_sim_prepare + bcg_sim_advance without output. `env`
 * may be NULL (fields stay as set).                                         */
int bcg_sim_prepare(bcg_ctx* ctx, bcg_env* env, const bcg_clock* clock);
int bcg_sim_advance(bcg_ctx* ctx, bcg_clock* clock, bcg_cell* moved, int32_t moved_cap, int32_t* n_div,
                    int32_t* lineage, int32_t lineage_cap);
int bcg_sim_step(bcg_ctx* ctx, bcg_env* env, bcg_clock* clock, int32_t* n_div, int32_t* lineage,
                 int32_t lineage_cap);

#ifdef __cplusplus
}
#endif
#endif /* BCG_H_ */
