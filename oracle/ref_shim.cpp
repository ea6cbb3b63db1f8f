// ref_shim.cpp - test-infrastructure harness around the reference's own
// mechanics functions, compiled together with the reference sources into
// oracle/_ref/libbcgref.so (see oracle/Makefile). It is the checker's
// checker: it lets the tests pin the oracle restatement (bcg_oracle.c)
// against the reference code itself. Nothing in the product links it.
//
// The step function mirrors the order of calls of the reference time loop,
// ABM/Simulation.cpp:330-413 (single thread, so the loop is serial), with
// the reference's MoveCell / GrowCell / DivideCell / getNeighbours doing the
// work. Field arrays are injected instead of computed.
#include <omp.h>
#include <fstream>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "Array.h"
#include "Cell.h"
#include "Cells.h"
#include "Constants.h"
#include "Forces.h"
#include "Integrate.h"
#include "Neighbours.h"
#include "UniformGrid.h"
#include "grow.h"
#include "tools.h"
#include "Compute.h"
#include "InputOutput.h"
#include "Multigrid.h"
#include "Nutrients.h"

#include "../include/bcg.h"

static_assert(sizeof(Cell) == sizeof(bcg_cell), "Cell layout");

static void set_globals(const bcg_params* p) {
    cellRadius = p->cellRadius;
    L_divide = p->L_divide;
    k_cc = p->k_cc;
    k_wc = p->k_wc;
    varL = p->varL;
    varAngle = p->varAngle;
    var_pos = p->var_pos;
    viscosity = p->viscosity;
    wall_rough = p->wall_rough;
    gamma_t = p->gamma_t;
    wall_mu = p->wall_mu;
    cell_mu = p->cell_mu;
    density_threshold = p->density_threshold;
    tension = p->tension;
    DH = p->DH;
    Rc = p->Rc;
    BoxLength = p->BoxLength;
    BoxX = p->BoxX;
    BoxY = p->BoxY;
    BoxZ = p->BoxZ;
    refinementGridHeight = p->refinementGridHeight;
    maxCells = p->maxCells;
    crossFeeding = p->crossFeeding != 0;
}

struct RefSim {
    bcg_params prm;
    int N;
    Cell* old_cells;
    Cell* new_cells;
    int** NL;
    int maxNeighbours;
    UniformGrid* grid;
    DoubleArray2D* Height;
    CoordArray2D* Normal;
    DoubleArray2D* Wall;
    EnvArray3D* Env;
    int* dividing;
    int* IDlist;
};

// OpenMP threads of the reference's own parallel regions (the nutrient
// sweeps of ABM/Multigrid.cpp and the move loop of ref_sim_step_omp); 1 for
// parity runs, where the reference's threaded sweeps race.
static int g_threads = 1;

extern "C" {

void ref_set_threads(int n) { g_threads = n > 0 ? n : 1; }

void ref_set_params(const bcg_params* p) { set_globals(p); omp_set_num_threads(g_threads); }
void ref_srand(unsigned s) { srand(s); }
int ref_rand(void) { return rand(); }

void ref_min_distance(const bcg_params* p, const bcg_cell* a, const bcg_cell* b, double* out7) {
    set_globals(p);
    double d;
    DoubleCoord c1, c2;
    min_distance(*(const Cell*)a, *(const Cell*)b, d, c1, c2);
    out7[0] = d;
    out7[1] = c1.x; out7[2] = c1.y; out7[3] = c1.z;
    out7[4] = c2.x; out7[5] = c2.y; out7[6] = c2.z;
}

void ref_F_cc(const bcg_params* p, const bcg_cell* a, const bcg_cell* b, double* out7) {
    set_globals(p);
    DoubleCoord F, r;
    double d;
    F_cc(*(const Cell*)a, *(const Cell*)b, F, r, d);
    out7[0] = F.x; out7[1] = F.y; out7[2] = F.z;
    out7[3] = r.x; out7[4] = r.y; out7[5] = r.z;
    out7[6] = d;
}

// out18: F1, F2, r1, r2, StaFric, DynFric (accumulated from zero)
void ref_F_cw(const bcg_params* p, const bcg_cell* c, double wall_z, double* out18) {
    set_globals(p);
    DoubleCoord F1(0, 0, 0), F2(0, 0, 0), r1(0, 0, 0), r2(0, 0, 0), sf(0, 0, 0), df(0, 0, 0);
    F_cw(*(const Cell*)c, wall_z, F1, F2, r1, r2, sf, df);
    const DoubleCoord v[6] = {F1, F2, r1, r2, sf, df};
    for (int i = 0; i < 6; i++) {
        out18[3 * i] = v[i].x; out18[3 * i + 1] = v[i].y; out18[3 * i + 2] = v[i].z;
    }
}

void* ref_sim_create(const bcg_params* p) {
    set_globals(p);
    omp_set_num_threads(g_threads);
    RefSim* s = new RefSim();
    s->prm = *p;
    s->N = 0;
    s->old_cells = new Cell[p->maxCells];
    s->new_cells = new Cell[p->maxCells];
    memset((void*)s->old_cells, 0, sizeof(Cell) * p->maxCells);
    memset((void*)s->new_cells, 0, sizeof(Cell) * p->maxCells);
    // ABM/CellMD3D.cpp:54 capacity rule, floored so that small test grids work
    int perBox = int(ceil(BoxLength * BoxLength * BoxLength /
                          (PI * cellRadius * cellRadius * L_divide * density_threshold / 2)));
    if (perBox < 64) perBox = 64;
    s->grid = new UniformGrid(BoxX, BoxY, BoxZ, perBox, BoxLength);
    s->maxNeighbours = s->grid->MaxCellsPerBox() * 27 + 1;
    s->NL = InitializeNeighbourList(p->maxCells, s->maxNeighbours);
    s->Height = new DoubleArray2D(BoxX * refinementGridHeight, BoxY * refinementGridHeight);
    s->Normal = new CoordArray2D(BoxX * refinementGridHeight, BoxY * refinementGridHeight);
    s->Wall = new DoubleArray2D(BoxX, BoxY);
    s->Env = new EnvArray3D(BoxX, BoxY, BoxZ);
    s->Height->Initialize(cellRadius);
    s->Normal->Initialize(DoubleCoord(0, 0, -1));
    s->Wall->Initialize(0.0);
    s->dividing = new int[p->maxCells];
    s->IDlist = new int[s->maxNeighbours + 2];
    return s;
}

void ref_sim_destroy(void* h) {
    RefSim* s = (RefSim*)h;
    delete[] s->old_cells;
    delete[] s->new_cells;
    for (int i = 0; i < s->prm.maxCells; i++) delete[] s->NL[i];
    delete[] s->NL;
    delete s->grid;
    delete s->Height;
    delete s->Normal;
    delete s->Wall;
    delete s->Env;
    delete[] s->dividing;
    delete[] s->IDlist;
    delete s;
}

// As LoadCells (ABM/InputOutput.cpp:422-426): grid insertion in ID order.
void ref_sim_set_cells(void* h, const bcg_cell* cells, int n) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    memcpy((void*)s->old_cells, cells, sizeof(Cell) * n);
    s->N = n;
    for (int i = 0; i < n; i++) s->grid->Add(i, s->grid->GetAddress(average(s->old_cells[i].Position)));
}

void ref_sim_set_fields(void* h, const double* height, const double* normal, const double* wall,
                        const double* gr1, const double* gr2) {
    RefSim* s = (RefSim*)h;
    const int nh = BoxX * refinementGridHeight * BoxY * refinementGridHeight;
    if (height) s->Height->SetData((double*)height, nh);
    if (normal) {
        int nx = BoxX * refinementGridHeight, ny = BoxY * refinementGridHeight;
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++) {
                const double* v = normal + 3 * (x * ny + y);
                s->Normal->At(x, y) = DoubleCoord(v[0], v[1], v[2]);
            }
    }
    if (wall) s->Wall->SetData((double*)wall, BoxX * BoxY);
    for (int x = 0; x < BoxX; x++)
        for (int y = 0; y < BoxY; y++)
            for (int z = 0; z < BoxZ; z++) {
                int i = (x * BoxY + y) * BoxZ + z;
                if (gr1) s->Env->At(x, y, z).GrowthRate1 = gr1[i];
                if (gr2) s->Env->At(x, y, z).GrowthRate2 = gr2[i];
            }
}

int ref_sim_num_cells(void* h) { return ((RefSim*)h)->N; }

void ref_sim_get_cells(void* h, bcg_cell* out) {
    RefSim* s = (RefSim*)h;
    memcpy(out, (void*)s->old_cells, sizeof(Cell) * s->N);
}

void ref_sim_rebuild(void* h) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    getNeighbours(s->old_cells, s->N, *s->grid, s->NL, s->maxNeighbours);
}

void ref_sim_rebuild_ids(void* h, const int* ids, int n) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    getNeighbours(s->old_cells, *s->grid, s->NL, s->maxNeighbours, (int*)ids, n);
}

// counts[N], lists[N*stride] in the reference's list order
void ref_sim_get_neighbours(void* h, int* counts, int* lists, int stride) {
    RefSim* s = (RefSim*)h;
    for (int i = 0; i < s->N; i++) {
        int c = s->NL[i][0];
        counts[i] = c;
        for (int k = 0; k < c && k < stride; k++) lists[i * stride + k] = s->NL[i][k + 1];
    }
}

// box contents in the reference's internal order: counts[nbox], ids[nbox*stride]
void ref_sim_get_boxes(void* h, int* counts, int* ids, int stride) {
    RefSim* s = (RefSim*)h;
    int b = 0;
    for (int x = 0; x < BoxX; x++)
        for (int y = 0; y < BoxY; y++)
            for (int z = 0; z < BoxZ; z++, b++) {
                int* cells;
                int c = s->grid->GetCells(x, y, z, cells);
                counts[b] = c;
                for (int k = 0; k < c && k < stride; k++) ids[b * stride + k] = cells[k];
            }
}

void ref_sim_sum_forces(void* h, double* F, double* T) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    for (int i = 0; i < s->N; i++) {
        DoubleCoord f, t, sf, df;
        IntCoord xy = s->grid->GetXY(s->grid->GetAddress(average(s->old_cells[i].Position)));
        sum_forces(s->old_cells[i], s->old_cells, s->NL[i], f, t, *s->Height, *s->Normal, *s->grid, xy,
                   *s->Wall, sf, df);
        F[3 * i] = f.x; F[3 * i + 1] = f.y; F[3 * i + 2] = f.z;
        T[3 * i] = t.x; T[3 * i + 1] = t.y; T[3 * i + 2] = t.z;
    }
}

// mean_stress (ABM/Compute.cpp:18-117) of every cell, as the output step
// calls it (ABM/Simulation.cpp:320-326).
void ref_sim_mean_stress(void* h, double* S5, double* F3) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    for (int i = 0; i < s->N; i++) {
        Tensor st;
        DoubleCoord fn;
        mean_stress(s->old_cells[i], s->old_cells, s->NL[i], *s->grid, *s->Wall, *s->Height, *s->Normal, st, fn);
        S5[5 * i] = st.xx; S5[5 * i + 1] = st.yy; S5[5 * i + 2] = st.zz; S5[5 * i + 3] = st.rr; S5[5 * i + 4] = st.tt;
        F3[3 * i] = fn.x; F3[3 * i + 1] = fn.y; F3[3 * i + 2] = fn.z;
    }
}

// The output step's cell file: mean_stress and Output of every cell
// (ABM/Simulation.cpp:320-326) into `path`.
int ref_sim_output_cells(void* h, const char* path, double t) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    for (int i = 0; i < s->N; i++) {
        Tensor st;
        DoubleCoord fn;
        mean_stress(s->old_cells[i], s->old_cells, s->NL[i], *s->grid, *s->Wall, *s->Height, *s->Normal, st, fn);
        Output(f, i, t, s->old_cells[i], st);
    }
    fclose(f);
    return 0;
}

// One step, ABM/Simulation.cpp:261-413 without the field update and output.
// Returns the number of divisions; lineage gets (parent, daughter) pairs.
int ref_sim_step(void* h, double dt, double t, int rebuild, int* lineage, int lineage_cap) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    int N_cells = s->N;
    if (rebuild) getNeighbours(s->old_cells, N_cells, *s->grid, s->NL, s->maxNeighbours);
    int numDivide = 0;
    for (int cellID = 0; cellID < N_cells; cellID++) {
        MoveCell(cellID, *s->grid, s->old_cells, s->new_cells, s->NL[cellID], dt, *s->Height, *s->Normal,
                 *s->Wall);
        GrowCell(s->new_cells[cellID], cellID, dt, s->dividing, numDivide, *s->Env, NULL, *s->grid);
    }
    int done = 0;
    for (int cellCount = 0; cellCount < numDivide; cellCount++) {
        if (N_cells > (maxCells - 1)) break;
        const int cellID = s->dividing[cellCount];
        DivideCell(cellID, N_cells, s->new_cells, *s->grid, s->NL[cellID], *s->Wall, *s->Height, *s->Normal, t);
        if (done < lineage_cap) {
            lineage[2 * done] = cellID;
            lineage[2 * done + 1] = N_cells;
        }
        done++;
        N_cells++;
        int IDlen = s->NL[cellID][0];
        memcpy(s->IDlist, s->NL[cellID], (IDlen + 1) * sizeof(int));
        s->IDlist[0] = cellID;
        s->IDlist[IDlen + 1] = N_cells - 1;
        IDlen += 2;
        getNeighbours(s->new_cells, *s->grid, s->NL, s->maxNeighbours, s->IDlist, IDlen);
    }
    Cell* tmp = s->new_cells;
    s->new_cells = s->old_cells;
    s->old_cells = tmp;
    s->N = N_cells;
    return done;
}

// ref_sim_step that also copies the moved and grown cells before the
// divisions (N of them) to `moved`: the array the output step saves as its
// Restart state (SaveCells(new_cells, ...), ABM/Simulation.cpp:349).
int ref_sim_step_split(void* h, double dt, double t, int rebuild, bcg_cell* moved, int* lineage, int lineage_cap) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    int N_cells = s->N;
    if (rebuild) getNeighbours(s->old_cells, N_cells, *s->grid, s->NL, s->maxNeighbours);
    int numDivide = 0;
    for (int cellID = 0; cellID < N_cells; cellID++) {
        MoveCell(cellID, *s->grid, s->old_cells, s->new_cells, s->NL[cellID], dt, *s->Height, *s->Normal,
                 *s->Wall);
        GrowCell(s->new_cells[cellID], cellID, dt, s->dividing, numDivide, *s->Env, NULL, *s->grid);
    }
    memcpy((void*)moved, (const void*)s->new_cells, sizeof(Cell) * N_cells);
    int done = 0;
    for (int cellCount = 0; cellCount < numDivide; cellCount++) {
        if (N_cells > (maxCells - 1)) break;
        const int cellID = s->dividing[cellCount];
        DivideCell(cellID, N_cells, s->new_cells, *s->grid, s->NL[cellID], *s->Wall, *s->Height, *s->Normal, t);
        if (done < lineage_cap) {
            lineage[2 * done] = cellID;
            lineage[2 * done + 1] = N_cells;
        }
        done++;
        N_cells++;
        int IDlen = s->NL[cellID][0];
        memcpy(s->IDlist, s->NL[cellID], (IDlen + 1) * sizeof(int));
        s->IDlist[0] = cellID;
        s->IDlist[IDlen + 1] = N_cells - 1;
        IDlen += 2;
        getNeighbours(s->new_cells, *s->grid, s->NL, s->maxNeighbours, s->IDlist, IDlen);
    }
    Cell* tmp = s->new_cells;
    s->new_cells = s->old_cells;
    s->old_cells = tmp;
    s->N = N_cells;
    return done;
}

// The same step with the move/grow loop on OpenMP threads, as the
// reference's time loop runs it (ABM/Simulation.cpp:330-335: a static
// parallel for over MoveCell + GrowCell, whose bin moves and dividing-list
// appends take critical sections, ABM/Cells.cpp:34,50). The dividing order,
// and with it the division loop, then depends on thread timing: for timing
// only.
int ref_sim_step_omp(void* h, double dt, double t, int rebuild, int* lineage, int lineage_cap) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    omp_set_num_threads(g_threads);
    int N_cells = s->N;
    if (rebuild) getNeighbours(s->old_cells, N_cells, *s->grid, s->NL, s->maxNeighbours);
    int numDivide = 0;
#pragma omp parallel for schedule(static)
    for (int cellID = 0; cellID < N_cells; cellID++) {
        MoveCell(cellID, *s->grid, s->old_cells, s->new_cells, s->NL[cellID], dt, *s->Height, *s->Normal,
                 *s->Wall);
        GrowCell(s->new_cells[cellID], cellID, dt, s->dividing, numDivide, *s->Env, NULL, *s->grid);
    }
    int done = 0;
    for (int cellCount = 0; cellCount < numDivide; cellCount++) {
        if (N_cells > (maxCells - 1)) break;
        const int cellID = s->dividing[cellCount];
        DivideCell(cellID, N_cells, s->new_cells, *s->grid, s->NL[cellID], *s->Wall, *s->Height, *s->Normal, t);
        if (done < lineage_cap) {
            lineage[2 * done] = cellID;
            lineage[2 * done + 1] = N_cells;
        }
        done++;
        N_cells++;
        int IDlen = s->NL[cellID][0];
        memcpy(s->IDlist, s->NL[cellID], (IDlen + 1) * sizeof(int));
        s->IDlist[0] = cellID;
        s->IDlist[IDlen + 1] = N_cells - 1;
        IDlen += 2;
        getNeighbours(s->new_cells, *s->grid, s->NL, s->maxNeighbours, s->IDlist, IDlen);
    }
    Cell* tmp = s->new_cells;
    s->new_cells = s->old_cells;
    s->old_cells = tmp;
    s->N = N_cells;
    return done;
}

// ---------------------------------------------------------------- environment
// The field and nutrient passes of the reference time loop
// (ABM/Simulation.cpp:187-211, 269-312) on a RefSim's cells, grid, Height,
// Normal and Environment. Arrays the reference leaves uninitialised are
// pinned: BoxAverage outputs outside its window are zeroed after each call,
// and insideColonyDen gets one extra zero x-plane (the restriction reads
// x = BoxX at ABM/Multigrid.cpp:272).
static_assert(sizeof(LocalEnv) == sizeof(bcg_env_cell), "LocalEnv layout");
static_assert(sizeof(LocalAga) == sizeof(bcg_aga), "LocalAga layout");

static void set_env_globals(const bcg_env_params* p) {
    BoxX = p->BoxX;
    BoxY = p->BoxY;
    BoxZ = p->BoxZ;
    BoxZAgar = p->BoxZAgar;
    maxLevelsMG = p->maxLevelsMG;
    refinementGridHeight = p->refinementGridHeight;
    FilterLen = p->FilterLen;
    crossFeeding = p->crossFeeding != 0;
    NutrientGSI = p->NutrientGSI != 0;
    InterfaceCondition = p->InterfaceCondition;
    BoxLength = p->BoxLength;
    cellRadius = p->cellRadius;
    density_threshold = p->density_threshold;
    UpdateTime = p->UpdateTime;
    Cdt = p->Cdt;
    KC_aer = p->KC_aer; KC_ana = p->KC_ana; KW = p->KW; KO = p->KO; KA = p->KA; KI = p->KI;
    C_rate = p->C_rate; W_rate = p->W_rate; Carbon_to_Waste = p->Carbon_to_Waste; Maintenance_rate = p->Maintenance_rate;
    maxGrowthRate1 = p->maxGrowthRate1; maxGrowthRate2 = p->maxGrowthRate2;
    DiffColony = p->DiffColony; DiffAgar = p->DiffAgar; maxCarbon = p->maxCarbon; maxO2 = p->maxO2;
    maxO2_Agar = p->maxO2_Agar;
    qCarAer = p->qCarAer; qCarAna = p->qCarAna; qAceAer = p->qAceAer; qO2Car = p->qO2Car; qO2Ace = p->qO2Ace;
    pAceAer = p->pAceAer; pAceAna = p->pAceAna;
    DiffColonyCarbon = p->DiffColonyCarbon; DiffAgarCarbon = p->DiffAgarCarbon; DiffColonyO2 = p->DiffColonyO2;
    DiffAgarO2 = p->DiffAgarO2; DiffColonyAce = p->DiffColonyAce; DiffAgarAce = p->DiffAgarAce;
    maxGrowthRateCarAer = p->maxGrowthRateCarAer; maxGrowthRateCarAna = p->maxGrowthRateCarAna;
    maxGrowthRateAceAer = p->maxGrowthRateAceAer;
    Kw = p->Kw; Ka = p->Ka; Kb = p->Kb; Cb = p->Cb; c1 = p->c1;
    qCarAernt = p->qCarAernt; qCarAnant = p->qCarAnant; qAceAernt = p->qAceAernt; qO2Carnt = p->qO2Carnt;
    qO2Acent = p->qO2Acent; pAceAnant = p->pAceAnant;
}

struct RefEnv {
    bcg_env_params prm;
    RefSim* sim;
    EnvArray3D* oldEnv;
    AgaArray3D *agar[8], *agar_old[8];
    AgaArray2D *wall[8], *wall_old[8];
    DoubleArray3D *R, *R1, *R2, *RS, *RS1, *RS2, *D, *D1, *D2, *DS, *DS1, *DS2, *inside;
    DoubleArray2D *RW, *RW1, *RW2, *RWS, *RWS1, *RWS2, *W, *W1, *W2, *WS, *WS1, *WS2, *RH;
    DoubleArray2D* Filter;
    DoubleArray3D* Filter3D;
    Multigrid* mg;
    int bounds[5];
};

static void env_globals(RefEnv* e) {
    set_globals(&e->sim->prm);
    set_env_globals(&e->prm);
}

void* ref_env_create(const bcg_env_params* p, void* sim) {
    RefEnv* e = new RefEnv();
    e->prm = *p;
    e->sim = (RefSim*)sim;
    env_globals(e);
    e->oldEnv = new EnvArray3D(BoxX, BoxY, BoxZ);
    for (int l = 0; l < maxLevelsMG; l++) {
        e->agar[l] = new AgaArray3D(BoxX, BoxY, BoxZAgar);
        e->agar_old[l] = new AgaArray3D(BoxX, BoxY, BoxZAgar);
        e->wall[l] = new AgaArray2D(BoxX, BoxY);
        e->wall_old[l] = new AgaArray2D(BoxX, BoxY);
    }
    // the simulation's Environment starts as constructed (LocalEnv())
    for (int x = 0; x < BoxX; x++)
        for (int z = 0; z < BoxZ; z++) e->sim->Env->At(x, 0, z) = LocalEnv();
    DoubleArray3D** d3[] = {&e->R, &e->R1, &e->R2, &e->RS, &e->RS1, &e->RS2, &e->D, &e->D1, &e->D2,
                            &e->DS, &e->DS1, &e->DS2};
    for (auto q : d3) {
        *q = new DoubleArray3D(BoxX, BoxY, BoxZ);
        (*q)->Initialize(0.0);
    }
    e->inside = new DoubleArray3D(BoxX + 1, BoxY, BoxZ);
    e->inside->Initialize(0.0);
    DoubleArray2D** d2[] = {&e->RW, &e->RW1, &e->RW2, &e->RWS, &e->RWS1, &e->RWS2, &e->W, &e->W1, &e->W2,
                            &e->WS, &e->WS1, &e->WS2};
    for (auto q : d2) {
        *q = new DoubleArray2D(BoxX, BoxY);
        (*q)->Initialize(0.0);
    }
    e->RH = new DoubleArray2D(BoxX * refinementGridHeight, BoxY * refinementGridHeight);
    e->RH->Initialize(cellRadius);
    // filters exactly as ABM/Simulation.cpp:60-110 builds them
    e->Filter = new DoubleArray2D(FilterLen, 1);
    double value = 1.0 / (double)(FilterLen);
    for (int ii = 0; ii < FilterLen; ii++) e->Filter->Set(ii, 0, value);
    e->Filter3D = new DoubleArray3D(FilterLen, FilterLen, FilterLen);
    double norm = 0;
    for (int ii = 0; ii < FilterLen; ii++)
        for (int kk = 0; kk < FilterLen; kk++) {
            value = exp(-(ii - FilterLen / 2) * (ii - FilterLen / 2) - (kk - FilterLen / 2) * (kk - FilterLen / 2));
            norm = norm + value;
        }
    for (int ii = 0; ii < FilterLen; ii++)
        for (int kk = 0; kk < FilterLen; kk++) {
            value = exp(-(ii - FilterLen / 2) * (ii - FilterLen / 2) - (kk - FilterLen / 2) * (kk - FilterLen / 2));
            e->Filter3D->Set(ii, 0, kk, value / norm);
        }
    e->mg = new Multigrid(maxLevelsMG, 8, BoxX, BoxY, BoxZAgar);
    e->mg->setInsideColony(e->inside);
    memset(e->bounds, 0, sizeof e->bounds);
    return e;
}

// The output step's density and environment files (ABM/Simulation.cpp:198,
// 221): the reference's own writers.
int ref_env_output(void* h, const char* density_path, const char* env_path) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    FILE* f = fopen(density_path, "w");
    if (!f) return -1;
    e->DS->Output(f, 1);
    fclose(f);
    f = fopen(env_path, "w");
    if (!f) return -1;
    e->sim->Env->Output(f, 1);
    fclose(f);
    return 0;
}

void ref_env_destroy(void* h) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    delete e->mg;
    delete e->oldEnv;
    for (int l = 0; l < e->prm.maxLevelsMG; l++) {
        delete e->agar[l];
        delete e->agar_old[l];
        delete e->wall[l];
        delete e->wall_old[l];
    }
    for (auto q : {e->R, e->R1, e->R2, e->RS, e->RS1, e->RS2, e->D, e->D1, e->D2, e->DS, e->DS1, e->DS2, e->inside})
        delete q;
    for (auto q : {e->RW, e->RW1, e->RW2, e->RWS, e->RWS1, e->RWS2, e->W, e->W1, e->W2, e->WS, e->WS1, e->WS2, e->RH})
        delete q;
    delete e->Filter;
    delete e->Filter3D;
    delete e;
}

// BoxAverage, then its outputs outside the window it computes set to 0.
static void box_average(RefEnv* e, DoubleArray3D& rough, DoubleArray3D& out, DoubleArray2D& rwall, DoubleArray2D& wout,
                        int minx, int maxx, int miny, int maxy, int maxz) {
    BoxAverage(rough, out, rwall, wout, *e->Filter3D, int((e->Filter3D->m_Size.x) / 2), minx, maxx, miny, maxy, maxz,
               BoxX, BoxY, BoxZ);
    for (int x = 0; x < BoxX; x++) {
        const bool xin = x >= minx && x < maxx;
        if (!xin) wout.Set(x, 0, 0.0);
        for (int z = 0; z < BoxZ; z++)
            if (!xin || z + 1 > maxz) out.Set(x, 0, z, 0.0);
    }
}

// pass 0 INIT, 1 UPDATE, 2 HEIGHT (see bcg_env_fields)
void ref_env_fields(void* h, int pass) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    RefSim* s = e->sim;
    int minx, maxx, miny, maxy, maxz;
    const int FilterDim = FilterLen / 2;
    if (pass == 2) {
        GetHeight(*e->R, *s->Height, *s->grid, s->old_cells, minx, maxx, miny, maxy, maxz);
        Height_Average(*s->Height, *e->RH, minx, maxx, miny, maxy);
        Smooth(*e->RH, *s->Height, *e->Filter, FilterDim, minx, maxx, miny, maxy);
        GetSurfaceNormal(*s->Normal, *s->Height, minx, maxx, miny, maxy);
    } else {
        GetDensity(*e->R, *e->R1, *e->R2, *e->inside, *s->Height, *s->grid, s->old_cells, minx, maxx, miny, maxy, maxz);
        ShiftDensity(*e->R, *e->R1, *e->R2, *e->RW, *e->RW1, *e->RW2, *e->RS, *e->RS1, *e->RS2, *e->RWS, *e->RWS1,
                     *e->RWS2, BoxX, BoxY, BoxZ);
        box_average(e, *e->R, *e->D, *e->RW, *e->W, minx, maxx, miny, maxy, maxz);
        box_average(e, *e->R1, *e->D1, *e->RW1, *e->W1, minx, maxx, miny, maxy, maxz);
        box_average(e, *e->R2, *e->D2, *e->RW2, *e->W2, minx, maxx, miny, maxy, maxz);
        box_average(e, *e->RS, *e->DS, *e->RWS, *e->WS, minx, maxx, miny, maxy, maxz);
        box_average(e, *e->RS1, *e->DS1, *e->RWS1, *e->WS1, minx, maxx, miny, maxy, maxz);
        box_average(e, *e->RS2, *e->DS2, *e->RWS2, *e->WS2, minx, maxx, miny, maxy, maxz);
        if (pass == 1) {
            Height_Average(*s->Height, *e->RH, minx, maxx, miny, maxy);
            Smooth(*e->RH, *s->Height, *e->Filter, FilterDim, minx, maxx, miny, maxy);
        }
        GetSurfaceNormal(*s->Normal, *s->Height, minx, maxx, miny, maxy);
    }
    e->bounds[0] = minx;
    e->bounds[1] = maxx;
    e->bounds[2] = miny;
    e->bounds[3] = maxy;
    e->bounds[4] = maxz;
}

int ref_env_nutrients(void* h) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    omp_set_num_threads(g_threads);
    RefSim* s = e->sim;
    return UpdateEnvArrayMG_OxygenAcetate(s->Env, e->oldEnv, e->agar, e->agar_old, e->wall, e->wall_old, *e->DS,
                                          *e->DS1, *e->DS2, *e->WS, *e->WS1, *e->WS2, e->bounds[0], e->bounds[1],
                                          e->bounds[2], e->bounds[3], e->bounds[4], *s->Height, *e->inside, *e->mg);
}

void ref_env_get_fields(void* h, double* density, double* dshift, double* inside, double* wshift, double* height,
                        double* normal, int* bounds) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    for (int x = 0; x < BoxX; x++) {
        for (int z = 0; z < BoxZ; z++) {
            const int b = x * BoxZ + z;
            if (density) density[b] = e->D->Get(x, 0, z);
            if (dshift) dshift[b] = e->DS->Get(x, 0, z);
            if (inside) inside[b] = e->inside->Get(x, 0, z);
        }
        if (wshift) wshift[x] = e->WS->Get(x, 0);
    }
    const int nh = BoxX * refinementGridHeight, ny = BoxY * refinementGridHeight;
    for (int x = 0; x < nh; x++)
        for (int y = 0; y < ny; y++) {
            if (height) height[x * ny + y] = e->sim->Height->Get(x, y);
            if (normal) {
                DoubleCoord v = e->sim->Normal->Get(x, y);
                normal[3 * (x * ny + y)] = v.x;
                normal[3 * (x * ny + y) + 1] = v.y;
                normal[3 * (x * ny + y) + 2] = v.z;
            }
        }
    if (bounds) memcpy(bounds, e->bounds, sizeof e->bounds);
}

void ref_env_get_state(void* h, bcg_env_cell* colony, bcg_env_cell* colony_old, bcg_aga* agar, bcg_aga* agar_old,
                       bcg_aga* wall, bcg_aga* wall_old) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    for (int x = 0; x < BoxX; x++) {
        for (int z = 0; z < BoxZ; z++) {
            if (colony) memcpy(&colony[x * BoxZ + z], &e->sim->Env->At(x, 0, z), sizeof(bcg_env_cell));
            if (colony_old) memcpy(&colony_old[x * BoxZ + z], &e->oldEnv->At(x, 0, z), sizeof(bcg_env_cell));
        }
        for (int l = 0; l < maxLevelsMG; l++) {
            const size_t na = (size_t)BoxX * BoxZAgar;
            for (int z = 0; z < BoxZAgar; z++) {
                if (agar) memcpy(&agar[l * na + x * BoxZAgar + z], &e->agar[l]->At(x, 0, z), sizeof(bcg_aga));
                if (agar_old) memcpy(&agar_old[l * na + x * BoxZAgar + z], &e->agar_old[l]->At(x, 0, z), sizeof(bcg_aga));
            }
            if (wall) memcpy(&wall[l * BoxX + x], &e->wall[l]->At(x, 0), sizeof(bcg_aga));
            if (wall_old) memcpy(&wall_old[l * BoxX + x], &e->wall_old[l]->At(x, 0), sizeof(bcg_aga));
        }
    }
}

// ---------------------------------------------------------------- file I/O
// The reference's own writers and readers of its output files, on arrays in
// the library's layouts (agar of level l at l * BX * BY * BZA, x-major; wall
// of level l at l * BX * BY). The arrays are built with the current env
// globals (LocalAga() reads maxCarbon and maxO2), then overwritten.

// FieldAgar[level]->Append(f, n) for every level (ABM/Array.h:1447-1494)
int ref_write_agar(const char* path, const bcg_aga* data, int BX, int BY, int BZA, int levels, int n) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    const size_t na = (size_t)BX * BY * BZA;
    for (int l = 0; l < levels; l++) {
        AgaArray3D a(BX, BY, BZA);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++)
                for (int z = 0; z < BZA; z++)
                    memcpy(&a.At(x, y, z), &data[l * na + ((size_t)x * BY + y) * BZA + z], sizeof(LocalAga));
        a.Append(f, n);
    }
    fclose(f);
    return 0;
}

// FieldWall[level]->Append(f) for every level (ABM/Array.h:419-457)
int ref_write_wall(const char* path, const bcg_aga* data, int BX, int BY, int levels) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    for (int l = 0; l < levels; l++) {
        AgaArray2D w(BX, BY);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++) memcpy(&w.At(x, y), &data[(size_t)l * BX * BY + x * BY + y], sizeof(LocalAga));
        w.Append(f);
    }
    fclose(f);
    return 0;
}

// Two Input calls on one stream, as the field restart makes them
// (ABM/CellMD3D.cpp:105-125): a then b are read from `path`; both hold the
// values to keep where the file gives none.
int ref_read_agar_pair(const char* path, bcg_aga* a, bcg_aga* b, int BX, int BY, int BZA) {
    std::ifstream in(path);
    if (!in) return -1;
    bcg_aga* io[2] = {a, b};
    for (int k = 0; k < 2; k++) {
        AgaArray3D arr(BX, BY, BZA);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++)
                for (int z = 0; z < BZA; z++)
                    memcpy(&arr.At(x, y, z), &io[k][((size_t)x * BY + y) * BZA + z], sizeof(LocalAga));
        arr.Input(in);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++)
                for (int z = 0; z < BZA; z++)
                    memcpy(&io[k][((size_t)x * BY + y) * BZA + z], &arr.At(x, y, z), sizeof(LocalAga));
    }
    return 0;
}

int ref_read_wall_pair(const char* path, bcg_aga* a, bcg_aga* b, int BX, int BY) {
    std::ifstream in(path);
    if (!in) return -1;
    bcg_aga* io[2] = {a, b};
    for (int k = 0; k < 2; k++) {
        AgaArray2D arr(BX, BY);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++) memcpy(&arr.At(x, y), &io[k][x * BY + y], sizeof(LocalAga));
        arr.Input(in);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++) memcpy(&io[k][x * BY + y], &arr.At(x, y), sizeof(LocalAga));
    }
    return 0;
}

int ref_read_env_pair(const char* path, bcg_env_cell* a, bcg_env_cell* b, int BX, int BY, int BZ) {
    std::ifstream in(path);
    if (!in) return -1;
    bcg_env_cell* io[2] = {a, b};
    for (int k = 0; k < 2; k++) {
        EnvArray3D arr(BX, BY, BZ);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++)
                for (int z = 0; z < BZ; z++)
                    memcpy(&arr.At(x, y, z), &io[k][((size_t)x * BY + y) * BZ + z], sizeof(LocalEnv));
        arr.Input(in);
        for (int x = 0; x < BX; x++)
            for (int y = 0; y < BY; y++)
                for (int z = 0; z < BZ; z++)
                    memcpy(&io[k][((size_t)x * BY + y) * BZ + z], &arr.At(x, y, z), sizeof(LocalEnv));
    }
    return 0;
}

// SaveCells into a new file (ABM/InputOutput.cpp:434-448)
int ref_save_cells(const char* path, const bcg_cell* cells, int n, double t, double dt) {
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    SaveCells(f, (Cell*)cells, n, t, dt);
    fclose(f);
    return 0;
}

// LoadCells into the simulation (ABM/InputOutput.cpp:394-431): cells, their
// grid insertion in ID order, t and dt. -1 if the file holds more cells than
// the simulation's capacity (the reference does not check).
int ref_sim_load_cells(void* h, const char* path, double* t, double* dt) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long size = ftell(f);
    fclose(f);
    if (size < 16 || (size - 16) / (long)sizeof(Cell) > s->prm.maxCells) return -1;
    s->N = LoadCells((char*)path, s->old_cells, *s->grid, *t, *dt);
    return s->N;
}

// AddFirstCells (ABM/InputOutput.cpp:294-390) for one colony of `size`
// cells into the simulation; size 1 only (a larger colony never places its
// second cell and the reference exits).
int ref_sim_add_first_cells(void* h, int size, double radius) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    if (size != 1) return -1;
    Inputs ini;
    ini.ColonyNumber = 1;
    ini.ColonySeparation = 0;
    ini.ColonyRadius = radius;
    ini.ColonySize = size;
    ini.Fraction = 0;
    s->N = AddFirstCells(s->old_cells, L_divide, cellRadius, *s->grid, ini);
    return s->N;
}

// The rough wall's expression (ABM/Simulation.cpp:160-168), compiled with the
// reference's flags: BoxX * BoxY draws of this library's rand().
void ref_rough_wall(const bcg_params* p, double* out) {
    set_globals(p);
    for (int ii = 0; ii < BoxX; ii++)
        for (int jj = 0; jj < BoxY; jj++) out[ii * BoxY + jj] = ((float)rand() / RAND_MAX - 0.5) * wall_rough;
}

// The field part of a restart (ABM/CellMD3D.cpp:105-125) into the
// environment's arrays: Environment and oldEnvironment from one stream of
// <dir>/Environment/<index>.txt; every level reopens AgarField and WallField.
int ref_env_field_restart(void* h, const char* dir, int index) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    char env_name[600], agar_name[600], wal_name[600];
    snprintf(env_name, sizeof env_name, "%s/Environment/%d.txt", dir, index);
    snprintf(agar_name, sizeof agar_name, "%s/AgarField/%d.txt", dir, index);
    snprintf(wal_name, sizeof wal_name, "%s/WallField/%d.txt", dir, index);
    std::ifstream infile2(env_name);
    if (!infile2) return -1;
    e->sim->Env->Input(infile2);
    e->oldEnv->Input(infile2);
    for (int i = 0; i < maxLevelsMG; i++) {
        std::ifstream infile(agar_name);
        std::ifstream infile3(wal_name);
        if (!infile || !infile3) return -1;
        e->agar[i]->Input(infile);
        e->agar_old[i]->Input(infile);
        e->wall[i]->Input(infile3);
        e->wall_old[i]->Input(infile3);
    }
    return 0;
}

// The AgarField and WallField files of the environment's state
// (ABM/Simulation.cpp:222-226): every level appended to each.
int ref_env_append(void* h, const char* agar_path, const char* wall_path) {
    RefEnv* e = (RefEnv*)h;
    env_globals(e);
    FILE* fa = fopen(agar_path, "w");
    FILE* fw = fopen(wall_path, "w");
    if (!fa || !fw) return -1;
    for (int level = 0; level < maxLevelsMG; level++) {
        e->agar[level]->Append(fa, 1);
        e->wall[level]->Append(fw);
    }
    fclose(fa);
    fclose(fw);
    return 0;
}

// The rough wall of the run (ABM/Simulation.cpp:160-168) into the simulation.
void ref_sim_rough_wall(void* h) {
    RefSim* s = (RefSim*)h;
    set_globals(&s->prm);
    for (int ii = 0; ii < BoxX; ii++)
        for (int jj = 0; jj < BoxY; jj++) s->Wall->Set(ii, jj, ((float)rand() / RAND_MAX - 0.5) * wall_rough);
}

}  // extern "C"
