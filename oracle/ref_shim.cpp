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

extern "C" {

void ref_set_params(const bcg_params* p) { set_globals(p); omp_set_num_threads(1); }
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
    omp_set_num_threads(1);
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

}  // extern "C"
