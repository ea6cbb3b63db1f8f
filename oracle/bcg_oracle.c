/*
 * bcg_oracle.c - CPU restatement of the BacColonyGrowth mechanics hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against and
 * the CPU "port" baseline timed beside it. The product never links this file.
 *
 * Every function restates one reference function; the file:line it follows
 * (relative to AgentBasedModel-BacColonyGrowth/) is given above it. The
 * arithmetic is written in the reference's evaluation order so that, with FMA
 * contraction off, every result rounds exactly as the reference's does.
 * planar == 0 selects the 3-D generalisation (no reference counterpart):
 * y is not forced to zero, the wall contact point keeps the vertex y, and
 * the surface height/normal are looked up bilinearly in x and y.
 */
#include "bcg_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static const double PI_ = 3.141592653589793238462643383279502884197; /* ABM/tools.h:12 */

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vscale(v3 a, double c) { return mk(a.x * c, a.y * c, a.z * c); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* ABM/tools.h:121-124 (component y written as -a.x*b.z + a.z*b.x) */
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
static inline double dmax(double a, double b) { return a > b ? a : b; } /* ABM/tools.h:94 */
static inline double dmin(double a, double b) { return a < b ? a : b; } /* ABM/tools.h:104 */
static inline v3 P(const bcg_cell* c) { return mk(c->p[0], c->p[1], c->p[2]); }
static inline v3 Q(const bcg_cell* c) { return mk(c->q[0], c->q[1], c->q[2]); }
static inline v3 V(const double* a) { return mk(a[0], a[1], a[2]); }
static inline void st(double* a, v3 v) { a[0] = v.x; a[1] = v.y; a[2] = v.z; }
/* average(Segment): ABM/tools.h:149-152 */
static inline v3 cmof(const bcg_cell* c) { return vscale(vadd(P(c), Q(c)), 0.5); }

/* ---------------------------------------------------------------- RNG --- */
/* glibc random_r TYPE_3: r[0]=seed, r[i]=16807*r[i-1] mod (2^31-1) for
 * i<31, r[i]=r[i-31] for i=31..33, then r[i]=r[i-31]+r[i-3] (mod 2^32),
 * the first 310 outputs discarded, each draw returned as r>>1.            */
void or_glibc_srand(or_glibc_rng* g, uint32_t seed) {
    if (seed == 0) seed = 1;
    g->r[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        long long w = (16807LL * g->r[i - 1]) % 2147483647LL;
        if (w < 0) w += 2147483647LL;
        g->r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; i++) g->r[i] = g->r[i - 31];
    g->i = 0; /* ring slot of r[k] for the next k = 34, 35, ... is k % 34 */
    for (int k = 34; k < 344; k++) (void)or_glibc_rand(g);
}

int32_t or_glibc_rand(or_glibc_rng* g) {
    /* r[k] = r[k-31] + r[k-3]; with k % 34 == g->i, r[k-31] sits in slot
       (i+3)%34 and r[k-3] in slot (i+31)%34 */
    int i = g->i;
    uint32_t v = (uint32_t)g->r[(i + 3) % 34] + (uint32_t)g->r[(i + 31) % 34];
    g->r[i] = (int32_t)v;
    g->i = (i + 1) % 34;
    return (int32_t)(v >> 1);
}

/* The stream the oracle draws from: glibc's rand() restated (above), owned
 * by the oracle so that no other code in the process (runtimes that call
 * rand() themselves) can advance it. Starts as srand(1), as libc does.    */
static or_glibc_rng g_stream;
static int g_stream_seeded = 0;

void or_srand(uint32_t seed) {
    or_glibc_srand(&g_stream, seed);
    g_stream_seeded = 1;
}

int32_t or_rand(void) {
    if (!g_stream_seeded) or_srand(1);
    return or_glibc_rand(&g_stream);
}

static void rand_discard(int64_t n) {
    for (int64_t k = 0; k < n; k++) (void)or_rand();
}

/* A draw as the reference uses it: ((float)rand()/RAND_MAX - 0.5). */
static double draw_centered(void) { return (double)((float)or_rand() / RAND_MAX) - 0.5; }

/* ----------------------------------------------------- min_distance --- */
/* ABM/Neighbours.cpp:95-104 */
static double clampd(double n, double lo, double hi) { return n < lo ? lo : (n > hi ? hi : n); }

/* ABM/Neighbours.cpp:107-160 */
static void min_distance(const bcg_cell* a, const bcg_cell* b, double* d, v3* c1, v3* c2) {
    v3 p1 = P(a), q1 = Q(a), p2 = P(b), q2 = Q(b);
    v3 v1 = vsub(q1, p1), v2 = vsub(q2, p2), r = vsub(p1, p2);
    double l1 = vdot(v1, v1), l2 = vdot(v2, v2), f = vdot(v2, r), c = vdot(v1, r), bb = vdot(v1, v2);
    double denom = l1 * l2 - bb * bb, s, t;
    if (denom != 0) s = clampd((bb * f - c * l2) / denom, 0, 1);
    else s = 0;
    double tnom = bb * s + f;
    if (tnom < 0) { t = 0; s = clampd(-c / l1, 0, 1); }
    else if (tnom > l2) { t = 1; s = clampd((bb - c) / l1, 0, 1); }
    else t = tnom / l2;
    *c1 = vadd(p1, vscale(v1, s));
    *c2 = vadd(p2, vscale(v2, t));
    v3 cv = vsub(*c1, *c2);
    *d = sqrt(vdot(cv, cv));
}

void or_min_distance(const bcg_cell* a, const bcg_cell* b, double* o) {
    double d; v3 c1, c2;
    min_distance(a, b, &d, &c1, &c2);
    o[0] = d; o[1] = c1.x; o[2] = c1.y; o[3] = c1.z; o[4] = c2.x; o[5] = c2.y; o[6] = c2.z;
}

/* ----------------------------------------------------------- forces --- */
/* F_cc: ABM/Forces.cpp:13-84 */
static void F_cc(const bcg_params* pr, const bcg_cell* a, const bcg_cell* b, v3* F, v3* r, double* d) {
    v3 c1, c2;
    const double gamma_n = 100.0;
    double sigma = a->Radius + b->Radius;
    double L1 = a->Length + 4.0 / 3 * a->Radius;
    double M1 = L1 * a->Radius * a->Radius * PI_;
    double L2 = b->Length + 4.0 / 3 * b->Radius;
    double M2 = L2 * b->Radius * b->Radius * PI_;
    double Meff = M1 * M2 / (M1 + M2);
    min_distance(a, b, d, &c1, &c2);
    double delta = dmax(sigma - *d, 0.0);
    v3 vn = vscale(vsub(c2, c1), 1.0 / *d);
    double FE_mag = sqrt(a->Radius * b->Radius / (a->Radius + b->Radius)) * pr->k_cc * pow(delta, 3.0 / 2.0);
    v3 FE = vscale(vn, -FE_mag);
    *r = vscale(vadd(vscale(c1, b->Radius), vscale(c2, a->Radius)), 1 / sigma);
    if (FE_mag > DBL_EPSILON) {
        v3 rcm1 = vsub(*r, cmof(a)), rcm2 = vsub(*r, cmof(b));
        v3 va1 = vcross(V(a->AngularVelocity), rcm1), va2 = vcross(V(b->AngularVelocity), rcm2);
        v3 v1 = vadd(V(a->Velocity), va1), v2 = vadd(V(b->Velocity), va2);
        v3 dv = vsub(v2, v1);
        v3 dvn = vscale(vn, vdot(dv, vn));
        v3 FDn = vscale(dvn, gamma_n * Meff * delta);
        v3 dvt = vsub(dv, dvn);
        double v_mag = sqrt(vdot(dvt, dvt));
        v3 FDt = mk(0.0, 0.0, 0.0);
        if (v_mag > 3 * DBL_EPSILON)
            FDt = vscale(dvt, dmin(pr->gamma_t * Meff * sqrt(delta), pr->cell_mu * FE_mag / v_mag));
        *F = vadd(vadd(FE, FDt), FDn);
    } else {
        *F = FE;
    }
}

void or_F_cc(const bcg_params* p, const bcg_cell* a, const bcg_cell* b, double* o) {
    v3 F, r; double d;
    F_cc(p, a, b, &F, &r, &d);
    o[0] = F.x; o[1] = F.y; o[2] = F.z; o[3] = r.x; o[4] = r.y; o[5] = r.z; o[6] = d;
}

/* One vertex of F_cw: ABM/Forces.cpp:98-152 (first locus) and :157-201. */
static v3 F_cw_locus(const bcg_params* pr, const bcg_cell* c, v3 vert, double wall_x, double wall_z, double M,
                     v3* rr, v3* sf, v3* df) {
    const double gamma_n = 100.0;
    double sigma = c->Radius;
    v3 wall = mk(wall_x, pr->planar ? 0.0 : vert.y, wall_z);
    *rr = wall;
    double d = vert.z - wall.z;
    double delta = dmax(sigma - d, 0.0);
    double FE_mag = sqrt(c->Radius) * pr->k_wc * pow(delta, 3.0 / 2.0);
    v3 FE = mk(0.0, 0.0, FE_mag);
    if (FE_mag > DBL_EPSILON) {
        v3 rcm = vsub(*rr, cmof(c));
        v3 va = vcross(V(c->AngularVelocity), rcm);
        v3 v = vadd(V(c->Velocity), va);
        v3 dv = vscale(v, -1.0);
        v3 dvn = mk(0.0, 0.0, dv.z);
        v3 dvt = vsub(dv, dvn);
        double v_mag = sqrt(vdot(dvt, dvt));
        v3 FDt;
        if (v_mag > DBL_EPSILON) FDt = vscale(dvt, dmin(pr->gamma_t * M * sqrt(delta), pr->wall_mu * FE_mag / v_mag));
        else FDt = mk(0, 0, 0);
        v3 FDn = vscale(dvn, gamma_n * M * delta);
        *sf = vadd(*sf, vscale(dvt, pr->gamma_t * M * sqrt(delta)));
        *df = vadd(*df, vscale(dvt, pr->wall_mu * FE_mag / v_mag));
        return vadd(FE, vadd(FDt, FDn));
    }
    return mk(0.0, 0.0, 0.0);
}

/* F_cw: ABM/Forces.cpp:88-204. The first locus draws one rand() (:102).
 * With draw == 0 the draw is left to the caller (its value is multiplied by
 * var_pos, so only var_pos == 0 may skip it).                             */
static void F_cw_d(const bcg_params* pr, const bcg_cell* c, double Wall_z, v3* F1, v3* F2, v3* r1, v3* r2, v3* sf,
                   v3* df, int draw) {
    double L = c->Length + 4.0 / 3 * c->Radius;
    double M = L * c->Radius * c->Radius * PI_;
    double wx = c->p[0] + (draw ? draw_centered() : 0.0) * pr->var_pos;
    *F1 = F_cw_locus(pr, c, P(c), wx, Wall_z, M, r1, sf, df);
    *F2 = F_cw_locus(pr, c, Q(c), c->q[0], Wall_z, M, r2, sf, df);
}

static void F_cw(const bcg_params* pr, const bcg_cell* c, double Wall_z, v3* F1, v3* F2, v3* r1, v3* r2, v3* sf,
                 v3* df) {
    F_cw_d(pr, c, Wall_z, F1, F2, r1, r2, sf, df, 1);
}

void or_F_cw(const bcg_params* p, const bcg_cell* c, double wall_z, double* o) {
    v3 F1, F2, r1, r2, sf = mk(0, 0, 0), df = mk(0, 0, 0);
    F_cw(p, c, wall_z, &F1, &F2, &r1, &r2, &sf, &df);
    v3 v[6] = {F1, F2, r1, r2, sf, df};
    for (int i = 0; i < 6; i++) st(o + 3 * i, v[i]);
}

/* Array2D<double>::linear_interp (ABM/Array.h:281-288): weights in float.
 * Reads outside the array (the reference reads past the end when Ny == 1,
 * always with a zero weight) contribute nothing here.                     */
static double height_interp(const double* H, int nx, int ny, int x0, int y0, double dx, double dy) {
    float x = dx < 0 ? 1 + dx : dx;
    float y = dy < 0 ? 1 + dy : dy;
#define HH(i, j) (((i) >= 0 && (i) < nx && (j) >= 0 && (j) < ny) ? H[(i) * ny + (j)] : 0.0)
#define HP(i, j) (((i) * ny + (j)) >= 0 && ((i) * ny + (j)) < nx * ny ? H[(i) * ny + (j)] : 0.0)
    if (ny == 1) /* the reference's flat index x*Ny+y, so (x0, 1) is element x0+1 */
        return HP(x0, y0) * (1 - x) * (1 - y) + HP(x0 + 1, y0) * x * (1 - y) + HP(x0, y0 + 1) * (1 - x) * y +
               HP(x0 + 1, y0 + 1) * x * y;
    return HH(x0, y0) * (1 - x) * (1 - y) + HH(x0 + 1, y0) * x * (1 - y) + HH(x0, y0 + 1) * (1 - x) * y +
           HH(x0 + 1, y0 + 1) * x * y;
#undef HH
#undef HP
}

typedef struct {
    int nhx, nhy;          /* height / normal array size */
    const double* height;  /* [nhx][nhy]    */
    const double* normal;  /* [nhx][nhy][3] */
    const double* wall;    /* [BoxX][BoxY]  */
} fields_t;

/* Interpolated surface height at a point: ABM/Forces.cpp:222-233 (x only in
 * the planar model; the 3-D model applies the same rule in y).             */
static double surf_height(const bcg_params* pr, const fields_t* f, v3 pt) {
    const int ref = pr->refinementGridHeight;
    const double hx = pr->BoxLength / ref;
    int i = (int)(pt.x / hx + pr->BoxX / 2 * ref);
    double dx = pt.x / hx - (i - pr->BoxX / 2 * ref + 0.5);
    i = (int)ceil(dx) + i - 1;
    int j = 0;
    double dy = 0;
    if (!pr->planar) {
        j = (int)(pt.y / hx + pr->BoxY / 2 * ref);
        dy = pt.y / hx - (j - pr->BoxY / 2 * ref + 0.5);
        j = (int)ceil(dy) + j - 1;
    }
    return height_interp(f->height, f->nhx, f->nhy, i, j, dx, dy);
}

/* F_surf_tension: ABM/Forces.cpp:209-317 */
static void F_surf_tension(const bcg_params* pr, const bcg_cell* c, const fields_t* f, v3* F, v3* T) {
    const int ref = pr->refinementGridHeight;
    v3 cm = cmof(c);
    int xa = (int)((cm.x + pr->BoxX / 2 * pr->BoxLength) / (pr->BoxLength / ref));
    int ya = 0;
    if (!pr->planar) ya = (int)((cm.y + pr->BoxY / 2 * pr->BoxLength) / (pr->BoxLength / ref));
    /* the reference does not clamp (an index outside reads foreign memory) */
    if (xa < 0) xa = 0;
    if (xa > f->nhx - 1) xa = f->nhx - 1;
    if (ya < 0) ya = 0;
    if (ya > f->nhy - 1) ya = f->nhy - 1;
    double height_p = surf_height(pr, f, P(c));
    double height_q = surf_height(pr, f, Q(c));
    double height_cm = surf_height(pr, f, cm);
    if (height_cm > (cm.z + pr->BoxLength)) {
        *F = mk(0, 0, 0);
        *T = mk(0, 0, 0);
        return;
    }
    v3 N = V(f->normal + 3 * (xa * f->nhy + ya));
    v3 F1 = mk(0, 0, 0), F2 = mk(0, 0, 0);
    double dh = c->p[2] - (height_p - pr->DH / N.z);
    if (!(dh < 0.0)) {
        double distance = -dh * N.z;
        double F0 = 2 * PI_ * pr->tension * dmin(distance / (pr->cellRadius / 5.0), 1.0);
        F1 = vscale(N, F0);
    }
    dh = c->q[2] - (height_q - pr->DH / N.z);
    if (!(dh < 0.0)) {
        double distance = -dh * N.z;
        double F0 = 2 * PI_ * pr->tension * dmin(distance / (pr->cellRadius / 5.0), 1.0);
        F2 = vscale(N, F0);
    }
    *F = vadd(F1, F2);
    v3 r = vsub(P(c), cm);
    *T = vcross(r, F1);
    r = vsub(Q(c), cm);
    *T = vadd(*T, vcross(r, F2));
}

/* ------------------------------------------------------------- grid --- */
/* UniformGrid::GetAddress: ABM/UniformGrid.h:87-100 (int() truncates). */
static int get_address(const bcg_params* pr, v3 pos, int* xi, int* yi, int* zi) {
    int x = (int)(pos.x / pr->BoxLength + pr->BoxX / 2);
    int y = (int)(pos.y / pr->BoxLength + pr->BoxY / 2 + 0.5);
    int z = (int)(pos.z / pr->BoxLength);
    if (xi) { *xi = x; *yi = y; *zi = z; }
    if (!((x < pr->BoxX) & (y < pr->BoxY) & (z < pr->BoxZ) & (x >= 0) & (y >= 0) & (z >= 0))) return -1;
    return x * (pr->BoxY * pr->BoxZ) + y * pr->BoxZ + z;
}

struct or_sim {
    bcg_params p;
    int N;
    bcg_cell* old_cells;
    bcg_cell* new_cells;
    int nbox, boxcap;
    int* box_count;   /* [nbox]          */
    int* box_cells;   /* [nbox][boxcap]  */
    int* nl_count;    /* [maxCells]      */
    int* nl;          /* [maxCells][cap] */
    int cap;
    int* scratch;     /* gather buffer   */
    int scratch_cap;
    fields_t f;
    double* height;
    double* normal;
    double* wall;
    double* gr1;
    double* gr2;
    int* dividing;
    int* idlist;
    int64_t wall_draws;
    int err;
};

/* Box::Add / Box::Remove: ABM/UniformGrid.h:33-55 */
static int box_add(or_sim* s, int a, int id) {
    if (a < 0) return BCG_ERR_OUT_OF_DOMAIN;
    if (s->box_count[a] + 1 >= s->boxcap) return BCG_ERR_CAPACITY;
    s->box_cells[(size_t)a * s->boxcap + s->box_count[a]++] = id;
    return 0;
}
static int box_remove(or_sim* s, int a, int id) {
    if (a < 0) return BCG_ERR_OUT_OF_DOMAIN;
    int* b = s->box_cells + (size_t)a * s->boxcap;
    int ii = 0;
    while (ii < s->box_count[a] && b[ii] != id) ii++;
    if (ii == s->box_count[a]) return BCG_ERR_STATE;
    b[ii] = b[s->box_count[a] - 1];
    s->box_count[a]--;
    return 0;
}

/* reduceNeighbours' test of one pair (ABM/Neighbours.cpp:72-88): 1 when
 * neighbour o stays in the list of cell c.                                 */
static int keep_neighbour(const bcg_cell* c, const bcg_cell* o, double Rc) {
    v3 cm = cmof(c);
    const double reach = c->Length / 2.0 + Rc * c->Radius;
    v3 v = vsub(P(o), cm);
    double dp = sqrt(vdot(v, v));
    v = vsub(Q(o), cm);
    double dq = sqrt(vdot(v, v));
    if ((dp < reach) || (dq < reach)) {
        double d; v3 c1, c2;
        min_distance(c, o, &d, &c1, &c2);
        return d < Rc * c->Radius;
    }
    return 0;
}

/* The test of pairs (a[i], b[i]), i < n. */
void or_keep_batch(const bcg_cell* a, const bcg_cell* b, int n, double Rc, int* out) {
    for (int i = 0; i < n; i++) out[i] = keep_neighbour(&a[i], &b[i], Rc);
}

/* UniformGrid::GetNeighbours (ABM/UniformGrid.h:141-170) followed by
 * reduceNeighbours (ABM/Neighbours.cpp:57-93) for one cell.               */
static int gather_into(or_sim* s, const bcg_cell* cells, int id, int* scratch) {
    const bcg_params* pr = &s->p;
    int x0, y0, z0;
    const bcg_cell* c = &cells[id];
    if (get_address(pr, cmof(c), &x0, &y0, &z0) < 0) return BCG_ERR_OUT_OF_DOMAIN;
    int n = 0;
    for (int za = (z0 - 1 > 0 ? z0 - 1 : 0); za < (z0 + 2 < pr->BoxZ ? z0 + 2 : pr->BoxZ); za++)
        for (int ya = (y0 - 1 > 0 ? y0 - 1 : 0); ya < (y0 + 2 < pr->BoxY ? y0 + 2 : pr->BoxY); ya++)
            for (int xa = (x0 - 1 > 0 ? x0 - 1 : 0); xa < (x0 + 2 < pr->BoxX ? x0 + 2 : pr->BoxX); xa++) {
                int a = xa * (pr->BoxY * pr->BoxZ) + ya * pr->BoxZ + za;
                const int* b = s->box_cells + (size_t)a * s->boxcap;
                for (int k = 0; k < s->box_count[a]; k++)
                    if (b[k] != id) {
                        if (n >= s->scratch_cap) return BCG_ERR_NEIGHBOUR_OVERFLOW;
                        scratch[n++] = b[k];
                    }
            }
    int m = 0;
    for (int k = 0; k < n; k++) {
        if (keep_neighbour(c, &cells[scratch[k]], pr->Rc)) {
            if (m >= s->cap) return BCG_ERR_NEIGHBOUR_OVERFLOW;
            s->nl[(size_t)id * s->cap + m++] = scratch[k];
        }
    }
    s->nl_count[id] = m;
    return 0;
}

static int gather(or_sim* s, const bcg_cell* cells, int id) { return gather_into(s, cells, id, s->scratch); }

/* sum_forces: ABM/Forces.cpp:332-381 */
static void sum_forces_d(or_sim* s, const bcg_cell* c, const bcg_cell* cells, const int* nl, int nn, int wx, int wy,
                         v3* Fnet, v3* Tnet, v3* sf, v3* df, int64_t* draws, int draw) {
    const bcg_params* pr = &s->p;
    *Fnet = mk(0, 0, 0);
    *Tnet = mk(0, 0, 0);
    v3 F, F2, T, r, r2;
    v3 cm = cmof(c);
    double d;
    for (int k = 0; k < nn; k++) {
        F_cc(pr, c, &cells[nl[k]], &F, &r, &d);
        *Fnet = vadd(*Fnet, F);
        r = vsub(r, cm);
        *Tnet = vadd(*Tnet, vcross(r, F));
    }
    *sf = mk(0, 0, 0);
    *df = mk(0, 0, 0);
    if (dmin(c->q[2], c->p[2]) < 1.2 * c->Radius) {
        double wall_z = s->f.wall[wx * pr->BoxY + wy];
        F_cw_d(pr, c, wall_z, &F, &F2, &r, &r2, sf, df, draw);
        (*draws)++;
        *Fnet = vadd(*Fnet, F);
        r = vsub(r, cm);
        *Tnet = vadd(*Tnet, vcross(r, F));
        *Fnet = vadd(*Fnet, F2);
        r2 = vsub(r2, cm);
        *Tnet = vadd(*Tnet, vcross(r2, F2));
    }
    /* F_v: ABM/Forces.cpp:320-329 */
    F = vscale(V(c->Velocity), -pr->viscosity * 6.0 * PI_ * c->Radius);
    T = vscale(V(c->AngularVelocity), -pr->viscosity * 6.0 * PI_ * c->Radius);
    *Fnet = vadd(*Fnet, F);
    *Tnet = vadd(*Tnet, T);
    F_surf_tension(pr, c, &s->f, &F, &T);
    *Fnet = vadd(*Fnet, F);
    *Tnet = vadd(*Tnet, T);
    if (pr->planar) {
        Tnet->z = 0;
        Fnet->y = 0;
    }
}

static void sum_forces(or_sim* s, const bcg_cell* c, const bcg_cell* cells, const int* nl, int nn, int wx, int wy,
                       v3* Fnet, v3* Tnet, v3* sf, v3* df) {
    sum_forces_d(s, c, cells, nl, nn, wx, wy, Fnet, Tnet, sf, df, &s->wall_draws, 1);
}

/* Half kick of UpdatePositions / UpdateVelocities: ABM/Integrate.cpp:47-65
 * (identical to :106-124). Writes the new velocity and angular velocity.  */
static void half_kick(double dt, v3 F, v3 T, const bcg_cell* o, bcg_cell* n) {
    double R2 = o->Radius * o->Radius;
    double L2 = o->Length * o->Length;
    double M = PI_ * R2 * (o->Length + 4.0 / 3.0 * o->Radius);
    double Ixy = PI_ * R2 * (o->Length * L2 / 12.0 + o->Radius * (8.0 * R2 / 15.0 + L2 / 3.0));
    double Iz = PI_ * R2 * R2 * (o->Length / 2.0 + 8.0 * o->Radius / 15.0);
    v3 z = vsub(Q(o), P(o));
    z = vscale(z, 1.0 / o->Length);
    v3 w = V(o->AngularVelocity);
    v3 vaz = vscale(z, vdot(z, w));
    v3 vaxy = vsub(w, vaz);
    v3 Tz = vscale(z, vdot(z, T));
    v3 Txy = vsub(T, Tz);
    st(n->Velocity, vadd(V(o->Velocity), vscale(F, dt / (2.0 * M))));
    st(n->AngularVelocity, vadd(vadd(vaz, vscale(Tz, dt / (2.0 * Iz))), vadd(vaxy, vscale(Txy, dt / (2.0 * Ixy)))));
}

/* UpdatePositions: ABM/Integrate.cpp:34-90 */
static void update_positions(int planar, double dt, v3 F, v3 T, const bcg_cell* o, bcg_cell* n) {
    *n = *o;
    half_kick(dt, F, T, o, n);
    v3 cm = cmof(o);
    v3 rcmp = vsub(P(o), cm), rcmq = vsub(Q(o), cm);
    v3 w = V(n->AngularVelocity), v = V(n->Velocity);
    v3 vp = vadd(v, vcross(w, rcmp)), vq = vadd(v, vcross(w, rcmq));
    st(n->p, vadd(P(o), vscale(vp, dt)));
    st(n->q, vadd(Q(o), vscale(vq, dt)));
    if (planar) {
        n->p[1] = 0;
        n->q[1] = 0;
    }
}

/* UpdateVelocities: ABM/Integrate.cpp:93-127 */
static void update_velocities(int planar, double dt, v3 F, v3 T, const bcg_cell* o, bcg_cell* n) {
    *n = *o;
    half_kick(dt, F, T, o, n);
    if (planar) n->Velocity[1] = 0;
}

/* grow: ABM/grow.cpp:17-102 (live part; the lambda interpolation is
 * commented out in the reference). */
static int grow(or_sim* s, double dt, bcg_cell* c) {
    const bcg_params* pr = &s->p;
    int a = get_address(pr, cmof(c), NULL, NULL, NULL);
    if (a < 0) return BCG_ERR_OUT_OF_DOMAIN;
    double gr;
    if (pr->crossFeeding) gr = (c->Type == 1) ? s->gr1[a] : s->gr2[a];
    else gr = s->gr1[a];
    gr = gr * 1.5850;
    v3 v = vsub(Q(c), P(c));
    double dL = c->Length * dt * gr;
    v3 dv = vscale(v, dL * 0.5 / c->Length);
    c->Length += dL;
    st(c->p, vsub(P(c), dv));
    st(c->q, vadd(Q(c), dv));
    c->GrowthRate = gr;
    return 0;
}

/* divide: ABM/grow.cpp:107-186. Draw order: dl, dangle, ddir1 xyz,
 * ddir2 xyz, daughter angular velocity z. */
static void divide(const bcg_params* pr, bcg_cell* mother, bcg_cell* daughter, double t) {
    double dl = draw_centered() * pr->varL;
    double dangle = draw_centered() * pr->varAngle;
    v3 pq = vsub(P(mother), Q(mother));
    double ddx1 = draw_centered(), ddy1 = draw_centered(), ddz1 = draw_centered();
    v3 ddir1 = mk(ddx1, ddy1, ddz1);
    ddir1 = vsub(ddir1, vscale(pq, vdot(ddir1, pq) / vdot(pq, pq)));
    ddir1 = vscale(ddir1, 1 / sqrt(vdot(ddir1, ddir1)));
    v3 dOrien1 = vscale(ddir1, dangle);
    double ddx2 = draw_centered(), ddy2 = draw_centered(), ddz2 = draw_centered();
    v3 ddir2 = mk(ddx2, ddy2, ddz2);
    ddir2 = vsub(ddir2, vscale(pq, vdot(ddir2, pq) / vdot(pq, pq)));
    ddir2 = vscale(ddir2, 1 / sqrt(vdot(ddir2, ddir2)));
    v3 dOrien2 = vscale(ddir2, dangle);

    v3 uv = vsub(P(mother), Q(mother));
    double L = mother->Length;
    uv = vscale(uv, 1 / L);
    st(daughter->q, Q(mother));
    daughter->Radius = pr->cellRadius;
    mother->Length = L / 2 - mother->Radius + dl;
    daughter->Length = L / 2 + mother->Radius - 2 * daughter->Radius - dl;
    daughter->GrowthRate = mother->GrowthRate;
    memcpy(daughter->Velocity, mother->Velocity, sizeof(double) * 3);
    double wz = draw_centered() * 0.001;
    st(daughter->AngularVelocity, mk(0, 0, wz));
    st(mother->AngularVelocity, mk(0, 0, 0));
    st(mother->q, vsub(P(mother), vscale(uv, mother->Length)));
    daughter->time_q = mother->time_q;
    mother->time_q = t;
    daughter->time_p = t;
    mother->age_p++;
    daughter->age_q = mother->age_q + 1;
    st(daughter->p, vsub(Q(mother), vscale(uv, daughter->Radius + mother->Radius)));
    st(daughter->q, vsub(P(daughter), vscale(uv, daughter->Length)));
    st(mother->p, vadd(P(mother), vscale(dOrien1, mother->Length)));
    st(mother->q, vsub(Q(mother), vscale(dOrien1, mother->Length)));
    st(daughter->p, vadd(P(daughter), vscale(dOrien2, daughter->Length)));
    st(daughter->q, vsub(Q(daughter), vscale(dOrien2, daughter->Length)));
    daughter->Type = mother->Type;
    daughter->Ancestor = mother->Ancestor;
    mother->Ldiv = pr->L_divide / 2 + pr->cellRadius + mother->Length;
    daughter->Ldiv = pr->L_divide / 2 + pr->cellRadius + daughter->Length;
}

/* -------------------------------------------------------- simulation --- */
or_sim* or_sim_create(const bcg_params* p) {
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    s->p = *p;
    s->nbox = p->BoxX * p->BoxY * p->BoxZ;
    s->boxcap = 256;
    s->box_count = (int*)calloc(s->nbox, sizeof(int));
    s->box_cells = (int*)malloc(sizeof(int) * (size_t)s->nbox * s->boxcap);
    s->cap = p->maxNeighbours;
    s->nl_count = (int*)calloc(p->maxCells, sizeof(int));
    s->nl = (int*)malloc(sizeof(int) * (size_t)p->maxCells * s->cap);
    s->scratch_cap = 27 * s->boxcap;
    s->scratch = (int*)malloc(sizeof(int) * s->scratch_cap);
    s->old_cells = (bcg_cell*)calloc(p->maxCells, sizeof(bcg_cell));
    s->new_cells = (bcg_cell*)calloc(p->maxCells, sizeof(bcg_cell));
    int ref = p->refinementGridHeight;
    s->f.nhx = p->BoxX * ref;
    s->f.nhy = p->BoxY * ref;
    size_t nh = (size_t)s->f.nhx * s->f.nhy;
    s->height = (double*)malloc(sizeof(double) * nh);
    s->normal = (double*)malloc(sizeof(double) * nh * 3);
    for (size_t i = 0; i < nh; i++) {
        s->height[i] = p->cellRadius;
        s->normal[3 * i] = 0; s->normal[3 * i + 1] = 0; s->normal[3 * i + 2] = -1;
    }
    s->wall = (double*)calloc((size_t)p->BoxX * p->BoxY, sizeof(double));
    s->gr1 = (double*)calloc(s->nbox, sizeof(double));
    s->gr2 = (double*)calloc(s->nbox, sizeof(double));
    s->f.height = s->height;
    s->f.normal = s->normal;
    s->f.wall = s->wall;
    s->dividing = (int*)malloc(sizeof(int) * p->maxCells);
    s->idlist = (int*)malloc(sizeof(int) * (s->cap + 2));
    return s;
}

void or_sim_destroy(or_sim* s) {
    free(s->box_count); free(s->box_cells); free(s->nl_count); free(s->nl); free(s->scratch);
    free(s->old_cells); free(s->new_cells); free(s->height); free(s->normal); free(s->wall);
    free(s->gr1); free(s->gr2); free(s->dividing); free(s->idlist); free(s);
}

/* LoadCells grid insertion in ID order: ABM/InputOutput.cpp:422-426 */
int or_sim_set_cells(or_sim* s, const bcg_cell* cells, int n) {
    if (n > s->p.maxCells) return BCG_ERR_CAPACITY;
    memcpy(s->old_cells, cells, sizeof(bcg_cell) * n);
    memset(s->box_count, 0, sizeof(int) * s->nbox);
    memset(s->nl_count, 0, sizeof(int) * s->p.maxCells);
    s->N = n;
    for (int i = 0; i < n; i++) {
        int rc = box_add(s, get_address(&s->p, cmof(&cells[i]), NULL, NULL, NULL), i);
        if (rc) return rc;
    }
    return 0;
}

void or_sim_set_fields(or_sim* s, const double* height, const double* normal, const double* wall,
                       const double* gr1, const double* gr2) {
    size_t nh = (size_t)s->f.nhx * s->f.nhy;
    if (height) memcpy(s->height, height, sizeof(double) * nh);
    if (normal) memcpy(s->normal, normal, sizeof(double) * nh * 3);
    if (wall) memcpy(s->wall, wall, sizeof(double) * s->p.BoxX * s->p.BoxY);
    if (gr1) memcpy(s->gr1, gr1, sizeof(double) * s->nbox);
    if (gr2) memcpy(s->gr2, gr2, sizeof(double) * s->nbox);
}

int or_sim_num_cells(or_sim* s) { return s->N; }
void or_sim_get_cells(or_sim* s, bcg_cell* out) { memcpy(out, s->old_cells, sizeof(bcg_cell) * s->N); }
int64_t or_sim_wall_draws(or_sim* s) { return s->wall_draws; }

int or_sim_rebuild(or_sim* s) {
    for (int i = 0; i < s->N; i++) {
        int rc = gather(s, s->old_cells, i);
        if (rc) return rc;
    }
    return 0;
}

int or_sim_rebuild_ids(or_sim* s, const int* ids, int n) {
    for (int i = 0; i < n; i++) {
        int rc = gather(s, s->old_cells, ids[i]);
        if (rc) return rc;
    }
    return 0;
}

void or_sim_get_neighbours(or_sim* s, int* counts, int* lists, int stride) {
    for (int i = 0; i < s->N; i++) {
        counts[i] = s->nl_count[i];
        for (int k = 0; k < s->nl_count[i] && k < stride; k++) lists[(size_t)i * stride + k] = s->nl[(size_t)i * s->cap + k];
    }
}

void or_sim_get_boxes(or_sim* s, int* counts, int* ids, int stride) {
    for (int a = 0; a < s->nbox; a++) {
        counts[a] = s->box_count[a];
        for (int k = 0; k < s->box_count[a] && k < stride; k++) ids[(size_t)a * stride + k] = s->box_cells[(size_t)a * s->boxcap + k];
    }
}

void or_sim_sum_forces(or_sim* s, double* F, double* T) {
    int64_t saved = s->wall_draws;
    for (int i = 0; i < s->N; i++) {
        int x, y, z;
        get_address(&s->p, cmof(&s->old_cells[i]), &x, &y, &z);
        v3 f, t, sf, df;
        /* diagnostic pass: F_cw still takes its rand() draws */
        sum_forces(s, &s->old_cells[i], s->old_cells, s->nl + (size_t)i * s->cap, s->nl_count[i], x, y, &f, &t, &sf, &df);
        st(F + 3 * i, f);
        st(T + 3 * i, t);
    }
    s->wall_draws = saved;
}

/* mean_stress: ABM/Compute.cpp:18-117, every cell as the output step calls
 * it (ABM/Simulation.cpp:320-326); F_cw takes its rand() draws.            */
void or_sim_mean_stress(or_sim* s, double* S5, double* F3) {
    const bcg_params* pr = &s->p;
    for (int i = 0; i < s->N; i++) {
        const bcg_cell* c = &s->old_cells[i];
        v3 cm = cmof(c);
        int x, y, z;
        get_address(pr, cm, &x, &y, &z);
        double R = c->Radius;
        double V = c->Length * PI_ * R * R + 4.0 / 3.0 * PI_ * R * R * R;
        double xx = 0, yy = 0, zz = 0;
        v3 Fn = mk(0, 0, 0), F, F2, r, r2, T, sf, df;
        double d;
        const int* nl = s->nl + (size_t)i * s->cap;
        for (int k = 0; k < s->nl_count[i]; k++) {
            F_cc(pr, c, &s->old_cells[nl[k]], &F, &r, &d);
            Fn = vadd(Fn, F);
            r = vsub(r, cm);
            xx -= F.x * r.x;
            yy -= F.y * r.y;
            zz -= F.z * r.z;
        }
        if (dmin(c->q[2], c->p[2]) < 1.2 * R) {
            double wall_z = s->f.wall[x * pr->BoxY + y];
            F_cw(pr, c, wall_z, &F, &F2, &r, &r2, &sf, &df);
            Fn = vadd(Fn, F);
            r = vsub(r, cm);
            xx -= F.x * r.x;
            yy -= F.y * r.y;
            zz -= F.z * r.z;
            Fn = vadd(Fn, F2);
            r2 = vsub(r2, cm);
            xx -= F2.x * r2.x;
            yy -= F2.y * r2.y;
            zz -= F2.z * r2.z;
        }
        F_surf_tension(pr, c, &s->f, &F, &T);
        Fn = vadd(Fn, F);
        r = vsub(P(c), cm);
        r.z += R;
        r2 = vsub(Q(c), cm);
        r2.z += R;
        xx -= F.x * r.x * 0.5 - F.x * r2.x * 0.5;
        yy -= F.y * r.y * 0.5 - F.y * r2.y * 0.5;
        zz -= F.z * r.z * 0.5 - F.z * r2.z * 0.5;
        xx /= V;
        yy /= V;
        zz /= V;
        double rr, tt, Lcm = sqrt(cm.x * cm.x + cm.y * cm.y);
        if (Lcm > DBL_EPSILON) {
            double ux = cm.x * (1.0 / Lcm), uy = cm.y * (1.0 / Lcm);
            rr = xx * ux + yy * uy;
            tt = -xx * uy + yy * ux;
        } else {
            rr = xx + yy;
            tt = 0.0;
        }
        double* o = S5 + 5 * (size_t)i;
        o[0] = xx;
        o[1] = yy;
        o[2] = zz;
        o[3] = rr;
        o[4] = tt;
        st(F3 + 3 * i, Fn);
    }
}

/* One step of ABM/Simulation.cpp:330-413 (single thread). */
int or_sim_step(or_sim* s, double dt, double t, int rebuild, int* lineage, int lineage_cap) {
    const bcg_params* pr = &s->p;
    int N = s->N;
    if (rebuild) {
        int rc = or_sim_rebuild(s);
        if (rc) return rc;
    }
    int numDivide = 0;
    for (int id = 0; id < N; id++) {
        /* MoveCell: ABM/Cells.cpp:14-39 */
        const bcg_cell* o = &s->old_cells[id];
        bcg_cell* n = &s->new_cells[id];
        int x, y, z;
        int olda = get_address(pr, cmof(o), &x, &y, &z);
        if (olda < 0) return BCG_ERR_OUT_OF_DOMAIN;
        /* integrate: ABM/Integrate.cpp:11-31 */
        v3 F, T, sf, df;
        bcg_cell temp;
        const int* nl = s->nl + (size_t)id * s->cap;
        sum_forces(s, o, s->old_cells, nl, s->nl_count[id], x, y, &F, &T, &sf, &df);
        update_positions(pr->planar, dt, F, T, o, &temp);
        sum_forces(s, &temp, s->old_cells, nl, s->nl_count[id], x, y, &F, &T, &sf, &df);
        update_velocities(pr->planar, dt, F, T, &temp, n);
        st(n->DynFric, df);
        st(n->StaFric, sf);
        int newa = get_address(pr, cmof(n), NULL, NULL, NULL);
        if (newa < 0) return BCG_ERR_OUT_OF_DOMAIN;
        if (newa != olda) {
            int rc = box_remove(s, olda, id);
            if (!rc) rc = box_add(s, newa, id);
            if (rc) return rc;
        }
        /* GrowCell: ABM/Cells.cpp:42-57 */
        int rc = grow(s, dt, n);
        if (rc) return rc;
        if (n->Length >= n->Ldiv) s->dividing[numDivide++] = id;
    }
    int done = 0;
    for (int k = 0; k < numDivide; k++) {
        if (N > (pr->maxCells - 1)) break;
        const int id = s->dividing[k];
        /* DivideCell: ABM/Cells.cpp:60-81 */
        bcg_cell* m = &s->new_cells[id];
        bcg_cell* d = &s->new_cells[N];
        memset(d, 0, sizeof *d);
        int rc = box_remove(s, get_address(pr, cmof(m), NULL, NULL, NULL), id);
        if (rc) return rc;
        divide(pr, m, d, t);
        rc = box_add(s, get_address(pr, cmof(m), NULL, NULL, NULL), id);
        if (!rc) rc = box_add(s, get_address(pr, cmof(d), NULL, NULL, NULL), N);
        if (rc) return rc;
        if (done < lineage_cap) { lineage[2 * done] = id; lineage[2 * done + 1] = N; }
        done++;
        N++;
        /* neighbour patch: ABM/Simulation.cpp:396-403 */
        int len = s->nl_count[id];
        s->idlist[0] = id;
        memcpy(s->idlist + 1, s->nl + (size_t)id * s->cap, sizeof(int) * len);
        s->idlist[len + 1] = N - 1;
        len += 2;
        for (int i = 0; i < len; i++) {
            rc = gather(s, s->new_cells, s->idlist[i]);
            if (rc) return rc;
        }
    }
    bcg_cell* tmp = s->new_cells;
    s->new_cells = s->old_cells;
    s->old_cells = tmp;
    s->N = N;
    return done;
}

/* ---------------------------------------------- OpenMP form of the step ---
 * The same step with the per-cell work on all threads: the list rebuild
 * (each thread its own gather buffer; the boxes do not change while it
 * runs) and MoveCell + GrowCell (each cell reads the old buffer and its own
 * list, writes only its new record). What the reference does in order stays
 * in order on one thread, in ascending ID: the box moves (the boxes'
 * contents order depends on it), the dividing list, the wall-contact draws
 * (taken after the loop: with var_pos == 0 their values are not used), and
 * the division loop. The result equals or_sim_step bit for bit.            */
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads(void) { return omp_get_max_threads(); }

int or_sim_step_par(or_sim* s, double dt, double t, int rebuild, int* lineage, int lineage_cap) {
    const bcg_params* pr = &s->p;
    if (pr->var_pos != 0.0) return BCG_ERR_UNSUPPORTED;
    const int N = s->N;
    int err = 0;
    if (rebuild) {
#pragma omp parallel reduction(| : err)
        {
            int* scratch = (int*)malloc(sizeof(int) * s->scratch_cap);
#pragma omp for schedule(dynamic, 256)
            for (int i = 0; i < N; i++) {
                int rc = gather_into(s, s->old_cells, i, scratch);
                if (rc) err |= 1 << (-rc);
            }
            free(scratch);
        }
        if (err) {
            for (int c = 1; c < 16; c++)
                if (err & (1 << c)) return -c;
        }
    }
    int* olda = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    int* newa = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    int64_t draws = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : draws) reduction(| : err)
    for (int id = 0; id < N; id++) {
        const bcg_cell* o = &s->old_cells[id];
        bcg_cell* n = &s->new_cells[id];
        int x, y, z;
        olda[id] = get_address(pr, cmof(o), &x, &y, &z);
        if (olda[id] < 0) {
            err |= 1 << (-BCG_ERR_OUT_OF_DOMAIN);
            continue;
        }
        v3 F, T, sf, df;
        bcg_cell temp;
        const int* nl = s->nl + (size_t)id * s->cap;
        sum_forces_d(s, o, s->old_cells, nl, s->nl_count[id], x, y, &F, &T, &sf, &df, &draws, 0);
        update_positions(pr->planar, dt, F, T, o, &temp);
        sum_forces_d(s, &temp, s->old_cells, nl, s->nl_count[id], x, y, &F, &T, &sf, &df, &draws, 0);
        update_velocities(pr->planar, dt, F, T, &temp, n);
        st(n->DynFric, df);
        st(n->StaFric, sf);
        newa[id] = get_address(pr, cmof(n), NULL, NULL, NULL);
        if (newa[id] < 0) {
            err |= 1 << (-BCG_ERR_OUT_OF_DOMAIN);
            continue;
        }
        int rc = grow(s, dt, n);
        if (rc) err |= 1 << (-rc);
    }
    int rc = 0;
    if (err) {
        for (int c = 1; c < 16 && !rc; c++)
            if (err & (1 << c)) rc = -c;
    }
    int numDivide = 0;
    for (int id = 0; id < N && !rc; id++) {
        if (newa[id] != olda[id]) {
            rc = box_remove(s, olda[id], id);
            if (!rc) rc = box_add(s, newa[id], id);
        }
        if (s->new_cells[id].Length >= s->new_cells[id].Ldiv) s->dividing[numDivide++] = id;
    }
    free(olda);
    free(newa);
    if (rc) return rc;
    s->wall_draws += draws;
    rand_discard(draws);
    /* divisions and list patch, in order (as or_sim_step) */
    int Nn = N, done = 0;
    for (int k = 0; k < numDivide; k++) {
        if (Nn > (pr->maxCells - 1)) break;
        const int id = s->dividing[k];
        bcg_cell* m = &s->new_cells[id];
        bcg_cell* d = &s->new_cells[Nn];
        memset(d, 0, sizeof *d);
        rc = box_remove(s, get_address(pr, cmof(m), NULL, NULL, NULL), id);
        if (rc) return rc;
        divide(pr, m, d, t);
        rc = box_add(s, get_address(pr, cmof(m), NULL, NULL, NULL), id);
        if (!rc) rc = box_add(s, get_address(pr, cmof(d), NULL, NULL, NULL), Nn);
        if (rc) return rc;
        if (done < lineage_cap) { lineage[2 * done] = id; lineage[2 * done + 1] = Nn; }
        done++;
        Nn++;
        int len = s->nl_count[id];
        s->idlist[0] = id;
        memcpy(s->idlist + 1, s->nl + (size_t)id * s->cap, sizeof(int) * len);
        s->idlist[len + 1] = Nn - 1;
        len += 2;
        for (int i = 0; i < len; i++) {
            rc = gather(s, s->new_cells, s->idlist[i]);
            if (rc) return rc;
        }
    }
    bcg_cell* tmp = s->new_cells;
    s->new_cells = s->old_cells;
    s->old_cells = tmp;
    s->N = Nn;
    return done;
}
