/*
 * dem_oracle.c -- CPU FP64 restatement of the reference's bonded-particle
 * (BPM) loading step, used ONLY as the checker of the HIP path and as the
 * timed CPU baseline.  TEST INFRASTRUCTURE: the product never links it.
 *
 * The reference (MSMS-LAB/DEM-Rock-Calibration) runs this loop in FP32 CUDA
 * kernels.  This file restates the live kernels' semantics in FP64 with plain
 * loops and the reference's own data flow (8-cell stencil over a stable cell
 * sort, full row scans, slot-order gathers).  Every function cites the
 * reference file:line it follows.
 *
 * Floating-point contract (shared with the device code by specification, not
 * by code): no FMA contraction (-ffp-contract=off), IEEE sqrt and division,
 * left-to-right evaluation exactly as written, correctly rounded 1/sqrt(x)
 * restated as 1.0/sqrt(x).  Per particle, forces accumulate bond slots in slot
 * order, then contact slots in slot order, then the top and bottom platens.
 *
 * Generalisations over the reference (which has one radius p_r and one
 * modulus m_E): per-particle radius r_i, mass m_i = rho*4/3*pi*r_i^3, inertia
 * 0.4*m_i*r_i^2; contact effective radius r_i*r_j/(r_i+r_j), effective mass
 * m_i*m_j/(m_i+m_j), overlap r_i+r_j-rho; separate contact modulus E_contact;
 * per-bond radius ratio*min(r_i,r_j) and per-bond lognormal strengths.  With
 * equal radii and E_contact == E_bond they are algebraically the reference
 * formulas.
 *
 * Parity status of each part is recorded next to it ("pinned by ...").
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.1415926535897932384626433832795 /* MC_pi, md_math_constants.h */
#define DAMP (-1.8257)                       /* -1.8257 = -2*sqrt(5/6), md_calculateforcesdem3.cu:248 */
#define TWO_THIRDS (2.0 / 3.0)               /* MCf_2d3 */
#define RED_BLOCK 32                         /* particles per platen partial (four device waves) */
#define CTRL_T 1024                          /* threads of the final platen-force sum */
#define LANES 8                              /* lanes sharing one particle on the device */

struct ora_ctx {
    int64_t n;
    int32_t S;
    int32_t nthreads;
    double *R, *V, *W, *F, *M;          /* [3n] x | y | z blocks               */
    double *rad, *mass, *dtdm, *dtdI;   /* per particle                        */
    int32_t* nbr;                       /* [n*S], -1 = empty                   */
    uint8_t* type;                      /* 1 bond, 0 contact                   */
    int8_t* bidx;                       /* bond-table index of a bond slot     */
    double *invL0, *sc, *tc;            /* [n*S]                               */
    double *pn, *ot, *mn, *mt;          /* [n*S*3]                             */
    /* cells */
    int64_t ncell;
    int32_t *key, *sorted, *cstart, *ccount;
    rf_material mat;
    rf_derived der;
    rf_box box;
    double acut2, inv_a;
    /* test */
    rf_platen_cfg cfg;
    int has_test;
    double top, bottom, vw, size_axis, size0, strain, stress, peak_stress, peak_strain;
    double sfl0, sfl1, inv_area, area, rr_scope, zcut, zero_stress, ld, destruction;
    double scope_c[3];
    int64_t step, steps_stop, stepsave;
    int loading, fracture, done;
    double* curve;                      /* [3*cap]                             */
    int64_t* bcount;                    /* [cap]                               */
    int64_t curve_cap;
    double *part_b, *part_t;            /* per-particle platen contributions   */
    int64_t* ord;                       /* particle at place p of the platen-sum order */
    int64_t nblk;
    double *blk_b, *blk_t;              /* block partials                      */
    char err[256];
};

static void set_err(ora_ctx* c, const char* m) {
    snprintf(c->err, sizeof(c->err), "%s", m);
}

const char* ora_error(const ora_ctx* c) { return c ? c->err : "null context"; }

/* sizeof of the public structs as the C compiler lays them out (ABI check). */
int64_t ora_sizeof(int32_t which) {
    switch (which) {
        case 0: return (int64_t)sizeof(rf_material);
        case 1: return (int64_t)sizeof(rf_box);
        case 2: return (int64_t)sizeof(rf_sample);
        case 3: return (int64_t)sizeof(rf_options);
        case 4: return (int64_t)sizeof(rf_derived);
        case 5: return (int64_t)sizeof(rf_platen_cfg);
        case 6: return (int64_t)sizeof(rf_test_state);
        case 7: return (int64_t)sizeof(rf_counters);
        case 8: return (int64_t)sizeof(rf_state_view);
        case 9: return (int64_t)sizeof(rf_fire_cfg);
        case 10: return (int64_t)sizeof(rf_fire_result);
        case 11: return (int64_t)sizeof(rf_result);
        case 12: return (int64_t)sizeof(rf_adjoint_cfg);
        default: return -1;
    }
}

void ora_set_threads(ora_ctx* c, int32_t nthreads) {
    c->nthreads = nthreads > 0 ? nthreads : 1;
}

/* ---------------------------------------------------------------- a1 ---- */
/* potential_data::CalculateParameters, md_data_types.h:155-187.  Pinned by
 * closed-form values in tests (T_R, T_b, hm_E, p_A). */
int ora_derive(const rf_material* m, double r_min, double r_max, rf_derived* d) {
    if (!m || !d || !(r_min > 0.0) || r_max < r_min) return RF_ERR_ARG;
    d->r_min = r_min;
    d->r_max = r_max;
    d->G_bond = 0.5 * m->E_bond / (1.0 + m->nu);
    d->G_contact = 0.5 * m->E_contact / (1.0 + m->nu);
    d->hm_E = 0.5 * m->E_contact / (1.0 - m->nu * m->nu);
    d->hm_G = 0.5 * d->G_contact / (2.0 - m->nu);
    double le = log(m->restitution);
    d->p_A = le / sqrt(PI * PI + le * le);
    double r = r_min;
    d->T_rayleigh = PI * r * sqrt(2.0 * m->density * (1.0 + m->nu)) /
                    (sqrt(m->E_contact) * (0.163 * m->nu + 0.8766));
    double b = m->bond_radius_ratio > 0.0 ? m->bond_radius_ratio * r : m->bond_radius;
    double bS = PI * b * b;
    double pV = 4.0 * (1.0 / 3.0) * PI * r * r * r;
    double pm = m->density * pV;
    double t0 = m->E_bond * bS / (2.0 * r);
    double t1 = d->G_bond * bS / (2.0 * r);
    d->T_bond = 2.0 * sqrt(pm / (t0 > t1 ? t0 : t1));
    if (m->dt > 0.0)
        d->dt = m->dt;
    else
        d->dt = (d->T_rayleigh < d->T_bond) ? 0.05 * d->T_rayleigh : 0.05 * d->T_bond;
    return RF_OK;
}

/* sample_data::SetSpaceSizeSi / CalculateParameters (md_data_types.h:87-124)
 * and cell_data::CalculateParameters (:42-55), in FP64. */
int ora_box_reference(const double space[3], double cell, rf_box* b) {
    if (!space || !b || !(cell > 0.0)) return RF_ERR_ARG;
    double A = 1.01 * cell;
    double B[3];
    for (int k = 0; k < 3; ++k) {
        B[k] = A + space[k];
        b->center[k] = 0.5 * (A + B[k]);
        b->half[k] = 0.5 * space[k];
    }
    b->hidden[0] = A + 0.5 * space[0];
    b->hidden[1] = A + 0.5 * space[1];
    b->hidden[2] = (floor(B[2] / cell) + 3 - 0.5) * cell;
    b->cell = cell;
    double inv = 1.0 / cell;
    b->dims[0] = (int32_t)floor(B[0] * inv) + 1;
    b->dims[1] = (int32_t)floor(B[1] * inv) + 1;
    b->dims[2] = (int32_t)floor(B[2] * inv) + 4;
    b->pad_ = 0;
    return RF_OK;
}

/* ------------------------------------------------------- strengths ---- */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Lognormal draw keyed by the unordered pair (lo < hi) so both directed
 * slots of a bond get the same value.  mean*exp(s*z - s^2/2), s^2 = ln(1+cv^2),
 * z standard normal by Box-Muller.  cv == 0 gives the constant mean (the
 * reference's single m_Ec / m_Gc). */
double ora_bond_strength(uint64_t seed, int64_t lo, int64_t hi, int32_t stream, double mean,
                         double cv) {
    if (!(cv > 0.0)) return mean;
    uint64_t base = seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(lo + 1)) ^
                    (0xC2B2AE3D27D4EB4FULL * (uint64_t)(hi + 1)) ^
                    (0x165667B19E3779F9ULL * (uint64_t)(stream + 1));
    uint64_t h1 = mix64(base);
    uint64_t h2 = mix64(h1 ^ 0xD6E8FEB86659FD93ULL);
    double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double z = sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
    double s2 = log(1.0 + cv * cv);
    double s = sqrt(s2);
    return mean * exp(s * z - 0.5 * s2);
}

/* ---------------------------------------------------------- create ---- */
static void* zalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* Order of the platen-force sum.  The device stores particles sorted by the
 * Morton code of their quarter-cell coordinates q = floor(4 x / cell) (21 bits
 * each; bit 3b of the code is bit b of qx, 3b+1 of qy, 3b+2 of qz), ties by
 * id (RF_ORDER_SPATIAL), or in id order (RF_ORDER_GIVEN), and sums the
 * per-particle platen forces over blocks of 32 stored particles.  Summing
 * over the same blocks is the only place where the storage order reaches a
 * result.  Pinned by a numpy restatement (tests/test_oracle_lists.py). */
static uint64_t spread_bits3(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

static uint64_t quarter(double x, double s) {
    double q = floor(x * s);
    if (!(q > 0.0)) return 0;
    return q >= 2097151.0 ? 2097151u : (uint64_t)q;
}

typedef struct { uint64_t key; int64_t id; } order_key;

static int cmp_order(const void* a, const void* b) {
    const order_key* x = (const order_key*)a;
    const order_key* y = (const order_key*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

static void set_order(ora_ctx* c, int32_t mode) {
    const int64_t n = c->n;
    if (mode == RF_ORDER_GIVEN) {
        for (int64_t i = 0; i < n; ++i) c->ord[i] = i;
        return;
    }
    const double s = 4.0 * (1.0 / c->box.cell);
    order_key* k = (order_key*)malloc(n * sizeof(order_key));
    for (int64_t i = 0; i < n; ++i) {
        k[i].key = spread_bits3(quarter(c->R[i], s)) | (spread_bits3(quarter(c->R[n + i], s)) << 1) |
                   (spread_bits3(quarter(c->R[2 * n + i], s)) << 2);
        k[i].id = i;
    }
    qsort(k, n, sizeof(order_key), cmp_order);
    for (int64_t p = 0; p < n; ++p) c->ord[p] = k[p].id;
    free(k);
}

/* Storage order the device uses for this sample (default RF_ORDER_SPATIAL);
 * call before the first step, while R holds the sample's positions. */
int ora_set_order(ora_ctx* c, int32_t mode) {
    if (mode != RF_ORDER_SPATIAL && mode != RF_ORDER_GIVEN) return RF_ERR_ARG;
    set_order(c, mode);
    return RF_OK;
}

int ora_get_order(const ora_ctx* c, int64_t* out) {
    memcpy(out, c->ord, c->n * sizeof(int64_t));
    return RF_OK;
}

ora_ctx* ora_create(const rf_sample* s, const rf_material* mat, const rf_box* box, int32_t slots,
                    int32_t nthreads) {
    if (!s || !mat || !box || s->n <= 0 || slots <= 0) return NULL;
    ora_ctx* c = (ora_ctx*)zalloc(sizeof(ora_ctx));
    int64_t n = s->n;
    c->n = n;
    c->S = slots;
    c->nthreads = nthreads > 0 ? nthreads : 1;
    c->mat = *mat;
    c->box = *box;
    /* working grid: one more cell layer beyond the top x and y faces, so the
     * whole space box lies in cells 1..dims-2 (the reference's last cell
     * column can need a stencil cell past its grid); membership, hence every
     * list order, is unchanged */
    c->box.dims[0] += 1;
    c->box.dims[1] += 1;
    size_t n3 = (size_t)(3 * n) * sizeof(double);
    c->R = zalloc(n3); c->V = zalloc(n3); c->W = zalloc(n3); c->F = zalloc(n3); c->M = zalloc(n3);
    c->rad = zalloc(n * sizeof(double)); c->mass = zalloc(n * sizeof(double));
    c->dtdm = zalloc(n * sizeof(double)); c->dtdI = zalloc(n * sizeof(double));
    size_t ns = (size_t)n * slots;
    c->nbr = malloc(ns * sizeof(int32_t));
    for (size_t q = 0; q < ns; ++q) c->nbr[q] = -1;
    c->type = zalloc(ns);
    c->bidx = malloc(ns);
    memset(c->bidx, -1, ns);
    c->invL0 = zalloc(ns * sizeof(double)); c->sc = zalloc(ns * sizeof(double));
    c->tc = zalloc(ns * sizeof(double));
    c->pn = zalloc(3 * ns * sizeof(double)); c->ot = zalloc(3 * ns * sizeof(double));
    c->mn = zalloc(3 * ns * sizeof(double)); c->mt = zalloc(3 * ns * sizeof(double));
    c->key = zalloc(n * sizeof(int32_t)); c->sorted = zalloc(n * sizeof(int32_t));
    c->ncell = (int64_t)c->box.dims[0] * c->box.dims[1] * c->box.dims[2];
    c->cstart = zalloc(c->ncell * sizeof(int32_t)); c->ccount = zalloc(c->ncell * sizeof(int32_t));
    c->part_b = zalloc(n * sizeof(double)); c->part_t = zalloc(n * sizeof(double));
    c->ord = zalloc(n * sizeof(int64_t));
    c->nblk = (n + RED_BLOCK - 1) / RED_BLOCK;
    c->blk_b = zalloc(c->nblk * sizeof(double)); c->blk_t = zalloc(c->nblk * sizeof(double));
    double rmin = s->r[0], rmax = s->r[0];
    for (int64_t i = 0; i < n; ++i) {
        c->R[i] = s->x[i]; c->R[n + i] = s->y[i]; c->R[2 * n + i] = s->z[i];
        c->rad[i] = s->r[i];
        if (s->r[i] < rmin) rmin = s->r[i];
        if (s->r[i] > rmax) rmax = s->r[i];
    }
    ora_derive(mat, rmin, rmax, &c->der);
    double dt = c->der.dt;
    for (int64_t i = 0; i < n; ++i) {
        /* p_V, p_m, p_I, dt_d_m, dt_d_I: md_data_types.h:159-163, 185-186 */
        double r = c->rad[i];
        double pV = 4.0 * (1.0 / 3.0) * PI * r * r * r;
        double pm = mat->density * pV;
        double pI = 0.4 * pm * r * r;
        c->mass[i] = pm;
        c->dtdm[i] = (1.0 / pm) * dt;
        c->dtdI[i] = dt / pI;
    }
    c->acut2 = mat->acut * mat->acut;
    c->inv_a = 1.0 / box->cell;
    c->has_test = 0;
    set_order(c, RF_ORDER_SPATIAL);
    return c;
}

void ora_destroy(ora_ctx* c) {
    if (!c) return;
    free(c->R); free(c->V); free(c->W); free(c->F); free(c->M);
    free(c->rad); free(c->mass); free(c->dtdm); free(c->dtdI);
    free(c->nbr); free(c->type); free(c->bidx); free(c->invL0); free(c->sc); free(c->tc);
    free(c->pn); free(c->ot); free(c->mn); free(c->mt);
    free(c->key); free(c->sorted); free(c->cstart); free(c->ccount); free(c->ord);
    free(c->part_b); free(c->part_t); free(c->blk_b); free(c->blk_t);
    free(c->curve); free(c->bcount);
    free(c);
}

/* ----------------------------------------------------------- cells ---- */
/* d_CalculateCellIndex (md_calculate_cell.cu:13-41), stable sort by cell
 * (cub::DeviceRadixSort::SortPairs over identity ids, md_cell_distribution.cu:87)
 * and cell start/count (d_DetermineCellPointer, md_calculate_cell.cu:43-75).
 * Pinned by: stable counting sort equals any stable sort by key; tests check
 * against numpy's stable argsort.  Particles must lie in cells 1..dims-2 of
 * the working grid (the reference's plus one layer past the top x and y
 * faces) so the 8-cell stencil stays inside it; the reference reads past its
 * grid for particles in the upper half of its last x or y cell. */
static int cell_coords(const ora_ctx* c, double x, double y, double z, int32_t* ix, int32_t* iy,
                       int32_t* iz) {
    double fx = floor(x * c->inv_a), fy = floor(y * c->inv_a), fz = floor(z * c->inv_a);
    if (!(fx >= 1.0 && fx <= c->box.dims[0] - 2 && fy >= 1.0 && fy <= c->box.dims[1] - 2 &&
          fz >= 1.0 && fz <= c->box.dims[2] - 2))
        return 0;
    *ix = (int32_t)fx; *iy = (int32_t)fy; *iz = (int32_t)fz;
    return 1;
}

int ora_cells(ora_ctx* c) {
    int64_t n = c->n;
    int32_t nx = c->box.dims[0], ny = c->box.dims[1];
    for (int64_t i = 0; i < n; ++i) {
        int32_t ix, iy, iz;
        if (!cell_coords(c, c->R[i], c->R[n + i], c->R[2 * n + i], &ix, &iy, &iz)) {
            char m[160];
            snprintf(m, sizeof m, "particle %lld outside the cell grid", (long long)i);
            set_err(c, m);
            return RF_ERR_GRID;
        }
        c->key[i] = ix + iy * nx + iz * ny * nx;
    }
    memset(c->ccount, 0, c->ncell * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) c->ccount[c->key[i]]++;
    int32_t acc = 0;
    for (int64_t q = 0; q < c->ncell; ++q) {
        c->cstart[q] = acc;
        acc += c->ccount[q];
    }
    int32_t* fill = zalloc(c->ncell * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) {
        int32_t k = c->key[i];
        c->sorted[c->cstart[k] + fill[k]++] = (int32_t)i;
    }
    free(fill);
    return RF_OK;
}

int ora_get_cells(const ora_ctx* c, int32_t* key, int32_t* sorted, int32_t* start, int32_t* count,
                  int64_t ncell_cap) {
    if (key) memcpy(key, c->key, c->n * sizeof(int32_t));
    if (sorted) memcpy(sorted, c->sorted, c->n * sizeof(int32_t));
    if (ncell_cap < c->ncell) return RF_ERR_ARG;
    if (start) memcpy(start, c->cstart, c->ncell * sizeof(int32_t));
    if (count) memcpy(count, c->ccount, c->ncell * sizeof(int32_t));
    return RF_OK;
}

int64_t ora_num_cells(const ora_ctx* c) { return c->ncell; }

/* The 8 stencil cells in the reference's visiting order: own, +x, +y, +z,
 * xy, xz, yz, xyz, the side chosen by the sign of the offset from the cell
 * centre (md_interaction_list_construct.cu:77-117,
 * md_interaction_list_reconstructbpm.cu:96-134). */
static void stencil(const ora_ctx* c, int64_t i, int64_t cells[8]) {
    int64_t n = c->n;
    double x = c->R[i], y = c->R[n + i], z = c->R[2 * n + i];
    double a = c->box.cell;
    int64_t nx = c->box.dims[0], nxy = (int64_t)c->box.dims[0] * c->box.dims[1];
    double fx = floor(x * c->inv_a), fy = floor(y * c->inv_a), fz = floor(z * c->inv_a);
    int64_t base = (int64_t)fx + (int64_t)fy * nx + (int64_t)fz * nxy;
    int64_t sx = copysign(1.0, x - (fx + 0.5) * a) > 0 ? 1 : -1;
    int64_t sy = copysign(1.0, y - (fy + 0.5) * a) > 0 ? 1 : -1;
    int64_t sz = copysign(1.0, z - (fz + 0.5) * a) > 0 ? 1 : -1;
    cells[0] = base;
    cells[1] = base + sx;
    cells[2] = base + sy * nx;
    cells[3] = base + sz * nxy;
    cells[4] = base + sx + sy * nx;
    cells[5] = base + sx + sz * nxy;
    cells[6] = base + sy * nx + sz * nxy;
    cells[7] = base + sx + sy * nx + sz * nxy;
}

static double dist2(const ora_ctx* c, int64_t i, int64_t j) {
    int64_t n = c->n;
    double dx = c->R[j] - c->R[i];
    double dy = c->R[n + j] - c->R[n + i];
    double dz = c->R[2 * n + j] - c->R[2 * n + i];
    return dx * dx + dy * dy + dz * dz;
}

static void zero_hist(ora_ctx* c, int64_t q) {
    for (int k = 0; k < 3; ++k) {
        c->pn[3 * q + k] = 0.0; c->ot[3 * q + k] = 0.0;
        c->mn[3 * q + k] = 0.0; c->mt[3 * q + k] = 0.0;
    }
}

/* ------------------------------------------------------------- a4 ---- */
/* d_ConstructInteractionList + addlink (md_interaction_list_construct.cu:
 * 17-37, 53-130): rows filled contiguously from slot 0 with every j != i in
 * the stencil at rr < acut^2; rows start empty (memset 0xFF,
 * md_interaction_list_construction.cu:64).  A row longer than S is an error
 * here (the reference writes into the next particle's row).  Pinned by the
 * brute-force pair set {j : rr < acut^2} on small samples. */
int ora_construct_rows(ora_ctx* c) {
    int rc = ora_cells(c);
    if (rc) return rc;
    int64_t n = c->n, S = c->S;
    int64_t bad = -1;
    for (int64_t q = 0; q < n * S; ++q) c->nbr[q] = -1;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i) {
        int64_t cells[8];
        stencil(c, i, cells);
        int64_t cnt = 0;
        for (int s = 0; s < 8; ++s) {
            int64_t cc = cells[s];
            int32_t st = c->cstart[cc], ct = c->ccount[cc];
            for (int32_t p = 0; p < ct; ++p) {
                int64_t j = c->sorted[st + p];
                double rr = dist2(c, i, j);
                if (rr < c->acut2 && j != i) {
                    if (cnt >= S) {
#pragma omp critical
                        bad = i;
                        continue;
                    }
                    c->nbr[i * S + cnt] = (int32_t)j;
                    ++cnt;
                }
            }
        }
    }
    if (bad >= 0) {
        char m[160];
        snprintf(m, sizeof m, "interaction row overflow at particle %lld (slots=%d)", (long long)bad,
                 c->S);
        set_err(c, m);
        return RF_ERR_OVERFLOW;
    }
    return RF_OK;
}

static double bond_radius(const ora_ctx* c, int64_t i, int64_t j) {
    if (c->mat.bond_radius_ratio > 0.0) {
        double ri = c->rad[i], rj = c->rad[j];
        return c->mat.bond_radius_ratio * (ri < rj ? ri : rj);
    }
    return c->mat.bond_radius;
}

/* ------------------------------------------------------------- a5 ---- */
/* d_ConstructBoundInteractions (md_boundinteractions_construct.cu:35-106):
 * type 1 if 1e-18 < rho^2 < b_d^2 with b_d = r_i + r_j + gap (reference:
 * 2 p_r + gap, mp_determineconjgradient.cu:25); 1/L0 = 1/sqrt(rho^2); every
 * history of an occupied slot zeroed, the previous normal included (A7).
 * Each particle's bonds are numbered 0,1,2,... in ascending slot order: that
 * number is the bond's position in the per-particle contribution tree.
 * Pinned by the closed form rho < r_i + r_j + gap and L0 = rho. */
int ora_construct_bonds(ora_ctx* c) {
    int64_t n = c->n, S = c->S;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i) {
        int nb = 0;
        for (int64_t k = 0; k < S; ++k) {
            int64_t q = i * S + k;
            int64_t j = c->nbr[q];
            c->bidx[q] = -1;
            if (j < 0) continue;
            double rm2 = dist2(c, i, j);
            double bd = c->rad[i] + c->rad[j] + c->mat.bond_gap;
            if (rm2 < bd * bd && rm2 > 1e-18) {
                c->invL0[q] = 1.0 / sqrt(rm2);
                c->type[q] = 1;
                c->bidx[q] = (int8_t)nb++;
            } else {
                c->type[q] = 0;
            }
            zero_hist(c, q);
        }
    }
    return RF_OK;
}

int ora_assign_strengths(ora_ctx* c) {
    int64_t n = c->n, S = c->S;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < S; ++k) {
            int64_t q = i * S + k;
            int64_t j = c->nbr[q];
            if (j < 0 || c->type[q] != 1) {
                c->sc[q] = 0.0;
                c->tc[q] = 0.0;
                continue;
            }
            int64_t lo = i < j ? i : j, hi = i < j ? j : i;
            c->sc[q] = ora_bond_strength(c->mat.strength_seed, lo, hi, 0, c->mat.sigma_c,
                                         c->mat.sigma_c_cv);
            c->tc[q] = ora_bond_strength(c->mat.strength_seed, lo, hi, 1, c->mat.tau_c,
                                         c->mat.tau_c_cv);
        }
    return RF_OK;
}

/* ------------------------------------------------------------- a7 ---- */
/* d_DeleteFarLinks (md_interaction_list_reconstructbpm.cu:141-186): slots
 * with rr > acut^2 or rr < 1e-10 emptied, type 0, histories zeroed; 1/L0 is
 * left as it was (A5).  Pinned by the brute-force pair set. */
int ora_prune(ora_ctx* c) {
    int64_t n = c->n, S = c->S;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < S; ++k) {
            int64_t q = i * S + k;
            int64_t j = c->nbr[q];
            if (j < 0) continue;
            double rr = dist2(c, i, j);
            if (rr > c->acut2 || rr < 1e-10) {
                c->nbr[q] = -1;
                c->type[q] = 0;
                zero_hist(c, q);
            }
        }
    return RF_OK;
}

/* ------------------------------------------------------------- a6 ---- */
/* d_ReConstructInteractionListbpm + addnewlinkbpm
 * (md_interaction_list_reconstructbpm.cu:13-46, 80-139): every stencil
 * candidate with 1e-10 < rr < acut^2 not yet in the row goes, as a contact
 * slot, into the first empty slot.  With RF_QUIRK_SLOT0_SENTINEL, particle 0
 * reproduces the reference's use of 0 as "no free slot yet": it takes the
 * first empty slot after slot 0 and slot 0 only when that is the only one
 * (:28-33).  A full row is an error here (the reference overwrites IL[0]).
 * Pinned by the brute-force pair set and a hand-built free-slot case. */
int ora_append(ora_ctx* c) {
    int rc = ora_cells(c);
    if (rc) return rc;
    int64_t n = c->n, S = c->S;
    int64_t bad = -1;
    int quirk0 = (c->mat.quirks & RF_QUIRK_SLOT0_SENTINEL) != 0;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i) {
        int64_t cells[8];
        stencil(c, i, cells);
        int32_t* row = c->nbr + i * S;
        for (int s = 0; s < 8; ++s) {
            int64_t cc = cells[s];
            int32_t st = c->cstart[cc], ct = c->ccount[cc];
            for (int32_t p = 0; p < ct; ++p) {
                int64_t j = c->sorted[st + p];
                double rr = dist2(c, i, j);
                if (!(rr < c->acut2 && rr > 1e-10)) continue;
                int64_t free_k = -1, k;
                for (k = 0; k < S; ++k) {
                    if (row[k] == (int32_t)j) break;
                    if (row[k] < 0) {
                        if (free_k < 0) free_k = k;
                        else if (quirk0 && i == 0 && free_k == 0) free_k = k;
                    }
                }
                if (k < S) continue; /* already present */
                if (free_k < 0) {
#pragma omp critical
                    bad = i;
                    continue;
                }
                row[free_k] = (int32_t)j;
                c->type[i * S + free_k] = 0;
            }
        }
    }
    if (bad >= 0) {
        char m[160];
        snprintf(m, sizeof m, "interaction row overflow at particle %lld (slots=%d)", (long long)bad,
                 c->S);
        set_err(c, m);
        return RF_ERR_OVERFLOW;
    }
    return RF_OK;
}

/* -------------------------------------------------- summation ---- */
/* Per-particle sums of slot contributions.  The reference accumulates them
 * one slot after another (fs += ... in slot order).  Here, as on the device
 * where 8 lanes share a particle, entry e (a bond number or a slot) belongs
 * to lane e % 8; each lane adds its entries in ascending order starting from
 * 0.0, and the 8 lane sums are added by the tree v[t] = v[t] + v[t + off],
 * off = 4, 2, 1.  Unused entries are skipped.  The order differs from the
 * sequential sum only in rounding. */
static double lane_sum(const double* c, const unsigned char* used, int nent) {
    double v[LANES];
    for (int l = 0; l < LANES; ++l) {
        v[l] = 0.0;
        for (int e = l; e < nent; e += LANES)
            if (used[e]) v[l] = v[l] + c[e];
    }
    for (int off = LANES / 2; off >= 1; off >>= 1)
        for (int t = 0; t < off; ++t) v[t] = v[t] + v[t + off];
    return v[0];
}

/* ------------------------------------------------------------- a8 ---- */
/* Parallel bond, d_CalculateForcesDEM_22 with dd_Calculate_rijm_nij_2,
 * dd_Calculate_fijn_bpm_2, dd_Calculate_M123_bpm_2, dd_Calculate_mijn_bpm_2,
 * dd_Calculate_mijt_bpm_2, dd_Calculate_fijt_mijunsym_bpm_2 and
 * dd_CheckBreakConditionDEM_2 (md_calculateforcesdem3.cu:20-200, 427-486).
 * Bond contribution to particle i (summed by bond number):
 *   F: n*(E(rho-L0)/L0 * A) + o_t*(G/L0 * A)
 *   M: (m_n*J4 + m_t*c2J) + (0.5 r x o_t)*(G/L0 * A),  c2J = dt E/(4 L0) J4 (A1)
 * Pinned by two-particle closed forms (axial, shear, torsion, bending with
 * the A1 factor, break threshold) and F_ij = -F_ji. */
static void bond_row(ora_ctx* c, int64_t i, double fs[3], double ms[3]) {
    const int64_t n = c->n, S = c->S;
    const double* R = c->R; const double* V = c->V; const double* W = c->W;
    const double dt = c->der.dt, E = c->mat.E_bond, G = c->der.G_bond;
    const int a1 = (c->mat.quirks & RF_QUIRK_BEND_DOUBLE_SCALE) != 0;
    const double xi = R[i], yi = R[n + i], zi = R[2 * n + i];
    const double wix = W[i], wiy = W[n + i], wiz = W[2 * n + i];
    const double vix = V[i], viy = V[n + i], viz = V[2 * n + i];
    double cf[3][32], cm[3][32];
    unsigned char used[32];
    memset(used, 0, sizeof used);
    int nent = 0;
    for (int64_t k = 0; k < S; ++k) {
        int64_t q = i * S + k;
        int64_t j = c->nbr[q];
        if (j < 0 || c->type[q] != 1) continue;
        int bi = c->bidx[q];
        used[bi] = 1;
        if (bi + 1 > nent) nent = bi + 1;
        double rx = R[j] - xi, ry = R[n + j] - yi, rz = R[2 * n + j] - zi;
        double rijm = sqrt(rx * rx + ry * ry + rz * rz);
        double inv = 1.0 / rijm;
        double nx = rx * inv, ny = ry * inv, nz = rz * inv;
        double lam = c->invL0[q];
        double b = bond_radius(c, i, j);
        /* normal force */
        double cefn = E * (rijm - 1.0 / lam) * lam;
        double Sc = cefn;
        cefn *= b * b * PI;
        /* rotation matrix M(t = prev x n, phi) */
        double* pn = c->pn + 3 * q;
        double tx = pn[1] * nz - pn[2] * ny;
        double ty = pn[2] * nx - pn[0] * nz;
        double tz = pn[0] * ny - pn[1] * nx;
        double wjx = W[j], wjy = W[n + j], wjz = W[2 * n + j];
        double cw = ((wjx + wix) * nx + (wjy + wiy) * ny + (wjz + wiz) * nz) * dt * 0.5;
        double px = nx * cw, py = ny * cw, pz = nz * cw;
        pn[0] = nx; pn[1] = ny; pn[2] = nz;
        double M1x = 1.0 + tz * pz + ty * py, M1y = pz - tz - ty * px, M1z = -py - tz * px + ty;
        double M2x = tz - pz - tx * py, M2y = tz * pz + 1.0 + tx * px, M2z = -tz * py + px - tx;
        double M3x = -ty - tx * pz + py, M3y = -ty * pz + tx - px, M3z = ty * py + tx * px + 1.0;
        /* torsion */
        double dwx = wjx - wix, dwy = wjy - wiy, dwz = wjz - wiz;
        double c1 = nx * dwx + ny * dwy + nz * dwz;
        c1 *= dt * 0.5 * G * lam;
        double* mn = c->mn + 3 * q;
        double mnx = M1x * mn[0] + M1y * mn[1] + M1z * mn[2] + nx * c1;
        double mny = M2x * mn[0] + M2y * mn[1] + M2z * mn[2] + ny * c1;
        double mnz = M3x * mn[0] + M3y * mn[1] + M3z * mn[2] + nz * c1;
        mn[0] = mnx; mn[1] = mny; mn[2] = mnz;
        double J4 = PI * b * b * b * b;
        /* bending */
        double c2 = dt * 0.25 * E * lam;
        double* mt = c->mt + 3 * q;
        double c1b = nx * dwx + ny * dwy + nz * dwz;
        double mtx = M1x * mt[0] + M1y * mt[1] + M1z * mt[2] + (dwx - nx * c1b) * c2;
        double mty = M2x * mt[0] + M2y * mt[1] + M2z * mt[2] + (dwy - ny * c1b) * c2;
        double mtz = M3x * mt[0] + M3y * mt[1] + M3z * mt[2] + (dwz - nz * c1b) * c2;
        mt[0] = mtx; mt[1] = mty; mt[2] = mtz;
        double c2J = a1 ? c2 * J4 : J4;
        /* shear force and asymmetric moment */
        double hx = 0.5 * (wjx + wix), hy = 0.5 * (wjy + wiy), hz = 0.5 * (wjz + wiz);
        double vx = V[j] - vix - hy * rz + hz * ry;
        double vy = V[n + j] - viy - hz * rx + hx * rz;
        double vz = V[2 * n + j] - viz - hx * ry + hy * rx;
        double cn = nx * vx + ny * vy + nz * vz;
        double* ot = c->ot + 3 * q;
        double ox = (M1x * ot[0] + M1y * ot[1] + M1z * ot[2]) + (vx - nx * cn) * dt;
        double oy = (M2x * ot[0] + M2y * ot[1] + M2z * ot[2]) + (vy - ny * cn) * dt;
        double oz = (M3x * ot[0] + M3y * ot[1] + M3z * ot[2]) + (vz - nz * cn) * dt;
        double cef = G * lam;
        double Mc = cef;
        cef *= b * b * PI;
        ot[0] = ox; ot[1] = oy; ot[2] = oz;
        cf[0][bi] = nx * cefn + ox * cef;
        cf[1][bi] = ny * cefn + oy * cef;
        cf[2][bi] = nz * cefn + oz * cef;
        cm[0][bi] = (mnx * J4 + mtx * c2J) + 0.5 * (ry * oz - rz * oy) * cef;
        cm[1][bi] = (mny * J4 + mty * c2J) + 0.5 * (rz * ox - rx * oz) * cef;
        cm[2][bi] = (mnz * J4 + mtz * c2J) + 0.5 * (rx * oy - ry * ox) * cef;
        /* breakage */
        Mc *= sqrt(ox * ox + oy * oy + oz * oz);
        Sc += sqrt(mtx * mtx + mty * mty + mtz * mtz) * 4.0 * b;
        Mc += sqrt(mnx * mnx + mny * mny + mnz * mnz) * 2.0 * b;
        if (Sc > c->sc[q] || Mc > c->tc[q]) {
            c->type[q] = 0;
            ot[0] = 0.0; ot[1] = 0.0; ot[2] = 0.0;
            mn[0] = 0.0; mn[1] = 0.0; mn[2] = 0.0;
            mt[0] = 0.0; mt[1] = 0.0; mt[2] = 0.0;
        }
    }
    for (int k = 0; k < 3; ++k) {
        fs[k] = lane_sum(cf[k], used, nent);
        ms[k] = lane_sum(cm[k], used, nent);
    }
}

/* ------------------------------------------------------------- a9 ---- */
/* Hertz-Mindlin contact, d_CalculateForcesDEM_21 with dd_Calculate_fijn_ppHM_2,
 * dd_Calculate_fijtmijt_ppHM_2, dd_Calculate_mijroll_ppHM_2
 * (md_calculateforcesdem3.cu:202-339, 341-425).  Contribution of slot k:
 *   F: n*(Fn + damping) + Ft,   M: (n x Ft)*r_i [+ rolling W_i*(-mu_r|Fn|/|W_i|)]
 * Rolling torque is mu_r*|Fn| with no radius factor, as in the live code
 * (:328).  With RF_QUIRK_ROLL_RESET a particle with |W_i|^2 <= 1e-18 ends with
 * a zero contact-moment sum: the reference clears its running sum at every
 * overlapping contact (A2, :333-338), so the last clear wins.
 * Pinned by Hertz |F| = 4/3 E* sqrt(R*) d^1.5, Mindlin Kt = 8 G* sqrt(R* d),
 * the Coulomb cap and the head-on restitution. */
static void contact_row(ora_ctx* c, int64_t i, double fs[3], double ms[3]) {
    const int64_t n = c->n, S = c->S;
    const double* R = c->R; const double* V = c->V; const double* W = c->W;
    const double dt = c->der.dt, hmE = c->der.hm_E, hmG = c->der.hm_G, pA = c->der.p_A;
    const double mu = c->mat.friction, mur = c->mat.rolling_friction;
    const int a2 = (c->mat.quirks & RF_QUIRK_ROLL_RESET) != 0;
    const double xi = R[i], yi = R[n + i], zi = R[2 * n + i];
    const double wix = W[i], wiy = W[n + i], wiz = W[2 * n + i];
    const double vix = V[i], viy = V[n + i], viz = V[2 * n + i];
    const double ri = c->rad[i], mi = c->mass[i];
    const double wim = wix * wix + wiy * wiy + wiz * wiz;
    const int roll = wim > 1e-18;
    double cf[3][32], cm[3][32];
    unsigned char used[32];
    memset(used, 0, sizeof used);
    for (int64_t k = 0; k < S; ++k) {
        int64_t q = i * S + k;
        int64_t j = c->nbr[q];
        double* ot = c->ot + 3 * q;
        if (j >= 0 && c->type[q] == 0) {
            double rx = R[j] - xi, ry = R[n + j] - yi, rz = R[2 * n + j] - zi;
            double rijm = sqrt(rx * rx + ry * ry + rz * rz);
            double inv = 1.0 / rijm;
            double nx = rx * inv, ny = ry * inv, nz = rz * inv;
            double rj = c->rad[j];
            double rsum = ri + rj;
            if (rijm < rsum) {
                double mj = c->mass[j];
                double Rst = ri * rj / (ri + rj);
                double meff = mi * mj / (mi + mj);
                double c1 = sqrt(Rst * fabs(rsum - rijm));
                double Kn = 2.0 * hmE * c1;
                double Kt = 8.0 * hmG * c1;
                double wsx = rj * W[j] + ri * wix;
                double wsy = rj * W[n + j] + ri * wiy;
                double wsz = rj * W[2 * n + j] + ri * wiz;
                double vx = V[j] - vix + (ny * wsz - nz * wsy);
                double vy = V[n + j] - viy + (nz * wsx - nx * wsz);
                double vz = V[2 * n + j] - viz + (nx * wsy - ny * wsx);
                double vn = nx * vx + ny * vy + nz * vz;
                double fn = (rijm - rsum) * Kn * TWO_THIRDS;
                double fd = DAMP * pA * vn * sqrt(Kn * meff);
                double vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
                /* tangential history: project, rescale, increment, cap */
                c1 = nx * ot[0] + ny * ot[1] + nz * ot[2];
                double ox = ot[0] - nx * c1, oy = ot[1] - ny * c1, oz = ot[2] - nz * c1;
                c1 = ox * ox + oy * oy + oz * oz;
                if (c1 > 1e-18) {
                    c1 = 1.0 / c1;
                    c1 *= ot[0] * ot[0] + ot[1] * ot[1] + ot[2] * ot[2];
                    c1 = sqrt(c1);
                    ox *= c1; oy *= c1; oz *= c1;
                }
                ox += vtx * dt; oy += vty * dt; oz += vtz * dt;
                double ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
                double ftm = ftx * ftx + fty * fty + ftz * ftz;
                if (ftm > mu * mu * fn * fn) {
                    c1 = mu * fabs(fn) * (1.0 / sqrt(ftm));
                    ftx *= c1; fty *= c1; ftz *= c1;
                    c1 = 1.0 / Kt;
                    ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
                } else {
                    c1 = DAMP * pA * sqrt(Kt * meff);
                    ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
                }
                ot[0] = ox; ot[1] = oy; ot[2] = oz;
                used[k] = 1;
                cf[0][k] = nx * (fn + fd) + ftx;
                cf[1][k] = ny * (fn + fd) + fty;
                cf[2][k] = nz * (fn + fd) + ftz;
                double tqx = (ny * ftz - nz * fty) * ri;
                double tqy = (nz * ftx - nx * ftz) * ri;
                double tqz = (nx * fty - ny * ftx) * ri;
                if (roll) {
                    double cr = -mur * fabs(fn) * (1.0 / sqrt(wim));
                    cm[0][k] = tqx + wix * cr;
                    cm[1][k] = tqy + wiy * cr;
                    cm[2][k] = tqz + wiz * cr;
                } else {
                    cm[0][k] = tqx; cm[1][k] = tqy; cm[2][k] = tqz;
                }
            } else {
                ot[0] = 0.0; ot[1] = 0.0; ot[2] = 0.0;
            }
        } else if (j < 0) {
            ot[0] = 0.0; ot[1] = 0.0; ot[2] = 0.0;
        }
    }
    for (int k = 0; k < 3; ++k) {
        fs[k] = lane_sum(cf[k], used, (int)S);
        ms[k] = lane_sum(cm[k], used, (int)S);
    }
    if (a2 && !roll) {
        ms[0] = 0.0; ms[1] = 0.0; ms[2] = 0.0;
    }
}

/* ------------------------------------------------------------ a10 ---- */
/* Particle-platen Hertz-Mindlin, dd_Calculate_fijn_pwHM_2 and
 * dd_Calculate_fijtmijt_pwHM_2 (md_compression_hm.cu:20-154): R* = r_i,
 * m* = m_i, previous tangential overlap `op` (which the reference's shared
 * Oiwt makes zero for every single-platen contact, A3), including the
 * reference's rescale typo o.x*o.y (:91), reachable only when a particle
 * touches both platens.  Pinned by the single-sphere-on-plane Hertz law. */
typedef struct wall_out {
    double fn[3], ft[3], mt[3], madd[3], o[3];
} wall_out;

static void wall_contact(const ora_ctx* c, int64_t i, const double nw[3], const double wv[3],
                         double d, double Ew, double Gw, const double op[3], wall_out* o) {
    const int64_t n = c->n;
    const double* V = c->V; const double* W = c->W;
    const double dt = c->der.dt, pA = c->der.p_A, mu = c->mat.friction,
                 mur = c->mat.rolling_friction;
    const double ri = c->rad[i], mi = c->mass[i];
    const double wix = W[i], wiy = W[n + i], wiz = W[2 * n + i];
    double c1 = sqrt(ri * fabs(ri - d));
    double Kn = 2.0 * Ew * c1;
    double Kt = 8.0 * Gw * c1;
    double vx = wv[0] - V[i] + ri * (nw[1] * wiz - nw[2] * wiy);
    double vy = wv[1] - V[n + i] + ri * (nw[2] * wix - nw[0] * wiz);
    double vz = wv[2] - V[2 * n + i] + ri * (nw[0] * wiy - nw[1] * wix);
    double vn = nw[0] * vx + nw[1] * vy + nw[2] * vz;
    double fnm = (d - ri) * Kn * TWO_THIRDS;
    c1 = DAMP * pA * vn * sqrt(Kn * mi);
    o->fn[0] = nw[0] * (fnm + c1); o->fn[1] = nw[1] * (fnm + c1); o->fn[2] = nw[2] * (fnm + c1);
    double vtx = vx - nw[0] * vn, vty = vy - nw[1] * vn, vtz = vz - nw[2] * vn;
    c1 = nw[0] * op[0] + nw[1] * op[1] + nw[2] * op[2];
    double ox = op[0] - nw[0] * c1, oy = op[1] - nw[1] * c1, oz = op[2] - nw[2] * c1;
    c1 = ox * ox + oy * oy + oz * oz;
    if (c1 > 1e-18) {
        c1 = 1.0 / c1;
        c1 *= op[0] * op[1] + op[1] * op[1] + op[2] * op[2];
        c1 = sqrt(c1);
        ox *= c1; oy *= c1; oz *= c1;
    }
    ox += vtx * dt; oy += vty * dt; oz += vtz * dt;
    double ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
    double ftm = ftx * ftx + fty * fty + ftz * ftz;
    if (ftm > mu * mu * fnm * fnm) {
        c1 = mu * fabs(fnm) * (1.0 / sqrt(ftm));
        ftx *= c1; fty *= c1; ftz *= c1;
        c1 = 1.0 / Kt;
        ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
    } else {
        c1 = DAMP * pA * sqrt(Kt * mi);
        ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
    }
    o->o[0] = ox; o->o[1] = oy; o->o[2] = oz;
    o->ft[0] = ftx; o->ft[1] = fty; o->ft[2] = ftz;
    o->mt[0] = (nw[1] * ftz - nw[2] * fty) * ri;
    o->mt[1] = (nw[2] * ftx - nw[0] * ftz) * ri;
    o->mt[2] = (nw[0] * fty - nw[1] * ftx) * ri;
    double wim = wix * wix + wiy * wiy + wiz * wiz;
    if (wim > 1e-18) {
        c1 = -mur * fabs(fnm) * (1.0 / sqrt(wim));
        o->madd[0] = wix * c1; o->madd[1] = wiy * c1; o->madd[2] = wiz * c1;
    } else {
        o->madd[0] = 0.0; o->madd[1] = 0.0; o->madd[2] = 0.0;
    }
}

/* d_UniaxialCompression_hm (md_compression_hm.cu:156-301; axis z, scope
 * (x-cx)^2+(y-cy)^2 < RR and z < Zcut with absolute z) and
 * d_BrazilCompression_hm (:304-449; axis y, scope (x-cx)^2+(z-cz)^2 < RR and
 * z-cz < Zcut).  Adds the wall forces to F, M and returns the per-particle
 * normal-force terms that the partial sums reduce. */
static void platen_particle(ora_ctx* c, int64_t i, double* sb, double* st) {
    const int64_t n = c->n;
    const int kind = c->cfg.kind;
    const int ax = kind == RF_TEST_BRAZILIAN ? 1 : 2;
    double x = c->R[i], y = c->R[n + i], z = c->R[2 * n + i];
    double s_b = 0.0, s_t = 0.0;
    double rx, ry, rz, r2;
    if (kind == RF_TEST_BRAZILIAN) {
        rx = x - c->scope_c[0]; ry = y; rz = z - c->scope_c[2];
        r2 = rx * rx + rz * rz;
    } else {
        rx = x - c->scope_c[0]; ry = y - c->scope_c[1]; rz = z;
        r2 = rx * rx + ry * ry;
    }
    if (r2 < c->rr_scope && rz < c->zcut) {
        double pos = ax == 2 ? rz : ry;
        double Ew = c->cfg.wall_modulus_factor * c->der.hm_E;
        double Gw = c->cfg.wall_modulus_factor * c->der.hm_G;
        double o_state[3] = {0.0, 0.0, 0.0};
        double ri = c->rad[i];
        wall_out wo;
        if (c->top - pos < ri) {
            double nw[3] = {0.0, 0.0, 0.0}, wv[3] = {0.0, 0.0, 0.0};
            nw[ax] = 1.0; wv[ax] = c->vw;
            wall_contact(c, i, nw, wv, fabs(c->top - pos), Ew, Gw, o_state, &wo);
            s_t += wo.fn[ax];
            for (int k = 0; k < 3; ++k) {
                c->F[k * n + i] += wo.fn[k] + wo.ft[k];
                c->M[k * n + i] += wo.mt[k] + wo.madd[k];
                o_state[k] = wo.o[k];
            }
        } else {
            o_state[0] = 0.0; o_state[1] = 0.0; o_state[2] = 0.0;
        }
        if (pos - c->bottom < ri) {
            double nw[3] = {0.0, 0.0, 0.0}, wv[3] = {0.0, 0.0, 0.0};
            nw[ax] = -1.0; wv[ax] = -c->vw;
            wall_contact(c, i, nw, wv, fabs(c->bottom - pos), Ew, Gw, o_state, &wo);
            s_b += wo.fn[ax];
            for (int k = 0; k < 3; ++k) {
                c->F[k * n + i] += wo.fn[k] + wo.ft[k];
                c->M[k * n + i] += wo.mt[k] + wo.madd[k];
            }
        }
    }
    *sb = -s_b;
    *st = -s_t;
}

/* Deterministic platen-force sum, the same order as the device: per group
 * of RED_BLOCK = 32 consecutive particles (-0.0 for indices past n, as the
 * reference's idle threads store -sfb = -0.0) each run of 8 by a pairwise
 * tree (v[t] += v[t+off], off = 4, 2, 1) and the four runs as
 * (W0 + W2) + (W1 + W3), then CTRL_T = 1024 strided running sums over the
 * group partials and a 1024-wide tree.  The reference's FP32 64-thread tree and host
 * loop (md_compression_hm.cu:243-300, md_results.cu:56-69) sum the same terms
 * in another order. */
static double tree_n(double* v, int w) {
    for (int off = w / 2; off >= 1; off >>= 1)
        for (int t = 0; t < off; ++t) v[t] = v[t] + v[t + off];
    return v[0];
}

static double ctrl_sum(const double* blk, int64_t nblk) {
    double buf[CTRL_T];
    for (int t = 0; t < CTRL_T; ++t) {
        double acc = 0.0;
        for (int64_t b = t; b < nblk; b += CTRL_T) acc = acc + blk[b];
        buf[t] = acc;
    }
    return tree_n(buf, CTRL_T);
}

/* A group of RED_BLOCK = 32 particles in storage order (-0.0 past the
 * last): each run of 8 by the tree v[k] += v[k+off], off = 4, 2, 1, then the
 * four runs as (W0 + W2) + (W1 + W3). */
static double group_tree(double* v) {
    double w[4];
    for (int r = 0; r < 4; ++r) w[r] = tree_n(v + 8 * r, 8);
    return (w[0] + w[2]) + (w[1] + w[3]);
}

static void reduce_partials(ora_ctx* c) {
    double buf[RED_BLOCK];
    for (int64_t b = 0; b < c->nblk; ++b) {
        for (int t = 0; t < RED_BLOCK; ++t) {
            int64_t i = b * RED_BLOCK + t;
            buf[t] = i < c->n ? c->part_b[c->ord[i]] : -0.0;
        }
        c->blk_b[b] = group_tree(buf);
        for (int t = 0; t < RED_BLOCK; ++t) {
            int64_t i = b * RED_BLOCK + t;
            buf[t] = i < c->n ? c->part_t[c->ord[i]] : -0.0;
        }
        c->blk_t[b] = group_tree(buf);
    }
    c->sfl0 = ctrl_sum(c->blk_b, c->nblk);
    c->sfl1 = ctrl_sum(c->blk_t, c->nblk);
}

/* The last ora_forces' per-particle platen forces (bottom, top), by the
 * caller's particle id. */
int ora_get_particle_partials(const ora_ctx* c, double* bottom, double* top, int64_t cap) {
    if (cap < c->n) return RF_ERR_ARG;
    memcpy(bottom, c->part_b, c->n * sizeof(double));
    memcpy(top, c->part_t, c->n * sizeof(double));
    return RF_OK;
}

int ora_get_partials(const ora_ctx* c, double* bottom, double* top, int64_t cap) {
    if (cap < c->nblk) return RF_ERR_ARG;
    memcpy(bottom, c->blk_b, c->nblk * sizeof(double));
    memcpy(top, c->blk_t, c->nblk * sizeof(double));
    return RF_OK;
}

/* F = M = 0 (md_calculationuniaxialcompression.cu:267-268), bonds (:269),
 * contacts (:271), platens (:277).  F[i] = ((0 + bond sum) + contact sum) +
 * top + bottom, as the three kernels' `+=` order. */
int ora_forces(ora_ctx* c) {
    const int64_t n = c->n;
    const int platen = c->has_test && c->cfg.kind != RF_TEST_NONE;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double fb[3] = {0.0, 0.0, 0.0}, mb[3] = {0.0, 0.0, 0.0};
        bond_row(c, i, fb, mb);
        double fc[3] = {0.0, 0.0, 0.0}, mc[3] = {0.0, 0.0, 0.0};
        contact_row(c, i, fc, mc);
        for (int k = 0; k < 3; ++k) {
            double f = 0.0, m = 0.0;
            f += fb[k]; m += mb[k];
            f += fc[k]; m += mc[k];
            c->F[k * n + i] = f;
            c->M[k * n + i] = m;
        }
        if (platen)
            platen_particle(c, i, &c->part_b[i], &c->part_t[i]);
        else {
            c->part_b[i] = -0.0;
            c->part_t[i] = -0.0;
        }
    }
    reduce_partials(c);
    return RF_OK;
}

/* ------------------------------------------------------ a11 / a12 ---- */
/* d_CalculateIncrementsDEM (md_calculateincrementsdem.cu:19-49) and
 * d_ParallelepipedCutRestriction (md_parallelepipedrestriction.cu:73-160).
 * Pinned by exact kick-drift under constant force. */
int ora_integrate(ora_ctx* c) {
    const int64_t n = c->n;
    const double dt = c->der.dt;
    const double* cen = c->box.center; const double* hf = c->box.half;
    const double* hp = c->box.hidden;
#pragma omp parallel for schedule(static) num_threads(c->nthreads)
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            c->V[k * n + i] += c->F[k * n + i] * c->dtdm[i];
            c->R[k * n + i] += c->V[k * n + i] * dt;
            c->W[k * n + i] += c->M[k * n + i] * c->dtdI[i];
        }
        int out = 0;
        for (int k = 0; k < 3; ++k) {
            double d = c->R[k * n + i] - cen[k];
            if (d - hf[k] > 0.0) out = 1;
            if (d + hf[k] < 0.0) out = 1;
        }
        if (out)
            for (int k = 0; k < 3; ++k) {
                c->R[k * n + i] = hp[k];
                c->V[k * n + i] = 0.0;
                c->F[k * n + i] = 0.0;
            }
    }
    return RF_OK;
}

/* -------------------------------------------------------- a2, a13 ---- */
/* compression_data::CalculateUParameters / CalculateBParameters
 * (md_data_types.h:212-242, 380-409) and ResultsUInit (md_results.cu:16-29). */
int ora_set_test(ora_ctx* c, const rf_platen_cfg* cfg) {
    c->cfg = *cfg;
    c->has_test = 1;
    int ax = cfg->kind == RF_TEST_BRAZILIAN ? 1 : 2;
    double sz = cfg->size[ax];
    c->size_axis = sz;
    c->size0 = sz;
    c->bottom = cfg->center[ax] - 0.5 * sz;
    c->top = cfg->center[ax] + 0.5 * sz;
    for (int k = 0; k < 3; ++k) c->scope_c[k] = cfg->center[k];
    double hx = 0.5 * cfg->size[0];
    if (cfg->kind == RF_TEST_BRAZILIAN) {
        double hz = 0.5 * cfg->size[2];
        double Rs = (hx > hz) ? hx : 1.05 * hz;
        c->rr_scope = Rs * Rs;
        c->area = 0.5 * 2 * PI * hx * cfg->size[2];
        c->zero_stress = 0.5 * cfg->min_stress;
    } else {
        double Rs = 1.5 * hx;
        c->rr_scope = Rs * Rs;
        c->area = PI * (hx + cfg->r_ref) * (hx + cfg->r_ref);
        c->zero_stress = 0.33 * cfg->min_stress;
    }
    c->inv_area = 1.0 / c->area;
    c->zcut = c->box.center[2] + c->box.half[2];
    c->vw = 0.0;
    c->strain = 0.0; c->stress = 0.0; c->peak_stress = 0.0; c->peak_strain = 0.0;
    c->sfl0 = 0.0; c->sfl1 = 0.0; c->ld = 0.0; c->destruction = 0.0;
    c->step = 0; c->steps_stop = 0; c->stepsave = 0;
    c->loading = 0; c->fracture = 0; c->done = 0;
    free(c->curve); free(c->bcount);
    int64_t cap = cfg->curve_cap;
    if (cap <= 0) {
        int64_t dn = cfg->dnsave > 0 ? cfg->dnsave : 1;
        double est = 1.5 * cfg->max_strain * sz / (fabs(cfg->v_load) * c->der.dt * dn);
        cap = (int64_t)est + 16;
        if (cfg->max_steps > 0 && cfg->max_steps / dn + 16 < cap) cap = cfg->max_steps / dn + 16;
    }
    c->curve_cap = cap;
    c->curve = zalloc((size_t)(3 * (cap + 2)) * sizeof(double));
    c->bcount = zalloc((size_t)(cap + 2) * sizeof(int64_t));
    return RF_OK;
}

/* SumInteractions + d_SumUpInteractions (md_results.cu:189-272). */
static int64_t count_bonds(const ora_ctx* c) {
    int64_t s = 0;
    for (int64_t q = 0; q < c->n * c->S; ++q)
        if (c->nbr[q] >= 0 && c->type[q] == 1) ++s;
    return s;
}

/* SumForcesULoading / SumForcesBLoading (md_results.cu:56-145). */
static void sum_forces(ora_ctx* c) {
    int64_t dn = c->cfg.dnsave > 0 ? c->cfg.dnsave : 1;
    if (c->step % dn == 0) {
        if (c->stepsave < c->curve_cap) {
            c->curve[3 * c->stepsave] = (c->size_axis - c->size0) / c->size0;
            c->curve[3 * c->stepsave + 1] = c->sfl0 * c->inv_area;
            c->curve[3 * c->stepsave + 2] = c->sfl1 * c->inv_area;
        }
        if (fabs(c->vw) > 1e-9) ++c->stepsave;
    }
    c->stress = 0.5 * (fabs(c->sfl0) + fabs(c->sfl1)) * c->inv_area;
    if (!c->loading && c->stress > c->cfg.min_stress) c->loading = 1;
    if (c->cfg.kind == RF_TEST_BRAZILIAN) {
        if (c->peak_stress < c->stress || c->stress < c->cfg.cond_fracture * c->peak_stress) {
            c->peak_stress = c->stress;
            c->peak_strain = (c->size0 - c->size_axis) / c->size0;
        }
        if (c->loading && c->stress < c->zero_stress) c->fracture = 1;
    } else {
        if (c->peak_stress < c->stress) {
            c->peak_stress = c->stress;
            c->peak_strain = (c->size0 - c->size_axis) / c->size0;
        }
        if (c->loading &&
            (c->stress < c->zero_stress || c->stress < c->cfg.cond_fracture * c->peak_stress))
            c->fracture = 1;
    }
}

/* CompressU_S3 / CompressB_S3 (md_data_types.h:347-379, 462-494).  Both
 * platens move by half the closing increment; the wall velocity seen by the
 * contact law is the full V (A4). */
static void controller(ora_ctx* c) {
    const double dt = c->der.dt;
    double pS = c->cfg.kind == RF_TEST_BRAZILIAN ? 0.5 * (-c->sfl0 + c->sfl1)
                                                 : 0.5 * fabs(-c->sfl0 + c->sfl1);
    double pD = fabs(c->sfl0 + c->sfl1);
    if (pS * c->inv_area < 1e-9) {
        c->vw = 10.0 * c->cfg.v_load;
    } else if (pD > 0.20 * pS && pS * c->inv_area > 2.0e+6 / 1e9) {
        if (fabs(c->vw) > 1e-15) c->steps_stop = 0;
        c->vw = 0;
        ++c->steps_stop;
        if (c->steps_stop > 30000) c->fracture = 1;
    } else {
        c->vw = c->cfg.v_load;
    }
    c->size_axis += c->vw * dt;
    c->top += 0.5 * c->vw * dt;
    c->bottom -= 0.5 * c->vw * dt;
    c->strain = (c->size_axis - c->size0) / c->size0;
}

/* One loading step in the order of md_calculationuniaxialcompression.cu:
 * 265-314 (Brazilian: md_calculationbraziltest.cu:185-231). */
static int one_step(ora_ctx* c) {
    int rc = ora_append(c);
    if (rc) return rc;
    ora_forces(c);
    int platen = c->has_test && c->cfg.kind != RF_TEST_NONE;
    if (platen) {
        int64_t dn = c->cfg.dnsave > 0 ? c->cfg.dnsave : 1;
        if (c->step % dn == 0 && c->stepsave < c->curve_cap) c->bcount[c->stepsave] = count_bonds(c);
        sum_forces(c);
    }
    ora_integrate(c);
    ora_prune(c);
    if (platen) controller(c);
    c->step++;
    return RF_OK;
}

static int loop_cond(const ora_ctx* c) {
    if (!c->has_test || c->cfg.kind == RF_TEST_NONE) return 1;
    if (c->cfg.max_steps > 0 && c->step >= c->cfg.max_steps) return 0;
    return fabs(c->strain) < c->cfg.max_strain && !c->fracture;
}

int64_t ora_step(ora_ctx* c, int64_t nsteps) {
    int64_t done = 0;
    while (done < nsteps) {
        if (!loop_cond(c)) {
            c->done = 1;
            break;
        }
        int rc = one_step(c);
        if (rc) return -rc;
        ++done;
    }
    if (!loop_cond(c)) c->done = 1;
    return done;
}

/* CalculateLinearityDeviation and CalculateDestruction (md_results.cu:147-187). */
int ora_post(ora_ctx* c) {
    if (c->cfg.kind == RF_TEST_UNIAXIAL) {
        double l_eps = 0, l_deps = 0, l_A = c->peak_stress / c->peak_strain, l_dS = 0, ld = 0;
        int64_t l_i = 0;
        const double* h = c->curve;
        do {
            ld += l_deps * l_dS * l_dS;
            ++l_i;
            if (l_i + 1 > c->curve_cap + 1) break;
            l_eps = -h[3 * l_i];
            l_deps = -0.5 * (h[3 * (l_i + 1)] - h[3 * (l_i - 1)]);
            l_dS = (l_A * l_eps - 0.5 * (h[3 * l_i + 2] - h[3 * l_i + 1]));
        } while (l_eps < c->peak_strain - 2 * l_deps && l_i < c->stepsave);
        ld /= (c->peak_strain * c->peak_stress * c->peak_stress);
        ld += 1.0;
        c->ld = ld;
    }
    if (c->stepsave >= 1 && c->bcount[0] > 0)
        c->destruction = (double)c->bcount[c->stepsave - 1] / (double)c->bcount[0];
    return RF_OK;
}

int ora_get_test_state(const ora_ctx* c, rf_test_state* o) {
    memset(o, 0, sizeof(*o));
    o->top = c->top; o->bottom = c->bottom; o->v = c->vw; o->size_axis = c->size_axis;
    o->size0 = c->size0; o->strain = c->strain; o->stress = c->stress;
    o->peak_stress = c->peak_stress; o->peak_strain = c->peak_strain;
    o->sfl0 = c->sfl0; o->sfl1 = c->sfl1;
    o->linearity_deviation = c->ld; o->destruction = c->destruction;
    o->step = c->step; o->steps_stop = c->steps_stop; o->n_curve = c->stepsave;
    o->loading = c->loading; o->fracture = c->fracture; o->done = c->done;
    o->kind = c->has_test ? c->cfg.kind : 0;
    return RF_OK;
}

int64_t ora_get_curve(const ora_ctx* c, double* out, int64_t cap) {
    int64_t m = c->stepsave < cap ? c->stepsave : cap;
    if (m > c->curve_cap) m = c->curve_cap;
    if (out && m > 0) memcpy(out, c->curve, (size_t)(3 * m) * sizeof(double));
    return m;
}

int64_t ora_get_bond_counts(const ora_ctx* c, int64_t* out, int64_t cap) {
    int64_t m = c->stepsave < cap ? c->stepsave : cap;
    if (m > c->curve_cap) m = c->curve_cap;
    if (out && m > 0) memcpy(out, c->bcount, (size_t)m * sizeof(int64_t));
    return m;
}

int ora_get_counters(const ora_ctx* c, rf_counters* o) {
    memset(o, 0, sizeof(*o));
    o->n = c->n;
    for (int64_t i = 0; i < c->n; ++i)
        for (int64_t k = 0; k < c->S; ++k) {
            int64_t q = i * c->S + k;
            int64_t j = c->nbr[q];
            if (j < 0) continue;
            o->occupied++;
            if (c->type[q] == 1) {
                o->bonds++;
            } else {
                double rr = dist2(c, i, j);
                double rs = c->rad[i] + c->rad[j];
                if (rr < rs * rs) o->contacts++;
            }
        }
    o->steps = c->step;
    return RF_OK;
}

int ora_build_bonds(ora_ctx* c) {
    int rc = ora_construct_rows(c);
    if (rc) return rc;
    ora_construct_bonds(c);
    ora_assign_strengths(c);
    ora_prune(c);
    return RF_OK;
}

static void copy_particles(double* dst, const double* src, int64_t n) {
    if (dst && src) memcpy(dst, src, (size_t)(3 * n) * sizeof(double));
}

int ora_get_state(const ora_ctx* c, rf_state_view* v) {
    int64_t n = c->n, ns = c->n * c->S;
    copy_particles(v->R, c->R, n); copy_particles(v->V, c->V, n); copy_particles(v->W, c->W, n);
    copy_particles(v->F, c->F, n); copy_particles(v->M, c->M, n);
    if (v->nbr) memcpy(v->nbr, c->nbr, ns * sizeof(int32_t));
    if (v->type) memcpy(v->type, c->type, ns);
    if (v->inv_L0) memcpy(v->inv_L0, c->invL0, ns * sizeof(double));
    if (v->prev_n) memcpy(v->prev_n, c->pn, 3 * ns * sizeof(double));
    if (v->o_t) memcpy(v->o_t, c->ot, 3 * ns * sizeof(double));
    if (v->m_n) memcpy(v->m_n, c->mn, 3 * ns * sizeof(double));
    if (v->m_t) memcpy(v->m_t, c->mt, 3 * ns * sizeof(double));
    if (v->sigma_c) memcpy(v->sigma_c, c->sc, ns * sizeof(double));
    if (v->tau_c) memcpy(v->tau_c, c->tc, ns * sizeof(double));
    if (v->bidx) memcpy(v->bidx, c->bidx, ns);
    return RF_OK;
}

int ora_set_state(ora_ctx* c, const rf_state_view* v) {
    int64_t n = c->n, ns = c->n * c->S;
    copy_particles(c->R, v->R, n); copy_particles(c->V, v->V, n); copy_particles(c->W, v->W, n);
    copy_particles(c->F, v->F, n); copy_particles(c->M, v->M, n);
    if (v->nbr) memcpy(c->nbr, v->nbr, ns * sizeof(int32_t));
    if (v->type) memcpy(c->type, v->type, ns);
    if (v->inv_L0) memcpy(c->invL0, v->inv_L0, ns * sizeof(double));
    if (v->prev_n) memcpy(c->pn, v->prev_n, 3 * ns * sizeof(double));
    if (v->o_t) memcpy(c->ot, v->o_t, 3 * ns * sizeof(double));
    if (v->m_n) memcpy(c->mn, v->m_n, 3 * ns * sizeof(double));
    if (v->m_t) memcpy(c->mt, v->m_t, 3 * ns * sizeof(double));
    if (v->sigma_c) memcpy(c->sc, v->sigma_c, ns * sizeof(double));
    if (v->tau_c) memcpy(c->tc, v->tau_c, ns * sizeof(double));
    if (v->bidx) memcpy(c->bidx, v->bidx, ns);
    return RF_OK;
}
