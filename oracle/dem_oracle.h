/*
 * dem_oracle.h -- CPU FP64 restatement of the reference's BPM loading step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (librockfit_dem.so) never links,
 * loads or calls it.
 *
 * It shares the plain-data struct definitions of include/rockfit_dem.h so
 * that a test can hand the same inputs to both sides; no code is shared.
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#include "../include/rockfit_dem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

int ora_derive(const rf_material* mat, double r_min, double r_max, rf_derived* out);
int ora_box_reference(const double space[3], double cell, rf_box* out);

ora_ctx* ora_create(const rf_sample* s, const rf_material* mat, const rf_box* box,
                    int32_t slots, int32_t nthreads);
void ora_destroy(ora_ctx* c);
const char* ora_error(const ora_ctx* c);
int64_t ora_sizeof(int32_t which);
void ora_set_threads(ora_ctx* c, int32_t nthreads);
/* Storage order of the device (RF_ORDER_SPATIAL, the default, or
 * RF_ORDER_GIVEN): the grouping of the platen-force sum.  Call before the
 * first step. */
int ora_set_order(ora_ctx* c, int32_t mode);
int ora_get_order(const ora_ctx* c, int64_t* out);

/* Cell keys, stable (cell, id) order, cell start/count (CellDistribution). */
int ora_cells(ora_ctx* c);
int ora_get_cells(const ora_ctx* c, int32_t* key, int32_t* sorted, int32_t* start,
                  int32_t* count, int64_t ncell_cap);
int64_t ora_num_cells(const ora_ctx* c);

/* t=0: rows, bonds, strengths, first prune. */
int ora_build_bonds(ora_ctx* c);
/* Individual phases (for unit tests). */
int ora_construct_rows(ora_ctx* c);   /* d_ConstructInteractionList       */
int ora_construct_bonds(ora_ctx* c);  /* d_ConstructBoundInteractions     */
int ora_assign_strengths(ora_ctx* c);
int ora_append(ora_ctx* c);           /* d_ReConstructInteractionListbpm  */
int ora_prune(ora_ctx* c);            /* d_DeleteFarLinks                 */
int ora_forces(ora_ctx* c);           /* F = M = 0; bonds; contacts; platens */
int ora_integrate(ora_ctx* c);        /* kick-drift + cut                 */

int ora_set_test(ora_ctx* c, const rf_platen_cfg* cfg);
/* Loading steps; returns the number executed, or a negative rf_status. */
int64_t ora_step(ora_ctx* c, int64_t nsteps);
int ora_post(ora_ctx* c);             /* linearity deviation, destruction */
int ora_get_test_state(const ora_ctx* c, rf_test_state* out);
int64_t ora_get_curve(const ora_ctx* c, double* out, int64_t cap);
int64_t ora_get_bond_counts(const ora_ctx* c, int64_t* out, int64_t cap);
int ora_get_counters(const ora_ctx* c, rf_counters* out);
int ora_get_state(const ora_ctx* c, rf_state_view* v);
int ora_set_state(ora_ctx* c, const rf_state_view* v);
int ora_get_partials(const ora_ctx* c, double* bottom, double* top, int64_t cap);
int ora_get_particle_partials(const ora_ctx* c, double* bottom, double* top, int64_t cap);

/* Lognormal per-bond strength draw (also restated by the product host code). */
/* FIRE relaxation of a packing (fire_oracle.c); same arguments as
 * rf_fire_relax plus the OpenMP thread count. */
int ora_fire_defaults(rf_fire_cfg* cfg);
int ora_fire_relax(int64_t n, double* x, double* y, double* z, const double* r,
                   const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out,
                   int32_t nthreads);

double ora_bond_strength(uint64_t seed, int64_t lo, int64_t hi, int32_t stream,
                         double mean, double cv);

#ifdef __cplusplus
}
#endif
#endif
