/*
 * rockfit_dem.h -- C-ABI of the MI355X (gfx950) bonded-particle DEM step.
 *
 * Plain C, plain pointers and sizes.  Every function returns an rf_status;
 * rf_last_error() gives the text of the last failure on the calling thread.
 *
 * Units: every length, mass, time, force and stress that crosses this
 * boundary is in the reference's reduced units (mm, mg, us; force 1e3 N,
 * stress 1e9 Pa, density 1e3 kg/m^3, velocity 1e3 m/s), the units the
 * reference uses on its device (md_phys_constants.h:5-14).
 *
 * Reference entry points this library replaces (file:line in
 * MSMS-LAB/DEM-Rock-Calibration):
 *   h_CalculationUniaxialCompression(md_task_data&, float*)   mp.h:16, md_calculationuniaxialcompression.cu:126
 *   h_CalculationBrazilTest(md_task_data&, float*)            mp.h:17, md_calculationbraziltest.cu:116
 *   h_LoadCylinderSample(mp_mdparameters_data&, uint, char[]) mp.h:22, md_createcylindersample.cu:441
 * and, below them, the device kernels of the loading loop (md.h:15-216).
 */
#ifndef ROCKFIT_DEM_H
#define ROCKFIT_DEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_ABI_VERSION 1

typedef enum rf_status {
    RF_OK = 0,
    RF_ERR_ARG = 1,       /* invalid argument or inconsistent sizes */
    RF_ERR_HIP = 2,       /* a HIP runtime call failed */
    RF_ERR_OVERFLOW = 3,  /* an interaction row (or candidate list) is full */
    RF_ERR_STATE = 4,     /* call out of order (e.g. step before build) */
    RF_ERR_GRID = 5,      /* a particle lies outside the padded cell grid */
    RF_ERR_ALLOC = 6,     /* device or host allocation failed */
    RF_ERR_IO = 7,        /* file could not be read or written */
    RF_ERR_NODEVICE = 8   /* no usable GPU */
} rf_status;

/* Reference quirks reproduced by default (bitmask, rf_material.quirks). */
#define RF_QUIRK_BEND_DOUBLE_SCALE  0x1u  /* bending moment scaled twice by dt*E/(4L0), md_calculateforcesdem3.cu:142-154 */
#define RF_QUIRK_ROLL_RESET         0x2u  /* |W_i|^2<=1e-18 clears the whole contact moment sum, md_calculateforcesdem3.cu:333-338 */
#define RF_QUIRK_SLOT0_SENTINEL     0x4u  /* particle 0 free-slot search uses 0 as "none", md_interaction_list_reconstructbpm.cu:28-33 */
#define RF_QUIRKS_REFERENCE (RF_QUIRK_BEND_DOUBLE_SCALE | RF_QUIRK_ROLL_RESET | RF_QUIRK_SLOT0_SENTINEL)

/* Material and contact-law parameters (potential_data, md_data_types.h:148-188,
 * values set in SetInitialParameters, mp_determineconjgradient.cu:15-47).
 * The reference has one radius and one modulus; here radii are per particle
 * and the contact modulus is separate from the bond modulus.  With equal
 * radii and E_contact == E_bond the formulas reduce to the reference ones. */
typedef struct rf_material {
    double E_bond;            /* bond Young's modulus (m_E)                      */
    double nu;                /* Poisson ratio (m_nu)                             */
    double E_contact;         /* Hertz-Mindlin Young's modulus (reference: m_E)  */
    double density;           /* m_ro                                            */
    double restitution;       /* p_e                                             */
    double friction;          /* p_mu                                            */
    double rolling_friction;  /* p_mur                                           */
    double bond_radius;       /* absolute bond radius b_r; used if ratio <= 0    */
    double bond_radius_ratio; /* b = ratio * min(r_i, r_j) when > 0              */
    double bond_gap;          /* bond if rho^2 < (r_i + r_j + gap)^2 (b_d)       */
    double sigma_c;           /* normal bond strength (m_Ec), mean if cv > 0     */
    double tau_c;             /* shear bond strength (m_Gc), mean if cv > 0      */
    double sigma_c_cv;        /* lognormal coefficient of variation, 0 = const  */
    double tau_c_cv;
    double acut;              /* neighbour cutoff a_farcut                       */
    double dt;                /* time step; <= 0 derives it (rf_derive)          */
    uint64_t strength_seed;   /* seed of the per-bond strength draw              */
    uint32_t quirks;          /* RF_QUIRK_* bitmask                              */
    uint32_t pad_;
} rf_material;

/* Space box, hidden point and cell grid (sample_data / cell_data,
 * md_data_types.h:24-124).  Particles leaving center +- half are moved to the
 * hidden point (md_parallelepipedrestriction.cu:73-160).  The cell grid has
 * its origin at 0, edge `cell` and dims[0..2] cells (the reference's Nr).  The
 * library works on that grid plus one cell layer past the top x and y faces,
 * so a particle must lie in cells 1..dims[0]-1, 1..dims[1]-1 and
 * 1..dims[2]-2, which hold the whole space box and the hidden point. */
typedef struct rf_box {
    double center[3];
    double half[3];
    double hidden[3];
    double cell;
    int32_t dims[3];
    int32_t pad_;
} rf_box;

/* A sample: N spheres, SoA coordinates and radii (reduced units). */
typedef struct rf_sample {
    int64_t n;
    const double* x;
    const double* y;
    const double* z;
    const double* r;
} rf_sample;

/* Order in which a context stores its particles (rf_options.order).
 * SPATIAL: sorted by the Morton code of (floor(4 x / cell), floor(4 y / cell),
 * floor(4 z / cell)) at creation, ties by id, so that particles close in
 * space sit close in memory and the partner gathers of a block of particles
 * hit the same cache lines; GIVEN: the caller's order.  The interface always
 * speaks in the caller's ids.  The order changes nothing but the grouping of
 * the platen-force sum (blocks of 32 particles in storage order, the only
 * per-particle sum across particles), so results of the two orders agree to
 * rounding and each is reproduced bit for bit by the CPU restatement run in
 * the same order. */
#define RF_ORDER_SPATIAL 0
#define RF_ORDER_GIVEN 1

typedef struct rf_options {
    int32_t slots;        /* interaction slots per particle (IonP), default 32  */
    int32_t cand_cap;     /* neighbour-candidate capacity per particle, <= 32 (GPU) */
    double skin;          /* candidate-list skin width (GPU, default 0.08 acut; halved
                             when a list overflows); never changes results */
    int32_t device;       /* HIP device ordinal                                */
    int32_t block;        /* ignored (the step kernel's workgroup is one wave) */
    void* stream;         /* hipStream_t the context runs on: one the caller
                             created on `device` (not the null stream), kept
                             by the caller and not destroyed by rf_destroy;
                             NULL: the context creates and owns its stream */
    int32_t order;        /* RF_ORDER_SPATIAL (default) or RF_ORDER_GIVEN        */
    int32_t pad_;
} rf_options;

/* Derived parameters (potential_data::CalculateParameters, md_data_types.h:155-187). */
typedef struct rf_derived {
    double G_bond, G_contact, hm_E, hm_G, p_A;
    double T_rayleigh, T_bond, dt;
    double r_min, r_max;
} rf_derived;

typedef enum rf_test_kind {
    RF_TEST_NONE = 0,      /* no platens: free DEM steps                    */
    RF_TEST_UNIAXIAL = 1,  /* z platens, CompressU_S3 controller            */
    RF_TEST_BRAZILIAN = 2  /* y platens, CompressB_S3 controller            */
} rf_test_kind;

/* Platen test configuration (compression_data::CalculateUParameters /
 * CalculateBParameters, md_data_types.h:212-242, 380-409; driver values at
 * mp_calculateonesample.cu:120-124). */
typedef struct rf_platen_cfg {
    int32_t kind;              /* rf_test_kind                                     */
    int32_t dnsave;            /* curve/bond-count sampling interval (dNsave)      */
    double v_load;             /* VLoad, negative compresses                       */
    double min_stress;         /* MinStress                                        */
    double max_strain;         /* MaxStrain                                        */
    double cond_fracture;      /* ConditionFracture                                */
    double size[3];            /* specimen size_real                               */
    double center[3];          /* specimen center_real                             */
    double r_ref;              /* radius used in the area formula (p_r)           */
    double wall_modulus_factor;/* 2 (uniaxial) or 1 (Brazilian) times hm_E, hm_G   */
    int64_t max_steps;         /* hard cap on loading steps (<=0: none)            */
    int64_t curve_cap;         /* capacity of the curve buffer (<=0: derived)      */
} rf_platen_cfg;

/* Controller and macro state (compression_data + result_data fields). */
typedef struct rf_test_state {
    double top, bottom, v, size_axis, size0, strain, stress;
    double peak_stress, peak_strain, sfl0, sfl1;
    double linearity_deviation, destruction;
    int64_t step;              /* loading steps executed                          */
    int64_t steps_stop;        /* StepsStop (paused-platen counter)                */
    int64_t n_curve;           /* stepsave                                         */
    int32_t loading, fracture, done, kind;
} rf_test_state;

/* Per-step counters: the sizes the algorithmic-bytes model is built from. */
typedef struct rf_counters {
    int64_t n;            /* particles                                            */
    int64_t occupied;     /* P: occupied directed slots                           */
    int64_t bonds;        /* Bd: directed bond slots                              */
    int64_t contacts;     /* Co: directed unbonded slots in overlap               */
    int64_t steps;        /* loading steps executed so far                        */
    int64_t rebuilds;     /* candidate-list rebuilds so far                       */
    int64_t candidates;   /* stored neighbour candidates (sum over particles)     */
} rf_counters;

/* Full per-particle and per-slot state, caller-owned buffers (any may be NULL
 * in rf_get_state to skip).  Particle arrays hold 3*n doubles (x block,
 * y block, z block).  Slot arrays are row-major [i*slots + k]; vectors per
 * slot are 3 consecutive doubles [(i*slots + k)*3 + c]. */
typedef struct rf_state_view {
    double* R; double* V; double* W; double* F; double* M;
    int32_t* nbr;         /* partner id, -1 = empty                                */
    uint8_t* type;        /* 1 = bond, 0 = contact slot                            */
    double* inv_L0;       /* 1/L0 per slot                                         */
    double* prev_n;       /* bond previous normal                                  */
    double* o_t;          /* tangential overlap / shear displacement (bond and contact) */
    double* m_n;          /* bond torsion history                                  */
    double* m_t;          /* bond bending history                                  */
    double* sigma_c;      /* per-slot bond strengths                               */
    double* tau_c;
    int8_t* bidx;         /* bond's position in the particle's force sum (its number
                             in ascending slot order at creation), -1 otherwise     */
} rf_state_view;

/* Opaque context: one sample, all device state. */
typedef struct rf_ctx rf_ctx;

const char* rf_last_error(void);
int32_t rf_version(void);

/* Derived parameters for a sample with radii in [r_min, r_max]. */
rf_status rf_derive(const rf_material* mat, double r_min, double r_max, rf_derived* out);

/* Reference grid set-up: space box of size `space` (sample_data::SetSpaceSizeSi,
 * CalculateParameters, md_data_types.h:87-124; cell_data::CalculateParameters
 * :42-55) with cell edge `cell` (= 2*acut).  Fills center, half, hidden, dims. */
rf_status rf_box_reference(const double space[3], double cell, rf_box* out);

rf_status rf_create(const rf_sample* sample, const rf_material* mat, const rf_box* box,
                    const rf_options* opt, rf_ctx** out);
rf_status rf_destroy(rf_ctx* ctx);

/* t=0 set-up of a loading run: cell sort, interaction rows
 * (d_ConstructInteractionList), bonds (d_ConstructBoundInteractions), per-bond
 * strengths and the first prune (d_DeleteFarLinks);
 * md_calculationuniaxialcompression.cu:155-158, 212. */
rf_status rf_build_bonds(rf_ctx* ctx);

/* Select the platen test (or RF_TEST_NONE) and reset the controller. */
rf_status rf_set_test(rf_ctx* ctx, const rf_platen_cfg* cfg);

/* Advance up to nsteps loading steps (stops early when the test ends).
 * *done_steps receives the number of steps executed. */
rf_status rf_step(rf_ctx* ctx, int64_t nsteps, int64_t* done_steps);

/* Run the selected test to its end (strain limit, fracture or max_steps) and
 * post-process (CalculateLinearityDeviation, CalculateDestruction). */
rf_status rf_run(rf_ctx* ctx, rf_test_state* out);

/* Results of one platen test in the form the reference's drivers leave them
 * (mdTD.Compress.FractureStress, FractureStrain, LinearityDeviation and
 * mdTD.R.Destruction, read back at mp_calculateonesample.cu:164-169), in
 * reduced units. */
typedef struct rf_result {
    double peak_stress;          /* UCS (uniaxial) or tensile strength (Brazilian) */
    double peak_strain;
    double secant_modulus;       /* peak_stress / peak_strain (md_results.cu:149)  */
    double linearity_deviation;  /* CalculateLinearityDeviation                    */
    double destruction;          /* CalculateDestruction                           */
    int64_t steps;               /* loading steps executed                         */
    int64_t curve_points;
    int32_t fracture;            /* the test ended on a fracture criterion         */
    int32_t kind;                /* rf_test_kind                                   */
} rf_result;

/* One whole uniaxial compression test on a built context: rf_set_test with
 * cfg (whose kind must be RF_TEST_UNIAXIAL), rf_run, results.  Replaces
 * h_CalculationUniaxialCompression(md_task_data&, float* h_R) (mp.h:16-22,
 * md_calculationuniaxialcompression.cu:126-339), which allocates, builds and
 * runs in one call; here rf_create + rf_build_bonds come first, so one
 * built sample can be tested more than once. */
rf_status rf_run_uniaxial(rf_ctx* ctx, const rf_platen_cfg* cfg, rf_result* out);
/* The same for the Brazilian test (cfg->kind RF_TEST_BRAZILIAN); replaces
 * h_CalculationBrazilTest (mp.h:16-22, md_calculationbraziltest.cu:116-264). */
rf_status rf_run_brazilian(rf_ctx* ctx, const rf_platen_cfg* cfg, rf_result* out);

rf_status rf_get_test_state(rf_ctx* ctx, rf_test_state* out);
/* Curve samples: 3 doubles each (strain, bottom stress, top stress), as
 * SaveSumForcesLoading (md_results.cu:168-180) in reduced units. */
rf_status rf_get_curve(rf_ctx* ctx, double* out, int64_t cap, int64_t* n);
/* Bond-count samples (d_SumUpInteractions, md_results.cu:189-272). */
rf_status rf_get_bond_counts(rf_ctx* ctx, int64_t* out, int64_t cap, int64_t* n);

rf_status rf_get_counters(rf_ctx* ctx, rf_counters* out);

/* Curve post-processing on the host (no device needed).  `curve` holds n
 * samples of 3 doubles (strain, bottom stress, top stress) as rf_get_curve
 * returns them; samples past n count as zero, as in the reference's
 * zero-filled buffer.
 *
 * CalculateLinearityDeviation (md_results.cu:147-165): with the secant
 * A = peak_stress / peak_strain, 1 + sum over samples of
 * deps_i (A eps_i - s_i)^2 / (peak_strain peak_stress^2), eps_i = -strain,
 * s_i = (top - bottom)/2, deps_i the central difference of eps; the sum runs
 * while eps_i < peak_strain - 2 deps_i.  rf_run reports the same value for
 * uniaxial tests. */
rf_status rf_linearity_deviation(const double* curve, int64_t n, double peak_strain,
                                 double peak_stress, double* out);

/* Moduli of a loading curve: the reference's secant peak_stress/peak_strain
 * (md_results.cu:149) and, without a reference counterpart, the least-squares
 * slope (and intercept) of s_i against eps_i over the ascending branch (the
 * samples up to the first one at peak_stress) whose stress lies in
 * [lo, hi] x peak_stress, with its coefficient of determination. */
typedef struct rf_moduli {
    double secant;
    double lsq, lsq_intercept, r2;
    int64_t points;      /* samples in the fit */
} rf_moduli;
rf_status rf_curve_moduli(const double* curve, int64_t n, double peak_strain, double peak_stress,
                          double lo, double hi, rf_moduli* out);
rf_status rf_get_state(rf_ctx* ctx, rf_state_view* out);
rf_status rf_set_state(rf_ctx* ctx, const rf_state_view* in);
int64_t rf_num_particles(const rf_ctx* ctx);
int32_t rf_num_slots(const rf_ctx* ctx);

/* Timing: rf_set_timing(ctx, k) with k > 0 times every k-th launch of
 * the fused step kernel (global step index divisible by k; in a chunk run as
 * a graph, its last step instead) with two HIP events that take the
 * dispatch's own begin and end timestamps (hipExtLaunchKernelGGL), so a
 * sample is the kernel's execution time without the launch gap; 0 switches
 * it off.  rf_last_timing returns, for
 * the last rf_step call, the summed time of the bracketed launches (ms), how
 * many launches were bracketed, the wall time of the call (ms) and the
 * number of steps it executed.  No counterpart in the reference. */
rf_status rf_set_timing(rf_ctx* ctx, int32_t every);
rf_status rf_last_timing(rf_ctx* ctx, double* step_kernel_ms, int64_t* timed_launches,
                         double* total_ms, int64_t* steps);

/* Steps per host synchronisation inside rf_step (default 256). */
rf_status rf_set_chunk(rf_ctx* ctx, int32_t steps_per_sync);

/* Set the chunk length (as rf_set_chunk) and capture now the HIP graph that
 * rf_step replays for a full chunk of that length (with timing on, the
 * chunk's first steps - 1 steps), so that the first rf_step does not pay for
 * the capture.  Runs no step.  Call after rf_set_test and rf_set_timing (the
 * graph bakes in the test's constants; they are re-captured if these change).
 * No counterpart in the reference. */
rf_status rf_capture_chunk(rf_ctx* ctx, int32_t steps);

/* Replay each full chunk of rf_step as one captured HIP graph (default on).
 * With timing on, the last step of each graphed chunk is launched on its own
 * with timing events (one timing sample per chunk).  Results are the same either
 * way.  No counterpart in the reference. */
rf_status rf_set_graph(rf_ctx* ctx, int32_t enable);

/* ------------------------------------------------------------------------
 * Sensitivities: forward-mode (tangent-linear) derivatives of the loading
 * response with respect to material parameters, up to RF_SENS_MAX per run.
 * Replaces the finite-difference Jacobian of the reference's calibration
 * (CalculateStepCG, mp_determineconjgradient.cu:84-142, NParam+1 forward
 * runs of CalculateOneSample) by one run carrying the parameter directions.
 * Between rf_sens_begin and the last rf_sens_step the context advances with
 * the sensitivity kernel (values equal the plain step's bit for bit: same
 * operations, same per-particle sum order); the slot table is held
 * (contact pairs crossing acut carry no force; a touching pair without a
 * slot, or a bond stretched past acut, ends the window with RF_ERR_STATE),
 * dt is a constant of the run: with a derived dt (rf_material.dt <= 0) a
 * direction that moves the time scale setting it (rf_derive: E_contact and
 * nu for the Rayleigh time, E_bond and the bond radius for the bond time) is
 * refused with RF_ERR_STATE.  The state at rf_sens_begin (positions,
 * velocities, histories) is treated as independent of the parameters: the
 * tangents start at zero there, so a window that starts after loading began
 * differentiates the rest of the run from that state on.
 * ------------------------------------------------------------------------ */
#define RF_SENS_MAX 4

typedef enum rf_param {
    RF_PARAM_E_BOND = 0,             /* bond Young's modulus (G follows via nu) */
    RF_PARAM_NU = 1,                 /* Poisson ratio (bond G, contact hm_E/hm_G) */
    RF_PARAM_E_CONTACT = 2,          /* contact Young's modulus               */
    RF_PARAM_FRICTION = 3,           /* sliding friction mu                   */
    RF_PARAM_ROLLING_FRICTION = 4,   /* rolling friction mu_r                 */
    RF_PARAM_BOND_RADIUS_RATIO = 5   /* bond radius / min(r_i, r_j)           */
} rf_param;

/* Derivatives of the material constants (E, G, hm_E, hm_G, mu, mu_r, p_A,
 * E_wall, G_wall, ratio, bond radius) with respect to one parameter. */
rf_status rf_sens_material_tangent(const rf_material* mat, double wall_modulus_factor,
                                   int32_t param, double* out11);
/* Zero the tangents and seed the parameter directions (params: nparam
 * rf_param values).  The curve tangents start at the next curve point. */
rf_status rf_sens_begin(rf_ctx* ctx, int32_t nparam, const int32_t* params);
/* Advance values and tangents by up to nsteps loading steps. */
rf_status rf_sens_step(rf_ctx* ctx, int64_t nsteps, int64_t* done_steps);
/* Tangents of the curve points recorded since rf_sens_begin:
 * out[(point*nparam + k)*2 + {0,1}] = d(bottom, top platen stress)/d(param k). */
rf_status rf_sens_get_curve(rf_ctx* ctx, double* out, int64_t cap, int64_t* n);

/* Frozen-branch windows.  A sensitivity window decides, at every step,
 * which bonds break, which contacts touch and which slip, whether rolling
 * friction acts and whether it chatters (then with the particle's spin held
 * fixed, as the tangents do), which particles touch a platen and slip on it,
 * and the platen velocity (controller).  Its tangents are the derivatives of
 * the run with all those decisions held.  RF_BRANCH_RECORD makes the next
 * windows of ctx (rf_sens_begin, rf_sens_step) record them, for up to `steps`
 * steps from the window's start; RF_BRANCH_REPLAY makes the next windows of
 * ctx take them from `recorded` (a context of the same sample on the same
 * device that recorded a window from the same starting step; it must stay
 * alive and record nothing else meanwhile) instead of deciding.  A replayed
 * window at the recorded parameters repeats the recorded run bit for bit;
 * at nearby parameters it is the frozen-branch model whose central
 * differences the tangents equal.  RF_BRANCH_DECIDE returns to deciding.
 * No counterpart in the reference. */
#define RF_BRANCH_DECIDE 0
#define RF_BRANCH_RECORD 1
#define RF_BRANCH_REPLAY 2
rf_status rf_sens_branches(rf_ctx* ctx, int32_t mode, int64_t steps, const rf_ctx* recorded);

/* Gradient of a loading window's macroscopic response in one call:
 * rf_sens_begin(params), rf_sens_step(steps), then over the curve points the
 * window recorded, with s_p = (|bottom| + |top|) / 2 the axial stress of
 * point p (SumForcesULoading, md_results.cu:56-99):
 *   values[0] = mean of s_p (mean axial stress),
 *   values[1] = s_q / e_q, the secant modulus at the first point q of highest
 *               s_p, e_q = -strain there (md_results.cu:149),
 *   grad[o * nparam + k] = d values[o] / d params[k].
 * The derivatives come from forward-mode tangents of the same run (one run
 * for all parameters), not from a reverse sweep; they replace the
 * forward-difference Jacobian of CalculateStepCG
 * (mp_determineconjgradient.cu:84-142).  The context needs a platen test
 * (rf_set_test); the window starts at its current step.  RF_ERR_STATE if
 * the test ends inside the window or the window records no loaded point. */
typedef struct rf_adjoint_cfg {
    int32_t nparam;                 /* 1..RF_SENS_MAX                          */
    int32_t params[RF_SENS_MAX];    /* rf_param values                         */
    int32_t pad_;
    int64_t steps;                  /* loading steps of the window             */
} rf_adjoint_cfg;
rf_status rf_adjoint_run(rf_ctx* ctx, const rf_adjoint_cfg* cfg, double* values, double* grad);

/* ------------------------------------------------------------------------
 * Batches: several samples advanced together, one step launch and one
 * controller launch per step for all of them.  No GPU counterpart in the
 * reference: it replaces running samples one after another
 * (TestSeveralSamples, mp_testseveralsamples.cu:119-161) or as separate
 * processes of a farm (NN_MP.py:71-215).  Every sample evolves exactly as it
 * would alone (same results bit for bit).  The contexts must be built, on
 * one device, and are not to be used on their own while the batch is used.
 * ------------------------------------------------------------------------ */
typedef struct rf_batch rf_batch;

rf_status rf_batch_create(rf_ctx* const* ctxs, int32_t count, rf_batch** out);
rf_status rf_batch_destroy(rf_batch* batch);
/* Advance every sample by up to nsteps steps (a sample stops early when its
 * test ends); done_steps (count entries, may be null) receives each count. */
rf_status rf_batch_step(rf_batch* batch, int64_t nsteps, int64_t* done_steps);
/* Run every sample's platen test to its end; out has count entries. */
rf_status rf_batch_run(rf_batch* batch, rf_test_state* out);
/* Wall time (ms) of the last rf_batch_step / rf_batch_run. */
rf_status rf_batch_last_ms(rf_batch* batch, double* total_ms);
/* The batch's launch autotuner (no reference counterpart): the first steps of
 * a batch run in 128-step chunks in each of five ways -- one shared launch per
 * step with the blocks ordered by sample (0), interleaved (1) or largest
 * sample first (2), each sample stepped alone, one after another (3), or
 * each sample on its own stream with all chunks launched before any is
 * waited for (4) -- timed on the host, and the fastest is kept.  The choice
 * never changes a result.  *order receives the choice (-1 before tuning),
 * ms_per_step (RF_BATCH_WAYS entries, may be null) the measured step times. */
#define RF_BATCH_WAYS 5
/* Fix the launch form (0..RF_BATCH_WAYS-1, as listed above) instead of
 * tuning it, or return to the autotuner (-1).  Results do not depend on it. */
rf_status rf_batch_set_form(rf_batch* batch, int32_t form);
rf_status rf_batch_tuning(rf_batch* batch, int32_t* order, double* ms_per_step);

/* ------------------------------------------------------------------------
 * FIRE relaxation of a packing (sample preparation).
 *
 * Replaces the relaxation loop of h_CreateCylinderSample
 * (md_createcylindersample.cu:353-420) with its parts SetFIREData
 * (md_firerelaxation.cu:72-86), d_CalculateForces (md_calculateforces.cu:20-67),
 * d_FdotVEntire (md_firerelaxation.cu:375+), d_CalculateDecrementsHalfStepFIRE
 * (:58-70), d_CalculateIncrementsFIRE (:18-56), d_CylinderRestrictionZ
 * (md_cylinderrestriction.cu:69-113) and d_ReduceParticleNumber
 * (md_reduceparticlenumber.cu:18-39).  Pair force for rho^2 < a^2,
 * a = r_i + r_j:  F_i += -D (a^2 / rho^2) (R_j - R_i).  The FIRE controller
 * runs on the device; the host waits once per call.
 * ------------------------------------------------------------------------ */
#define RF_QUIRK_FIRE_NO_INERTIA 0x8u /* V = F dt0/m overwritten each step, alpha mixing is the identity, md_firerelaxation.cu:28-52 */

typedef enum rf_container {
    RF_CONTAINER_NONE = 0,
    RF_CONTAINER_CYLINDER_Z = 1,        /* centres kept within radius and +-half_height of
                                           center; the velocity of a particle put back is
                                           reversed (d_CylinderRestrictionZ)             */
    RF_CONTAINER_CYLINDER_Z_ABSORB = 2  /* same bounds; only the outward velocity component
                                           is removed (no reference counterpart: reversed
                                           wall velocities keep the power F.V negative and
                                           stall FIRE when a packing is grown under
                                           pressure)                                      */
} rf_container;

/* Pair force of overlapping spheres (rho < a = r_i + r_j), on i:
 *   REFERENCE: -D (a^2 / rho^2) (R_j - R_i), magnitude D a^2 / rho
 *              (d_CalculateForces, md_calculateforces.cu:20-67);
 *   HARMONIC:  -D ((a - rho) / rho) (R_j - R_i), magnitude D (a - rho).
 * The harmonic law has no reference counterpart: its force vanishes at
 * contact, so FIRE converges on dense packings where the reference law
 * (finite force at contact) makes the power change sign every few steps. */
typedef enum rf_fire_law {
    RF_FIRE_LAW_REFERENCE = 0,
    RF_FIRE_LAW_HARMONIC = 1
} rf_fire_law;

typedef struct rf_fire_cfg {
    double D;                 /* repulsion strength (reduced force per mm)      */
    double dt0;               /* base step (us); dtmax = 100 dt0, dtmin = 1e-3 dt0 */
    double density;           /* mass m_i = density * 4/3 pi r_i^3 (reduced)    */
    double alpha0;            /* 0.1 in the reference loop (:349); 0.25 in SetFIREData */
    double dtgrow, dtshrink, alpha_shrink, dtmax_factor, dtmin_factor; /* 1.1, 0.5, 0.99, 100, 1e-3 */
    int32_t ndelay;           /* 20                                              */
    int32_t container;        /* rf_container                                    */
    int64_t np_negative_max;  /* 30000                                           */
    int64_t max_steps;        /* run at least this many steps (MaxStepsRelaxation) */
    int64_t hard_cap;         /* never more than this many steps (0: 10 max_steps) */
    double center[3];
    double radius, half_height, hide_z;   /* container; hide_z: z above which a
                                             particle counts as removed       */
    double hidden[3];         /* where removed particles are put              */
    double gap_tol;           /* stop only when min (rho - a)/a >= gap_tol (e.g. -0.01) */
    int32_t check_every;      /* steps between gap checks                      */
    int32_t delete_after;     /* from this step on, a check that fails removes the
                                 partner of the worst pair (<0: never)         */
    double skin;              /* neighbour-list skin (GPU only; never changes results) */
    uint32_t quirks;          /* RF_QUIRK_FIRE_NO_INERTIA                       */
    int32_t device;
    int32_t law;              /* rf_fire_law                                    */
    int32_t pad_;
} rf_fire_cfg;

typedef struct rf_fire_result {
    int64_t steps;            /* FIRE steps executed                           */
    int64_t np_positive, np_negative;
    int64_t removed;          /* particles moved to the hidden point           */
    double dt, alpha;         /* final FIRE state                              */
    double min_gap;           /* last checked min (rho - a)/a                  */
    double power;             /* last F.V                                      */
    int32_t stop_reason;      /* 1: converged after max_steps, 2: too many negative
                                 steps, 3: hard cap                            */
    int32_t pad_;
} rf_fire_result;

/* Fill cfg with the reference's FIRE constants (SetFIREData, loop :349). */
rf_status rf_fire_defaults(rf_fire_cfg* cfg);

/* Relax n particles in place: x, y, z (mm) are read and overwritten, r are
 * the radii.  `velocity` (3n: x block, y block, z block; may be null) is the
 * initial velocity and receives the final one. */
rf_status rf_fire_relax(int64_t n, double* x, double* y, double* z, const double* r,
                        const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out);

/* rf_fire_relax without any allocation: same arguments (host arrays in and
 * out) and the same results bit for bit, but every piece of device working
 * memory -- particle state, neighbour lists, cell tables, partials, the FIRE
 * control block and the staging of the host copies -- lives in one static
 * device array of 64 MiB per GPU, held until the process ends; no hipMalloc,
 * hipFree, stream creation or host allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (serving loops and
 * harnesses that forbid device allocation).  Limit: RF_ERR_ARG when the run
 * needs more than the array (about 140,000 particles at packing grids).
 * Runs on the null stream and waits for it.  Every call on a GPU uses that
 * GPU's one array: calls from several threads of the process are serialized
 * by a lock, and nothing else may use the array. */
rf_status rf_fire_relax_noalloc(int64_t n, double* x, double* y, double* z, const double* r,
                                const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out);

#ifdef __cplusplus
}
#endif

#endif /* ROCKFIT_DEM_H */
