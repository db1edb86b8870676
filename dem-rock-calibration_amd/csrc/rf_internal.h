// rf_internal.h -- device data layout and launch interface of the gfx950 DEM
// step.  Shared by the kernels (rf_kernels.hip) and the C-ABI host runtime
// (rf_runtime.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rockfit_dem.h"

namespace rf {

// Step kernel: one 8-lane group per particle, lane l of a group owning
// slots, bonds and candidates l, l+8, l+16, l+24; one wave (8 consecutive
// particles) per workgroup, so a wave's slot is refilled as soon as it ends.
// The platen forces are summed per group of 32 consecutive particles: each
// wave's 8 by a tree inside the wave, then the group's four waves by the
// controller (see step_body).
constexpr int kLanes = 8;
constexpr int kBlock = 64;                  // threads of a step workgroup: one wave
constexpr int kPerWave = kBlock / kLanes;   // 8 particles per wave
constexpr int kGroup = 32;                  // particles per platen-partial group
constexpr int kWavesPerGroup = kGroup / kPerWave;  // 4
constexpr int kPerLane = 4;                 // entries per lane and list
constexpr int kMaxSlots = kLanes * kPerLane;  // 32 slots per particle
constexpr int kMaxCand = kLanes * kPerLane;   // 32 candidates per particle
constexpr int kMaxBonds = kLanes * kPerLane;  // 32 bonds per particle
constexpr int kCtrlThreads = 1024;          // threads of the controller block

// Error bits raised by device code (Ctrl::err).
constexpr int kErrRowOverflow = 1;   // no free slot for a new contact
constexpr int kErrCandOverflow = 2;  // candidate list longer than cand_cap
constexpr int kErrGrid = 4;          // particle outside the padded cell grid
constexpr int kErrCurve = 8;         // curve buffer full
constexpr int kErrSensList = 16;     // a sensitivity window would change the slot table

// Parameter directions carried by one sensitivity run.
constexpr int kSensK = 4;
// Waves of the sensitivity step a CU holds at once: one per SIMD (its registers).
constexpr int kSensWavesPerCU = 4;

// One bond of a particle's bond table: 128 bytes, the 12 history values
// first (written back each step), then the constants.
struct BondRec {
    double pnx, pny, pnz;   // previous bond normal
    double otx, oty, otz;   // shear displacement
    double mnx, mny, mnz;   // torsion history
    double mtx, mty, mtz;   // bending history
    double invL0, sc, tc;   // 1/L0, normal and shear strength
    double L0;              // 1.0 / invL0, formed once (the same IEEE quotient)
};
static_assert(sizeof(BondRec) == 128, "bond record must be one 128-byte line");

// Device-resident arrays.  Particles are array-of-structs double4 so a
// neighbour gather is one 32-byte load:
//   pos = (x, y, z, radius), vel = (vx, vy, vz, mass), angv = (wx, wy, wz, 0).
// pos/vel/angv are double-buffered; Ctrl::cur selects the input buffer.
// Per-particle rows are row-major so the 32 lanes of a group read one
// contiguous run: slots [i*S + k], candidates [i*C + c], bonds [i*B + b].
struct DevState {
    double4* pos[2];
    double4* vel[2];
    double4* angv[2];
    double2* pk;       // (dt/m_i, dt/I_i)
    double4* F;        // last step's force  (x, y, z, 0)
    double4* M;        // last step's moment
    // position at the last candidate rebuild (x, y, z rounded to float) and
    // the squared displacement that asks for the next one (see rebuild_mark)
    float4* rbuild;
    // slot table
    int32_t* nbr;      // partner of slot k (valid where the occupancy bit is set)
    double4* ot;       // contact tangential overlap (x, y, z, 0)
    uint32_t* occ;     // occupied slots
    uint32_t* bnd;     // bond slots
    uint32_t* otnz;    // contact slots whose o_t may be non-zero
    // bond table (bonds numbered in ascending slot order at creation)
    int32_t* nbond;
    int8_t* bslot;     // slot of bond b, -1 = none
    int32_t* bpart;    // partner of bond b
    BondRec* brec;
    // candidate lists
    int32_t* cand; int8_t* cslot; int32_t* ncand;
    // cell sort
    int32_t* ckey; int32_t* ccount; int32_t* cstart; int32_t* cfill; int32_t* ctmp;
    int32_t* sorted;
    // caller's id of each stored particle (particles are stored in spatial
    // order; ids rank new pairs and order the cells, as in the reference)
    int32_t* orig;
    // reductions
    // per group of 32 particles, one entry per wave: [4 b + w]
    double* part_b; double* part_t; int32_t* part_bonds;
    // bit b: group b has a non-zero platen partial this step (set by the
    // step kernel, read and cleared by the controller)
    uint32_t* pmask;
    // curve
    double* curve; int64_t* bcount;
};

// Constant per run, passed by value.
struct DevParams {
    int64_t n;
    int32_t S, C, B;
    int32_t nblk;        // blocks of the step kernel
    double dt, E, G, hmE, hmG, pA, mu, mur;
    double acut2, cell, inv_cell, cand_r2, skin_half2;
    double b_ratio, b_abs, bond_gap;
    uint32_t quirks;
    int32_t nx, ny, nz;
    int64_t ncell;
    double cen[3], half[3], hid[3];
    int32_t kind;        // rf_test_kind
    int32_t axis;        // platen axis: 2 uniaxial, 1 Brazilian
    double Ew, Gw;       // wall moduli (factor * hm_E, factor * hm_G)
    double scope_c[3];
    double rr_scope, zcut;
    double v_load, min_stress, max_strain, cond_fracture, zero_stress, inv_area;
    int32_t dnsave, store_fm;
    int64_t max_steps, curve_cap;
    int64_t zero_idx;    // storage index of the caller's particle 0
};

// A value that is the same in every lane, held in scalar registers.
__device__ __forceinline__ double uniform(double v) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Rebuild test.  A candidate rebuild is due once a particle has moved more
// than skin/2 since the last one.  Only the timing of rebuilds depends on
// this test, never a result, so the mark is kept in single precision with a
// conservative threshold: x, y, z rounded to float (each off by at most
// 2^-24 |x|), and the squared threshold (skin/2 - eta)^2 rounded down, with
// eta = 2^-22 (|x| + |y| + |z|) four times the largest rounding of the mark.
// A particle that moved more than skin/2 therefore always asks for the
// rebuild; one that moved a little less may ask a step early.
__device__ __forceinline__ float4 rebuild_mark(const DevParams& p, const double4& x) {
    const double eta = 0x1p-22 * ((fabs(x.x) + fabs(x.y)) + fabs(x.z));
    const double th = sqrt(p.skin_half2) - eta;
    const double th2 = th > 0.0 ? th * th : 0.0;
    return make_float4(__double2float_rn(x.x), __double2float_rn(x.y), __double2float_rn(x.z),
                       __double2float_rd(th2));
}

__device__ __forceinline__ bool rebuild_due(const float4& m, double x, double y, double z) {
    const double ex = x - (double)m.x, ey = y - (double)m.y, ez = z - (double)m.z;
    return ex * ex + ey * ey + ez * ez > (double)m.w;
}

// Device control block: written by the controller kernel once per step.
struct Ctrl {
    double top, bottom, vw, size_axis, size0, strain, stress;
    double peak_stress, peak_strain, sfl0, sfl1;
    int64_t step, steps_stop, stepsave, last_bonds;
    int32_t loading, fracture, done, halt;
    int32_t cur, rebuild, err, pad;
    int64_t stop_at;  // halt once step reaches this (0: no limit; batched runs)
};

// Tangent state of a sensitivity run: kSensK directions, direction k of
// particle i at [k*n + i], of bond entry q at [(k*n*B + q)*12 + u] (the 12
// history values), of slot q at [k*n*S + q].  dmat[k] holds the derivatives
// of (E, G, hm_E, hm_G, mu, mu_r, p_A, E_wall, G_wall, bond radius ratio,
// bond radius) along direction k.
struct SensState {
    int32_t nl;        // lanes per particle = parameter directions in use
    int32_t nwaves;    // waves of a sensitivity step
    int32_t* wbase;    // wave w steps particles wbase[w] .. wbase[w + 1] - 1
    double* pp;        // per particle: bottom, top platen force (values)
    int32_t* pnb;      // per particle: bonds left
    double* tpp;       // per particle and direction: bottom, top tangents
    double4* tpos[2];
    double4* tvel[2];
    double4* tang[2];
    double* tbrec;
    double4* tot;
    double* tpart_b;
    double* tpart_t;
    double* tcurve;   // [point][k][bottom, top] stress tangents
    double dmat[kSensK][11];
    // Frozen-branch windows: br_mode 1 records every discrete decision of
    // the window, 2 replays a recorded run's decisions instead of taking
    // them (0: decide).  Per step s (from the window's first step br_step0)
    // and particle i, br_bits[(s*n + i)*kBrWords + w] holds: w 0 the contact
    // slots that touch, 1 the contact slots that slip, 2 the contact slots
    // whose rolling torque chatters, 3 the bond slots that break, 4 the
    // particle flags kBf*; br_w[s*n + i] the spin a chattering rolling torque
    // holds fixed; br_vw[s] the platen velocity the controller set.
    int32_t br_mode;
    int32_t pad2_;
    int64_t br_step0, br_steps;
    uint32_t* br_bits;
    double4* br_w;
    double* br_vw;
};

constexpr int kBrWords = 5;
// particle flags (word 4 of a frozen-branch record)
constexpr uint32_t kBfRoll = 1u;        // |W|^2 > 1e-18
constexpr uint32_t kBfScope = 2u;       // inside the platens' scope
constexpr uint32_t kBfTop = 4u;         // touches the top platen
constexpr uint32_t kBfTopSlip = 8u;
constexpr uint32_t kBfTopChat = 16u;
constexpr uint32_t kBfBot = 32u;        // touches the bottom platen
constexpr uint32_t kBfBotSlip = 64u;
constexpr uint32_t kBfBotChat = 128u;

// The platen velocity of a frozen-branch window: recorded (mode 1) or
// replayed (mode 2) by the controller.
struct VwLog {
    int32_t mode;
    int32_t pad_;
    int64_t step0, steps;
    double* vw;
};

hipError_t launch_sens_step(const DevState& s, const DevParams& p, const SensState& t, Ctrl* ctrl,
                            hipStream_t st);
hipError_t launch_controller_sens(const DevState& s, const DevParams& p, const SensState& t,
                                  Ctrl* ctrl, hipStream_t st);

// One sample of a batched launch: its state, constants and control block.
struct BatchEntry {
    DevState s;
    DevParams p;
    Ctrl* ctrl;
};

// Sets the text rf_last_error() returns and passes `st` through
// (rf_runtime.hip).
rf_status report(rf_status st, const char* fmt, ...);

// Launch wrappers (rf_kernels.hip).
hipError_t launch_cell_keys(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_cell_sort(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_construct_rows(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_construct_bonds(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_prune_initial(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_build_candidates(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_clear_flags(Ctrl* ctrl, hipStream_t st);
hipError_t launch_set_stop(Ctrl* ctrl, int64_t stop_at, hipStream_t st);
hipError_t launch_set_strengths(BondRec* brec, const double* sc, const double* tc, int64_t n,
                                hipStream_t st);
hipError_t launch_step(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
// The same launch with `start` and `stop` taking the dispatch's own begin and
// end timestamps (hipExtLaunchKernelGGL): the kernel's execution time only.
hipError_t launch_step_timed(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st,
                             hipEvent_t start, hipEvent_t stop);
hipError_t launch_controller(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st);
hipError_t launch_step_batch(const BatchEntry* e, const int32_t* blk_sample,
                             const int32_t* blk_local, int32_t nblk_total,
                             hipStream_t st);
hipError_t launch_controller_batch(const BatchEntry* e, int32_t count, hipStream_t st);
hipError_t launch_counters(const DevState& s, const DevParams& p, Ctrl* ctrl, int64_t* out3,
                           hipStream_t st);

}  // namespace rf
