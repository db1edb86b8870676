// rf_kernels.hip -- gfx950 kernels of the bonded-particle DEM loading step.
//
// One 8-lane group per particle, eight consecutive particles per wave64 and
// one wave per workgroup; lane l of a group owns slots, candidates and bonds
// l, l+8, l+16, l+24.  In one launch a group
//   1. prunes and appends its interaction row (slot lanes check the stored
//      partners, candidate lanes look for new ones),
//   2. evaluates its bonds and contacts (each lane its own entries, in
//      ascending order) and reduces the 8 lane sums with a fixed tree,
//   3. adds the platen forces, kick-drifts the particle, applies the cut
//      restriction and writes the next state buffer.
// Particles are stored in spatial (Morton) order (rf_create); the ids of
// the caller rank new pairs and order the cells.
// A one-block controller kernel then sums the platen partials, runs the
// stress/peak/fracture logic and moves the platens.  No float atomics: every
// sum has a fixed order, so results are bitwise reproducible and equal to the
// CPU restatement's.
//
// Reference kernels covered (MSMS-LAB/DEM-Rock-Calibration):
//   d_ReConstructInteractionListbpm, d_DeleteFarLinks  md_interaction_list_reconstructbpm.cu:13-186
//   d_CalculateForcesDEM_22 / _21                      md_calculateforcesdem3.cu:20-486
//   d_UniaxialCompression_hm / d_BrazilCompression_hm md_compression_hm.cu:20-449
//   d_CalculateIncrementsDEM                           md_calculateincrementsdem.cu:19-49
//   d_ParallelepipedCutRestriction                     md_parallelepipedrestriction.cu:73-160
//   SumForcesULoading/BLoading, CompressU_S3/B_S3      md_results.cu:56-145, md_data_types.h:347-494
//   CellDistribution, d_ConstructInteractionList,
//   d_ConstructBoundInteractions                        md_cell_distribution.cu:52-101,
//                                                       md_interaction_list_construct.cu:17-130,
//                                                       md_boundinteractions_construct.cu:35-106
//
// Floating point: compiled with -ffp-contract=off; sqrt and division are the
// IEEE-rounded FP64 operations; expressions are evaluated left to right as
// written, in the same order as the CPU restatement.
#include <hip/hip_ext.h>

#include "rf_internal.h"
#include "rf_physics.h"

#include "../../include/rockfit_dem.h"

namespace rf {


// Position of a stencil cell in the reference's visiting order, indexed by
// (dx | dy << 1 | dz << 2) where d = 1 means "the neighbour on the particle's
// side along that axis": own, x, y, z, xy, xz, yz, xyz.
// The order is 0, 1, 2, 4, 3, 5, 6, 7: only indices 3 (xy) and 4 (z) swap.
__device__ __forceinline__ uint32_t stencil_pos(int idx) {
    return idx == 3 ? 4u : (idx == 4 ? 3u : (uint32_t)idx);
}

// Bits of an 8-lane group (bit l = lane l of the group).
__device__ __forceinline__ uint32_t grp_ballot(bool pred) {
    const unsigned long long m = __ballot(pred);
    const int gw = (threadIdx.x >> 3) & 7;
    return (uint32_t)((m >> (8 * gw)) & 0xffu);
}

__device__ __forceinline__ uint32_t grp_or(uint32_t v) {
#pragma unroll
    for (int off = kLanes / 2; off >= 1; off >>= 1) v |= (uint32_t)__shfl_xor((int)v, off, kLanes);
    return v;
}

// Sum of the 8 lane partials by the tree v[t] = v[t] + v[t + off], off =
// 4, 2, 1; every lane gets the bitwise-identical total (a + b == b + a).
__device__ __forceinline__ double tree8(double v) {
#pragma unroll
    for (int off = kLanes / 2; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, kLanes);
    return v;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    const int lo = __shfl((int)(uint32_t)v, src, kLanes);
    const int hi = __shfl((int)(uint32_t)(v >> 32), src, kLanes);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// double4 of lane `src` of the wave.
__device__ __forceinline__ double4 shfl_d4(const double4& v, int src) {
    return make_double4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src));
}

__device__ __forceinline__ int nth_bit(uint32_t m, int r) {
    for (int t = 0; t < r; ++t) m &= m - 1u;
    return m ? __builtin_ctz(m) : -1;
}

// r-th free slot in the order the reference fills them.  Particle 0 with the
// sentinel quirk fills slot 0 last (md_interaction_list_reconstructbpm.cu:28-33).
__device__ __forceinline__ int free_slot(uint32_t freem, int r, bool q0) {
    if (q0 && (freem & 1u) && (freem & ~1u)) {
        const uint32_t f = freem & ~1u;
        const int c = __popc(f);
        if (r < c) return nth_bit(f, r);
        return r == c ? 0 : -1;
    }
    return r < __popc(freem) ? nth_bit(freem, r) : -1;
}

// Double-precision entry points of the pair physics (rf_physics.h).
using WallOut = WallOutT<double>;

__device__ __forceinline__ void wall_contact(const DevParams& p, double ri, double mi,
                                             const double4& vi, const double4& wi,
                                             double nwx, double nwy, double nwz,
                                             double wvx, double wvy, double wvz, double d,
                                             double opx, double opy, double opz, WallOut& o) {
    wall_contact_t<double>(mat_of(p), p, ri, mi, v4(vi), v4(wi), nwx, nwy, nwz, wvx, wvy, wvz, d,
                           opx, opy, opz, 0.0, o);
}

// ------------------------------------------------------------ step kernel --
#ifndef RF_XCD_REMAP
#define RF_XCD_REMAP 1
#endif
#ifndef RF_GATHER_BARRIER
#define RF_GATHER_BARRIER 1
#endif
#ifndef RF_BOND_SPREAD
#define RF_BOND_SPREAD 1
#endif
// Entries that few particles use (candidates 8..31, partners in slots
// 24..31) are loaded after the first batch, only in the waves that have any.
#ifndef RF_CAND_TAIL_LATE
#define RF_CAND_TAIL_LATE 1
#endif
// Three waves per SIMD: the step's register budget is 168 VGPRs (the
// rebuild mark is loaded at the integration, the slot partners of a contact
// round are read back from the row, and the slots 24..31 gather after the
// first batch), which hides a third more of its memory latency than two
// waves.
#ifndef RF_STEP_WAVES_PER_EU
#define RF_STEP_WAVES_PER_EU 3
#endif
#if RF_STEP_WAVES_PER_EU > 0
#define RF_STEP_ATTR __attribute__((amdgpu_waves_per_eu(RF_STEP_WAVES_PER_EU)))
#else
#define RF_STEP_ATTR
#endif
// The batched step (which looks up its sample's arrays and constants per
// workgroup) fits the same 168 registers once the platens' positions and
// speed sit in scalar registers.
#ifndef RF_BATCH_WAVES_PER_EU
#define RF_BATCH_WAVES_PER_EU 3
#endif
#if RF_BATCH_WAVES_PER_EU > 0
#define RF_BATCH_ATTR __attribute__((amdgpu_waves_per_eu(RF_BATCH_WAVES_PER_EU)))
#else
#define RF_BATCH_ATTR
#endif

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w runs on XCD
// w % 8).  Give each XCD one contiguous range of particle blocks instead, so
// the partner gathers of a block hit the L2 of its own XCD.  A bijection on
// [0, nblk) for any nblk.
__device__ __forceinline__ int xcd_block(int w, int nblk) {
#if RF_XCD_REMAP
    constexpr int kXcd = 8;
    const int x = w % kXcd, q = w / kXcd;
    const int base = nblk / kXcd, extra = nblk % kXcd;
    return x * base + (x < extra ? x : extra) + q;
#else
    (void)nblk;
    return w;
#endif
}

// x, y, z of a double4 (the w component left unread: 0.0).
__device__ __forceinline__ double4 ld3(const double4* a, int64_t q) {
    const double* d = reinterpret_cast<const double*>(a + q);
    const double2 xy = *reinterpret_cast<const double2*>(d);
    return make_double4(xy.x, xy.y, d[2], 0.0);
}

__device__ __forceinline__ bool bond_eval(const DevParams& p, const double4& pi, const double4& vi,
                                          const double4& wi, const double4& pj, const double4& vj,
                                          const double4& wj, const double2 (&r)[8],
                                          double (&c)[6], double2 (&h)[6]) {
    const double hist[12] = {r[0].x, r[0].y, r[1].x, r[1].y, r[2].x, r[2].y,
                             r[3].x, r[3].y, r[4].x, r[4].y, r[5].x, r[5].y};
    double hn[12];
    const bool brk = bond_eval_t<double>(mat_of(p), p, v4(pi), v4(vi), v4(wi), v4(pj), v4(vj),
                                         v4(wj), hist, r[6].x, r[7].y, r[6].y, r[7].x, c, hn);
#pragma unroll
    for (int u = 0; u < 6; ++u) h[u] = make_double2(hn[2 * u], hn[2 * u + 1]);
    return brk;
}

__device__ __forceinline__ void contact_eval(const DevParams& p, const double4& pi,
                                             const double4& vi, const double4& wi,
                                             const double4& pj, const double4& vj,
                                             const double4& wj, const double4& oprev, double wim,
                                             bool roll, double (&c)[6], double4& onew) {
    const double op[3] = {oprev.x, oprev.y, oprev.z};
    double on[3];
    contact_eval_t<double>(mat_of(p), p, v4(pi), v4(vi), v4(wi), v4(pj), v4(vj), v4(wj), op, wim,
                           roll, 0.0, c, on);
    onew = make_double4(on[0], on[1], on[2], 0.0);
}

// One 8-lane group per particle; lane l owns slots, candidates and bonds
// l + 8t (t = 0..3).  Per-particle sums: each lane adds its own entries in
// ascending order, then tree8.
//
// Memory schedule: every per-particle row (masks, slot row, candidate row,
// first bond index) is issued first, then the particle's state; then every
// gather that depends only on those (partners of slots 0..23 for the prune
// test, absent candidates, first bond's partner and record) is issued at
// once with clamped indices and its result masked afterwards, so a wave
// waits on two memory round trips before any force work.  What that batch
// leaves out keeps the step within the 168 registers of three waves per
// SIMD: the partners of slots 24..31 (gathered after it, in the waves that
// have any), (dt/m, dt/I) and the rebuild position (read at the
// integration), the partner of a contact slot (read back from the row in
// its contact round).  Contacts gather in rounds of one slot per lane.
// Bonds beyond the 8th and appends are rare paths.
// `wave`: particles 8 wave + k (k = 0..7, the wave's 8-lane groups); wave w
// = wave % 4 of the 32-particle group wave / 4.
__device__ __forceinline__ void step_body(const DevState& s, const DevParams& p,
                                          Ctrl* __restrict__ ctrl, const int wave) {
    const int l = threadIdx.x & (kLanes - 1);
    const int grp = threadIdx.x / kLanes;
    const int64_t gbase = (int64_t)wave * kPerWave;
    const int64_t i = gbase + grp;
    const bool valid = i < p.n;
    // The control block (written by the previous controller launch) and the
    // particle's own rows, which do not depend on it, are requested together;
    // the halt test (uniform for the whole grid) waits for both.
    const int32_t halt = ctrl->halt;
    const int cur = ctrl->cur;
    // the platens' positions and speed: uniform, kept in scalar registers
    // until the integration reads them
    const double top = uniform(ctrl->top), bottom = uniform(ctrl->bottom),
                 vw = uniform(ctrl->vw);
    const int S = p.S, C = p.C, B = p.B;
    const int64_t ic = valid ? i : 0;  // clamped index for the inactive groups' loads

    // ---- stage 1: per-particle loads (the rebuild mark is read at the
    // integration, where it is used)
    const uint32_t occ_raw = s.occ[ic], bnd_raw = s.bnd[ic], otnz_raw = s.otnz[ic];
    const int nc_raw = s.ncand[ic], nbd_raw = s.nbond[ic];
    const uint32_t occ_in = valid ? occ_raw : 0u;
    const uint32_t bnd_in = valid ? bnd_raw : 0u;
    uint32_t otnz = valid ? otnz_raw : 0u;
    const int nc = valid ? nc_raw : 0;
    const int nbd = valid ? nbd_raw : 0;
    int32_t jk[kPerLane], cj[kPerLane];
    int cs[kPerLane];
    // Candidate entries 8.. exist only in particles with more than 8
    // candidates (a few per cent): those are loaded below, once the counts
    // are in, and only in the waves that have such a particle; until then
    // they read as absent (entries past the count are never used).
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int k = l + kLanes * t;
        if (t < kPerLane - 1 || !RF_CAND_TAIL_LATE)
            jk[t] = s.nbr[ic * S + (k < S ? k : S - 1)];
        else
            jk[t] = (int32_t)ic;
        if (t == 0 || !RF_CAND_TAIL_LATE) {
            const int64_t qc = ic * C + (k < C ? k : C - 1);
            cj[t] = s.cand[qc];
            cs[t] = s.cslot[qc];
        } else {
            cj[t] = (int32_t)ic;
            cs[t] = -1;
        }
    }
    const int64_t qb0 = ic * B + l;
    const int bs0 = s.bslot[qb0];
    const int32_t bp0 = s.bpart[qb0];
    // bond l + 8 of the particle (evaluated by a lane without a bond, below)
    const int64_t qb1 = ic * B + (l + kLanes < B ? l + kLanes : B - 1);
    const int bs1 = s.bslot[qb1];
    const int32_t bp1 = s.bpart[qb1];

#if RF_CAND_TAIL_LATE
    // Slots 24..31 hold partners only in particles with many partners or a
    // fragmented row (a few per cent of waves); the others skip their loads.
    const bool hi_slots = __any((occ_in >> (kLanes * (kPerLane - 1))) != 0u);
    if (hi_slots) {  // wave-uniform
        const int k = l + kLanes * (kPerLane - 1);
        jk[kPerLane - 1] = s.nbr[ic * S + (k < S ? k : S - 1)];
    }
    if (__any(nc > kLanes)) {  // wave-uniform; rare
#pragma unroll
        for (int t = 1; t < kPerLane; ++t) {
            const int k = l + kLanes * t;
            const int64_t qc = ic * C + (k < C ? k : C - 1);
            cj[t] = s.cand[qc];
            cs[t] = s.cslot[qc];
        }
    }
#endif

    // both buffers' addresses in registers before the test, so the select
    // below is not turned back into a kernel-argument load after it
    using GPos = const __attribute__((address_space(1))) double4*;  // global memory
    GPos pos0 = (GPos)s.pos[0], pos1 = (GPos)s.pos[1];
    asm volatile("" : "+s"(pos0), "+s"(pos1));
    if (halt) return;
    const double4* __restrict__ pos = (const double4*)(cur ? pos1 : pos0);
    const double4* __restrict__ vel = cur ? s.vel[1] : s.vel[0];
    const double4* __restrict__ angv = cur ? s.angv[1] : s.angv[0];
    const double4 pi = pos[ic];
    const double4 vi = vel[ic];
    const double4 wi = ld3(angv, ic);

    // ---- stage 2: gathers (clamped to particle i where not needed)
    double4 pp[kPerLane];
    uint32_t cab = 0u;  // bit t: entry l + 8t is a candidate without a slot
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int k = l + kLanes * t;
        const bool had = k < S && ((occ_in >> k) & 1u);
#if RF_CAND_TAIL_LATE
        if (t == kPerLane - 1)
            pp[t] = pi;  // gathered after this batch, in the waves that need it
        else
#endif
            pp[t] = pos[had ? jk[t] : (int32_t)ic];
        cab |= (uint32_t)((k < nc) & (k < C) & (cs[t] < 0)) << t;
    }
    // Lists are rebuilt without the current partners, so the candidates to
    // test sit in the first entries: one gather per lane covers them.
    int tc = cab ? __builtin_ctz(cab) : kPerLane;
    int32_t jc = (int32_t)ic;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u)
        if (u == tc) jc = cj[u];
    double4 pcj = pos[jc];
    const bool pre0 = l < nbd && bs0 >= 0 && ((bnd_in >> bs0) & 1u);
    const bool pre1 = l + kLanes < nbd && bs1 >= 0 && ((bnd_in >> bs1) & 1u);
    // Bond spread: in a wave whose particles have at most 16 bonds, and no
    // more bonds beyond the 8th than lanes without a bond of their own,
    // bond l + 8 of lane `to`'s particle is evaluated by lane `to` in the
    // same round as every lane's bond l (the k-th such bond by the k-th
    // free lane), and its contributions go back to its owner lane, which
    // adds them after bond l as it would in a second round.  Lane 0 of a
    // group always keeps its own particle.
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long msrc = __ballot(l + kLanes < nbd);
    const unsigned long long midle = __ballot(l >= nbd && l > 0);
    const bool spread = RF_BOND_SPREAD && msrc != 0ull &&
                        __popcll(msrc) <= __popcll(midle) && !__any(nbd > 2 * kLanes);
    int from = lane, to = -1;  // lane whose bond l + 8 this lane evaluates / evaluating ours
    int32_t bjx = pre0 ? bp0 : (int32_t)ic;  // partner gathered for this lane's bond
    int64_t qbx = pre0 ? qb0 : ic * B;      // record gathered
    int bsx = -1;                            // slot (in its owner's table) of the spread bond
    if (spread) {
        unsigned long long a = msrc, f = midle;
        while (a) {  // wave-uniform
            const int s_ = __builtin_ctzll(a);
            const int t_ = __builtin_ctzll(f);
            a &= a - 1ull;
            f &= f - 1ull;
            if (lane == t_) from = s_;
            if (lane == s_) to = t_;
        }
        const int32_t bjs = __shfl(pre1 ? bp1 : (int32_t)ic, from);
        const int bss = __shfl(pre1 ? bs1 : -1, from);
        if (from != lane) {
            const int64_t io = gbase + from / kLanes;
            bjx = bjs;
            qbx = bss >= 0 ? io * B + kLanes + (from & (kLanes - 1)) : io * B;
            bsx = bss;
        }
    }
    const double4 pb0 = pos[bjx];
    const double4 vb0 = ld3(vel, bjx);
    const double4 wb0 = ld3(angv, bjx);
    double2 rec0[8];
    {
        const double2* rp = reinterpret_cast<const double2*>(s.brec + qbx);
#pragma unroll
        for (int u = 0; u < 8; ++u) rec0[u] = rp[u];
    }
    // keep every gather above issued before the first use of any of them
#if RF_GATHER_BARRIER
    __builtin_amdgcn_sched_barrier(0);
#endif

    // ---- 1a. prune (d_DeleteFarLinks) and the possible-contact prefilter
    const double ri = pi.w;
    uint32_t prn_l = 0u, pc_l = 0u;
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int k = l + kLanes * t;
        const bool had = k < S && ((occ_in >> k) & 1u);
#if RF_CAND_TAIL_LATE
        // Slots 24..31 hold partners only in particles with many partners or
        // a fragmented row (a few per cent of waves): their gather comes
        // here, in those waves only.
        if (t == kPerLane - 1) {
            if (!hi_slots) break;  // wave-uniform
            pp[t] = pos[had ? jk[t] : (int32_t)ic];
        }
#endif
        const double dx = pp[t].x - pi.x, dy = pp[t].y - pi.y, dz = pp[t].z - pi.z;
        const double rr = dx * dx + dy * dy + dz * dz;
        const double rs = ri + pp[t].w;
        // bitwise & keeps the tests branch-free (no load sunk into a branch)
        const bool far = (rr > p.acut2) | (rr < 1e-10);
        const bool near = rr < rs * rs * 1.000001;
        prn_l |= (uint32_t)(had & far) << k;
        pc_l |= (uint32_t)(had & near) << k;
    }
    uint32_t prune = 0u;
    if (__any(prn_l != 0u)) prune = grp_or(prn_l);  // wave-uniform; pruning is rare
    uint32_t occ = occ_in & ~prune;
    uint32_t bnd = bnd_in & ~prune;
    otnz &= ~prune;
    int nc_out = nc;
    if (prune) {  // uniform within the group; rare
        // A pruned partner may come back within acut before the next list
        // rebuild: it needs a candidate entry.  Entries that already name
        // the slot become candidates again (their link is cleared below);
        // the others are appended.
        uint32_t has_l = 0u;
#pragma unroll
        for (int t = 0; t < kPerLane; ++t) {
            const int c = l + kLanes * t;
            if ((c < nc) & (c < C) & (cs[t] >= 0)) has_l |= 1u << cs[t];
        }
        const uint32_t addm = prune & ~grp_or(has_l);
#pragma unroll
        for (int t = 0; t < kPerLane; ++t) {
            const int k = l + kLanes * t;
            if (!((addm >> k) & 1u)) continue;
            const int e = nc + __popc(addm & ((1u << k) - 1u));
            if (e < C) {
                s.cand[ic * C + e] = jk[t];
                s.cslot[ic * C + e] = -1;
            }
        }
        nc_out = nc + __popc(addm);
        if (nc_out > C) {  // list full: rebuild the lists before the next step
            nc_out = C;
            if (l == 0) ctrl->rebuild = 1;
        }
        // Candidates whose slot was pruned are absent again.  (The entries
        // the append below links to a slot are others: they had no slot.)
#pragma unroll
        for (int t = 0; t < kPerLane; ++t) {
            const int c = l + kLanes * t;
            if (c < nc && c < C && cs[t] >= 0 && ((prune >> cs[t]) & 1u)) s.cslot[ic * C + c] = -1;
        }
    }

    // ---- 1b. append (d_ReConstructInteractionListbpm)
    const double fx = floor(pi.x * p.inv_cell), fy = floor(pi.y * p.inv_cell),
                 fz = floor(pi.z * p.inv_cell);
    const double sx = copysign(1.0, pi.x - (fx + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    const double sy = copysign(1.0, pi.y - (fy + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    const double sz = copysign(1.0, pi.z - (fz + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    uint32_t spk = 0u;   // 4 bits per entry t: the stencil position of a new pair
    uint32_t nw_l = 0u;  // bit t: candidate l + 8t is a new pair
    uint32_t npc = 0u;   // bit t: the new pair may touch
    uint32_t crem = cab;
#pragma unroll 1
    for (;;) {  // one candidate per lane and round, ascending t
        const double4 pj = pcj;
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double rr = dx * dx + dy * dy + dz * dz;
        const double gx = floor(pj.x * p.inv_cell), gy = floor(pj.y * p.inv_cell),
                     gz = floor(pj.z * p.inv_cell);
        const int bx = gx == fx ? 0 : (gx == fx + sx ? 1 : -1);
        const int by = gy == fy ? 0 : (gy == fy + sy ? 1 : -1);
        const int bz = gz == fz ? 0 : (gz == fz + sz ? 1 : -1);
        const bool nw = (tc < kPerLane) & (rr < p.acut2) & (rr > 1e-10) & (bx >= 0) &
                        (by >= 0) & (bz >= 0);
        const uint64_t sp = stencil_pos(bx | (by << 1) | (bz << 2));
        const double rs = ri + pj.w;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u)
            if (u == tc && nw) spk |= (uint32_t)sp << (4 * u);
        nw_l |= (uint32_t)nw << tc;
        npc |= (uint32_t)(nw & (rr < rs * rs * 1.000001)) << tc;
        crem &= crem - 1u;  // clears bit tc
        if (!__any(crem != 0u)) break;
        tc = crem ? __builtin_ctz(crem) : kPerLane;
        jc = (int32_t)ic;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u)
            if (u == tc) jc = cj[u];
        pcj = pos[jc];
    }
    uint32_t nbt[kPerLane];
    uint32_t anynew = 0u;
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        nbt[t] = grp_ballot((nw_l >> t) & 1u);
        anynew |= nbt[t];
    }
    if (anynew) {  // uniform within the group; rare
        // Rank the new pairs by visiting order (stencil cell, then id) and
        // give them the free slots in the reference's fill order.
        int rank[kPerLane] = {0, 0, 0, 0};
        int total = 0;
        // visiting-order key of each new pair: (stencil position, caller id)
        uint64_t key[kPerLane];
#pragma unroll
        for (int t = 0; t < kPerLane; ++t)
            key[t] = ((nw_l >> t) & 1u)
                         ? ((uint64_t)((spk >> (4 * t)) & 15u) << 32) | (uint32_t)s.orig[cj[t]]
                         : 0ull;
#pragma unroll
        for (int t2 = 0; t2 < kPerLane; ++t2) {
            uint32_t m = nbt[t2];
            total += __popc(m);
            while (m) {
                const int L = __builtin_ctz(m);
                m &= m - 1u;
                const uint64_t kb = shfl_u64(key[t2], L);
#pragma unroll
                for (int t = 0; t < kPerLane; ++t)
                    if (((nw_l >> t) & 1u) && kb < key[t]) ++rank[t];
            }
        }
        const uint32_t freem = ~occ & (S >= 32 ? 0xffffffffu : ((1u << S) - 1u));
        const bool q0 = (p.quirks & RF_QUIRK_SLOT0_SENTINEL) && i == p.zero_idx;
        if (total > __popc(freem) && l == 0) atomicOr(&ctrl->err, kErrRowOverflow);
        uint32_t add_l = 0u, pcadd_l = 0u;
#pragma unroll
        for (int t = 0; t < kPerLane; ++t) {
            if (!((nw_l >> t) & 1u)) continue;
            const int sl = free_slot(freem, rank[t], q0);
            if (sl < 0) continue;
            s.nbr[ic * S + sl] = cj[t];
            s.cslot[ic * C + l + kLanes * t] = (int8_t)sl;
            add_l |= 1u << sl;
            if ((npc >> t) & 1u) pcadd_l |= 1u << sl;
        }
        const uint32_t added = grp_or(add_l);
        pc_l |= grp_or(pcadd_l);
        // the new row entries are read back by the contact rounds below
        __threadfence_block();
        occ |= added;
    }
    uint32_t pcm = 0u;
    if (__any(pc_l != 0u)) pcm = grp_or(pc_l);  // wave-uniform

    // ---- 2a. bonds: lane l owns bonds l + 8t (d_CalculateForcesDEM_22)
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    uint32_t brk_l = 0u;
    {
        const bool live = pre0 && ((bnd >> bs0) & 1u);
        double4 pe = pi, ve = vi, we = wi;  // the bond's own particle
        if (spread) {
            pe = shfl_d4(pi, from);
            ve = shfl_d4(vi, from);
            we = shfl_d4(wi, from);
        }
        double c[6];
        double2 h[6];
        const bool brk = bond_eval(p, pe, ve, we, pb0, vb0, wb0, rec0, c, h);
        if (live) {
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c[u];
            if (brk) {
                brk_l |= 1u << bs0;  // the entry is dead from now on
            } else {
                double2* wp = reinterpret_cast<double2*>(s.brec + qb0);
#pragma unroll
                for (int u = 0; u < 6; ++u) wp[u] = h[u];
            }
        }
        if (spread) {
            // a spread bond: history written by its evaluating lane, the
            // contributions and the break flag handed back to the owner
            const uint32_t bndo = (uint32_t)__shfl((int)bnd, from);
            if (bsx >= 0 && ((bndo >> bsx) & 1u) && !brk) {
                double2* wp = reinterpret_cast<double2*>(s.brec + qbx);
#pragma unroll
                for (int u = 0; u < 6; ++u) wp[u] = h[u];
            }
            const int tl = to >= 0 ? to : lane;
            double c1[6];
#pragma unroll
            for (int u = 0; u < 6; ++u) c1[u] = __shfl(c[u], tl);
            const int brk1 = __shfl((int)brk, tl);
            if (pre1 && ((bnd >> bs1) & 1u)) {
#pragma unroll
                for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c1[u];
                if (brk1) brk_l |= 1u << bs1;
            }
        }
    }
#pragma unroll 1
    for (int t = spread ? 2 : 1; t < kPerLane; ++t) {  // particles with more than 8 bonds
        const int b = l + kLanes * t;
        if (!__any(b < nbd)) break;  // wave-uniform
        const int64_t qb = ic * B + (b < B ? b : B - 1);
        // bond l + 8's slot and partner came with the first loads
        const int bks = !(b < nbd && b < B) ? -1 : t == 1 ? bs1 : (int)s.bslot[qb];
        const int32_t bjr = t == 1 ? bp1 : s.bpart[qb];
        const bool live = bks >= 0 && ((bnd >> bks) & 1u);
        const int32_t bj = live ? bjr : (int32_t)ic;
        const double4 pj = pos[bj];
        const double4 vj = ld3(vel, bj);
        const double4 wj = ld3(angv, bj);
        double2 rec[8];
        const double2* rp = reinterpret_cast<const double2*>(s.brec + (live ? qb : ic * B));
#pragma unroll
        for (int u = 0; u < 8; ++u) rec[u] = rp[u];
        double c[6];
        double2 h[6];
        const bool brk = bond_eval(p, pi, vi, wi, pj, vj, wj, rec, c, h);
        if (live) {
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c[u];
            if (brk) {
                brk_l |= 1u << bks;
            } else {
                double2* wp = reinterpret_cast<double2*>(s.brec + qb);
#pragma unroll
                for (int u = 0; u < 6; ++u) wp[u] = h[u];
            }
        }
    }
    const double Fbx = tree8(acc[0]), Fby = tree8(acc[1]), Fbz = tree8(acc[2]);
    const double Mbx = tree8(acc[3]), Mby = tree8(acc[4]), Mbz = tree8(acc[5]);
    if (__any(brk_l != 0u)) bnd &= ~grp_or(brk_l);  // wave-uniform; breaks are rare

    // ---- 2b. contacts: lane l owns slots l + 8t (d_CalculateForcesDEM_21)
    const double wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
    const bool roll = wim > 1e-18;
    const uint32_t cmask = occ & ~bnd;
    uint32_t mine = 0u;  // bit t: slot l + 8t needs the exact overlap test
#pragma unroll
    for (int t = 0; t < kPerLane; ++t)
        if (((cmask & pcm) >> (l + kLanes * t)) & 1u) mine |= 1u << t;
#pragma unroll
    for (int u = 0; u < 6; ++u) acc[u] = 0.0;
    uint32_t ovl_l = 0u;
#pragma unroll 1
    while (__any(mine != 0u)) {  // one slot per lane and round, ascending t
        const int t = mine ? __builtin_ctz(mine) : kPerLane;
        mine &= mine - 1u;
        const bool act = t < kPerLane;
        const int k = l + kLanes * (act ? t : 0);
        const int64_t qs = ic * S + k;
        // the slot's partner, read back from the row (which holds this
        // step's appends) rather than kept in registers since stage 1
        const int32_t j = act ? s.nbr[qs] : (int32_t)ic;
        const double4 pj = pos[j];
        const double4 vj = vel[j];
        const double4 wj = ld3(angv, j);
        const double4 ot = s.ot[qs];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const bool ov = act && sqrt(dx * dx + dy * dy + dz * dz) < ri + pj.w;
        if (ov) {
            const double4 oprev = ((otnz >> k) & 1u) ? ot : make_double4(0.0, 0.0, 0.0, 0.0);
            double c[6];
            double4 onew;
            contact_eval(p, pi, vi, wi, pj, vj, wj, oprev, wim, roll, c, onew);
            s.ot[qs] = onew;
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c[u];
            ovl_l |= 1u << k;
        }
    }
    // A wave without a touching pair sums nothing but +0.0: its sums are
    // +0.0 without the reductions.
    double Fcx = 0.0, Fcy = 0.0, Fcz = 0.0, Mcx = 0.0, Mcy = 0.0, Mcz = 0.0;
    if (__any(ovl_l != 0u)) {  // wave-uniform
        otnz = (otnz & ~cmask) | grp_or(ovl_l);
        Fcx = tree8(acc[0]); Fcy = tree8(acc[1]); Fcz = tree8(acc[2]);
        Mcx = tree8(acc[3]); Mcy = tree8(acc[4]); Mcz = tree8(acc[5]);
    } else {
        otnz &= ~cmask;
    }
    if ((p.quirks & RF_QUIRK_ROLL_RESET) && !roll) {
        Mcx = 0.0; Mcy = 0.0; Mcz = 0.0;
    }

    // ---- 3. totals, platens, kick-drift, cut (lane 0 of the group)
    double gb = -0.0, gt = -0.0;  // the particle's platen forces (lane 0; -0.0 past the last)
    int gn = 0;                   // and its bonds
    if (valid && l == 0) {
        const float4 rb = s.rbuild[ic];  // the rebuild mark, used last
        const double mi = vi.w;
        // (dt/m, dt/I) as the context formed them at creation, from the mass
        // and radius the particle carries: the same IEEE operations, so the
        // same bits as reading them back, without the memory round trip
        const double2 pk = make_double2((1.0 / mi) * p.dt, p.dt / (0.4 * mi * ri * ri));
        double part_b = -0.0, part_t = -0.0;
        double Fx = 0.0, Fy = 0.0, Fz = 0.0, Mx = 0.0, My = 0.0, Mz = 0.0;
        Fx += Fbx; Fy += Fby; Fz += Fbz;
        Mx += Mbx; My += Mby; Mz += Mbz;
        Fx += Fcx; Fy += Fcy; Fz += Fcz;
        Mx += Mcx; My += Mcy; Mz += Mcz;
        if (p.kind != RF_TEST_NONE) {
            double px_, py_, pz_, r2;
            if (p.kind == RF_TEST_BRAZILIAN) {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y; pz_ = pi.z - p.scope_c[2];
                r2 = px_ * px_ + pz_ * pz_;
            } else {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y - p.scope_c[1]; pz_ = pi.z;
                r2 = px_ * px_ + py_ * py_;
            }
            double s_b = 0.0, s_t = 0.0;
            if (r2 < p.rr_scope && pz_ < p.zcut) {
                const int ax = p.axis;
                const double ppos = ax == 2 ? pz_ : py_;
                double o0 = 0.0, o1 = 0.0, o2 = 0.0;
                WallOut wo;
                if (top - ppos < ri) {
                    const double nwy = ax == 1 ? 1.0 : 0.0, nwz = ax == 2 ? 1.0 : 0.0;
                    const double wvy = ax == 1 ? vw : 0.0, wvz = ax == 2 ? vw : 0.0;
                    wall_contact(p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                 fabs(top - ppos), o0, o1, o2, wo);
                    s_t += ax == 2 ? wo.fnz : wo.fny;
                    Fx += wo.fnx + wo.ftx; Fy += wo.fny + wo.fty; Fz += wo.fnz + wo.ftz;
                    Mx += wo.mtx + wo.max_; My += wo.mty + wo.may; Mz += wo.mtz + wo.maz;
                    o0 = wo.ox; o1 = wo.oy; o2 = wo.oz;
                }
                if (ppos - bottom < ri) {
                    const double nwy = ax == 1 ? -1.0 : 0.0, nwz = ax == 2 ? -1.0 : 0.0;
                    const double wvy = ax == 1 ? -vw : 0.0, wvz = ax == 2 ? -vw : 0.0;
                    wall_contact(p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                 fabs(bottom - ppos), o0, o1, o2, wo);
                    s_b += ax == 2 ? wo.fnz : wo.fny;
                    Fx += wo.fnx + wo.ftx; Fy += wo.fny + wo.fty; Fz += wo.fnz + wo.ftz;
                    Mx += wo.mtx + wo.max_; My += wo.mty + wo.may; Mz += wo.mtz + wo.maz;
                }
            }
            part_b = -s_b;
            part_t = -s_t;
        }
        const double nvx = vi.x + Fx * pk.x, nvy = vi.y + Fy * pk.x, nvz = vi.z + Fz * pk.x;
        double nrx = pi.x + nvx * p.dt, nry = pi.y + nvy * p.dt, nrz = pi.z + nvz * p.dt;
        const double nwx_ = wi.x + Mx * pk.y, nwy_ = wi.y + My * pk.y, nwz_ = wi.z + Mz * pk.y;
        double ovx = nvx, ovy = nvy, ovz = nvz;
        bool out = false;
        {
            const double dx = nrx - p.cen[0], dy = nry - p.cen[1], dz = nrz - p.cen[2];
            if (dx - p.half[0] > 0.0 || dx + p.half[0] < 0.0) out = true;
            if (dy - p.half[1] > 0.0 || dy + p.half[1] < 0.0) out = true;
            if (dz - p.half[2] > 0.0 || dz + p.half[2] < 0.0) out = true;
        }
        if (out) {
            nrx = p.hid[0]; nry = p.hid[1]; nrz = p.hid[2];
            ovx = 0.0; ovy = 0.0; ovz = 0.0;
            Fx = 0.0; Fy = 0.0; Fz = 0.0;
        }
        const int nxt = cur ^ 1;
        s.pos[nxt][i] = make_double4(nrx, nry, nrz, ri);
        s.vel[nxt][i] = make_double4(ovx, ovy, ovz, mi);
        s.angv[nxt][i] = make_double4(nwx_, nwy_, nwz_, 0.0);
        if (p.store_fm) {
            s.F[i] = make_double4(Fx, Fy, Fz, 0.0);
            s.M[i] = make_double4(Mx, My, Mz, 0.0);
        }
        s.occ[i] = occ;
        s.bnd[i] = bnd;
        s.otnz[i] = otnz;
        if (nc_out != nc) s.ncand[i] = nc_out;
        // a particle that has just been moved to the hidden point changes the
        // lists; one that sits there already (and stays: F = V = 0) does not
        const bool was_hidden = (pi.x == p.hid[0]) & (pi.y == p.hid[1]) & (pi.z == p.hid[2]);
        if ((out & !was_hidden) || (!was_hidden && rebuild_due(rb, nrx, nry, nrz)))
            ctrl->rebuild = 1;
        gb = part_b;
        gt = part_t;
        gn = __popc(bnd);
    }

    // Platen partial of the wave's 8 particles, tree v[k] = v[k] + v[k+off],
    // off = 4, 2, 1 (particle k's values in lane 8 k: shuffles by 32, 16 and
    // 8 lanes); lane 0 then holds the wave's W, and the controller adds the
    // group's four as (W0 + W2) + (W1 + W3).
#pragma unroll
    for (int off = 32; off >= kLanes; off >>= 1) {
        gb = gb + __shfl_down(gb, off);
        gt = gt + __shfl_down(gt, off);
        gn += __shfl_down(gn, off);
    }
    if (threadIdx.x == 0) {
        s.part_b[wave] = gb;
        s.part_t[wave] = gt;
        s.part_bonds[wave] = gn;
        const int g = wave / kWavesPerGroup;
        if (gb != 0.0 || gt != 0.0) atomicOr(&s.pmask[g >> 5], 1u << (g & 31));
    }
}

__global__ __launch_bounds__(kBlock) RF_STEP_ATTR void k_step(DevState s, DevParams p,
                                                            Ctrl* __restrict__ ctrl) {
    step_body(s, p, ctrl, xcd_block((int)blockIdx.x, (int)gridDim.x));
}

// Many samples in one launch: logical particle group G = L / 4 runs group
// blk_local[G] of sample blk_sample[G] (its four waves are L = 4 G .. 4 G + 3),
// so every sample runs exactly as it would alone; the host chooses the order
// of the table (the batch autotuner), which never changes a result.
__global__ __launch_bounds__(kBlock) RF_BATCH_ATTR void k_step_batch(
    const BatchEntry* __restrict__ e, const int32_t* __restrict__ blk_sample,
    const int32_t* __restrict__ blk_local) {
    const int L = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int G = L / kWavesPerGroup;
    const BatchEntry& b = e[blk_sample[G]];
    step_body(b.s, b.p, b.ctrl, blk_local[G] * kWavesPerGroup + L % kWavesPerGroup);
}

// -------------------------------------------------------- controller ------
// SumForcesULoading / SumForcesBLoading (md_results.cu:56-145) and
// CompressU_S3 / CompressB_S3 (md_data_types.h:347-379, 462-494), on device.
// Platen-force sum: kCtrlThreads strided running sums of the block partials,
// then a kCtrlThreads-wide tree (the CPU restatement uses the same order).
// The controller proper, on one thread: SumForcesULoading / SumForcesBLoading
// (md_results.cu:56-145) on the summed platen forces and bond count, the
// curve point, and CompressU_S3 / CompressB_S3 (md_data_types.h:347-379,
// 462-494).
// Whether `step` records a curve point and bond count (every dnsave steps).
// A 64-bit remainder is a long software routine; the values met in practice
// fit 32 bits, where the remainder is the same.
__device__ __forceinline__ bool sample_step(int64_t step, int64_t dnsave) {
    const int64_t dn = dnsave > 0 ? dnsave : 1;
    if (dn == 1) return true;
    if ((uint64_t)step <= 0xffffffffull && (uint64_t)dn <= 0xffffffffull)
        return (uint32_t)step % (uint32_t)dn == 0u;
    return step % dn == 0;
}

__device__ __forceinline__ void ctrl_logic(const DevState& s, const DevParams& p,
                                           Ctrl* __restrict__ ctrl, const Ctrl& c0, double sum_b,
                                           double sum_t, long long sum_n, bool save,
                                           const VwLog* lg) {
    Ctrl c = c0;
    if (p.kind != RF_TEST_NONE) {
        c.sfl0 = sum_b;
        c.sfl1 = sum_t;
        if (save) {  // sample_step(c.step, p.dnsave)
            if (c.stepsave < p.curve_cap) {
                s.bcount[c.stepsave] = sum_n;
                s.curve[3 * c.stepsave] = (c.size_axis - c.size0) / c.size0;
                s.curve[3 * c.stepsave + 1] = c.sfl0 * p.inv_area;
                s.curve[3 * c.stepsave + 2] = c.sfl1 * p.inv_area;
            }
            c.last_bonds = sum_n;
            if (fabs(c.vw) > 1e-9) ++c.stepsave;
        }
        c.stress = 0.5 * (fabs(c.sfl0) + fabs(c.sfl1)) * p.inv_area;
        if (!c.loading && c.stress > p.min_stress) c.loading = 1;
        if (p.kind == RF_TEST_BRAZILIAN) {
            if (c.peak_stress < c.stress || c.stress < p.cond_fracture * c.peak_stress) {
                c.peak_stress = c.stress;
                c.peak_strain = (c.size0 - c.size_axis) / c.size0;
            }
            if (c.loading && c.stress < p.zero_stress) c.fracture = 1;
        } else {
            if (c.peak_stress < c.stress) {
                c.peak_stress = c.stress;
                c.peak_strain = (c.size0 - c.size_axis) / c.size0;
            }
            if (c.loading && (c.stress < p.zero_stress || c.stress < p.cond_fracture * c.peak_stress))
                c.fracture = 1;
        }
        // platen controller
        const double pS = p.kind == RF_TEST_BRAZILIAN ? 0.5 * (-c.sfl0 + c.sfl1)
                                                      : 0.5 * fabs(-c.sfl0 + c.sfl1);
        const double pD = fabs(c.sfl0 + c.sfl1);
        if (pS * p.inv_area < 1e-9) {
            c.vw = 10.0 * p.v_load;
        } else if (pD > 0.20 * pS && pS * p.inv_area > 2.0e+6 / 1e9) {
            if (fabs(c.vw) > 1e-15) c.steps_stop = 0;
            c.vw = 0;
            ++c.steps_stop;
            if (c.steps_stop > 30000) c.fracture = 1;
        } else {
            c.vw = p.v_load;
        }
        if (lg && lg->mode) {  // frozen-branch window: the velocity recorded / replayed
            const int64_t ls = c.step - lg->step0;
            if (ls >= 0 && ls < lg->steps) {
                if (lg->mode == 1)
                    lg->vw[ls] = c.vw;
                else
                    c.vw = lg->vw[ls];
            }
        }
        c.size_axis += c.vw * p.dt;
        c.top += 0.5 * c.vw * p.dt;
        c.bottom -= 0.5 * c.vw * p.dt;
        c.strain = (c.size_axis - c.size0) / c.size0;
    }
    c.step += 1;
    c.cur ^= 1;
    if (p.kind != RF_TEST_NONE) {
        const bool go = (p.max_steps <= 0 || c.step < p.max_steps) &&
                        (fabs(c.strain) < p.max_strain && !c.fracture);
        if (!go) c.done = 1;
    }
    c.halt = (c.done || c.rebuild || c.err || (c.stop_at > 0 && c.step >= c.stop_at)) ? 1 : 0;
    *ctrl = c;
}

// Work a controller variant adds between the partial sums' loads and their
// tree (every thread of the block calls it, with the step's control block).
struct NoMid {
    __device__ __forceinline__ void operator()(const Ctrl&) const {}
};

// kQ: groups a thread has in flight at once (the registers of 2 kQ double4
// loads).
template <int kQ = 4, class Mid = NoMid>
__device__ __forceinline__ void ctrl_body(const DevState& s, const DevParams& p,
                                          Ctrl* __restrict__ ctrl, const VwLog* lg = nullptr,
                                          const Mid& mid = Mid()) {
    __shared__ double sb[kCtrlThreads];
    __shared__ double stt[kCtrlThreads];
    __shared__ long long sn[kCtrlThreads];
    const int t = threadIdx.x;
    const int nblk = p.nblk;
    const double4* __restrict__ pb = reinterpret_cast<const double4*>(s.part_b);
    const double4* __restrict__ pt = reinterpret_cast<const double4*>(s.part_t);
    // the partials are read before the halt flag is known (a halted step
    // leaves them as they were), so these loads and the control block's
    // (uniform: scalar loads) are in flight together; nothing in this
    // kernel writes the control block before its last store
    const Ctrl c0 = *ctrl;
    // Running sum t adds groups t, t + kCtrlThreads, ... in ascending order.
    // Group b's partial: its four waves' entries as (W0 + W2) + (W1 + W3).
    // Only the groups the step flagged are read: the others' partials are
    // zeros, and adding a zero leaves these sums' bits unchanged (they start
    // at +0.0, so they are never -0.0).  The flags of kQ groups per thread
    // are read together (unconditionally, so that no load waits inside a
    // branch), then their flagged partials together.
    double ab = 0.0, at = 0.0;
    for (int m0 = 0; kCtrlThreads * m0 < nblk; m0 += kQ) {  // uniform
        uint32_t fm = 0u;  // bit q: group t + kCtrlThreads (m0 + q) flagged
#pragma unroll
        for (int q = 0; q < kQ; ++q) {  // (past the last group: word 0, no flag)
            const int b = t + kCtrlThreads * (m0 + q);
            const int bc = b < nblk ? b : 0;
            const uint32_t word = s.pmask[bc >> 5];
            fm |= ((word >> (bc & 31)) & (uint32_t)(b < nblk)) << q;
        }
        // all flags in before the first partial load: the in-order load
        // counter would otherwise make a later flag wait for earlier partials
        __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0)
        double4 vb[kQ], vt[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            if ((fm >> q) & 1u) {
                const int b = t + kCtrlThreads * (m0 + q);
                vb[q] = pb[b];
                vt[q] = pt[b];
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            if ((fm >> q) & 1u) {
                ab = ab + ((vb[q].x + vb[q].z) + (vb[q].y + vb[q].w));
                at = at + ((vt[q].x + vt[q].z) + (vt[q].y + vt[q].w));
            }
        }
    }
    // bond counts: only on the steps that record a bond-count sample
    const bool save = sample_step(c0.step, p.dnsave);
    long long nb = 0;
    if (p.kind != RF_TEST_NONE && save) {
#pragma unroll 4
        for (int b = t; b < nblk; b += kCtrlThreads) {
            const int4 vn = reinterpret_cast<const int4*>(s.part_bonds)[b];
            nb += (vn.x + vn.y) + (vn.z + vn.w);
        }
    }
    __syncthreads();  // every flag read: clear them for the next step
    for (int q = t; q < (nblk + 31) / 32; q += kCtrlThreads) s.pmask[q] = 0u;
    if (c0.halt) return;
    mid(c0);
    // tree v[t] = v[t] + v[t + off], off = kCtrlThreads / 2 .. 1.  The
    // levels down to 64 values are formed by wave 0 alone: lane t reads the
    // 16 sums t + 64 k and adds them in the tree's pairing, so the block
    // needs one barrier; then one shuffle per level.
    static_assert(kCtrlThreads == 1024, "the tree below is written for 1024 sums");
    sb[t] = ab;
    stt[t] = at;
    sn[t] = nb;
    __syncthreads();
    if (t >= 64) return;
    double xb[16], xt[16];
    long long xn[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        xb[k] = sb[t + 64 * k];
        xt[k] = stt[t + 64 * k];
        xn[k] = sn[t + 64 * k];
    }
    // x[k] holds v[t + 64 k]; level off = 64 m pairs x[k] with x[k + m]
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) {
#pragma unroll
        for (int k = 0; k < m; ++k) {
            xb[k] = xb[k] + xb[k + m];
            xt[k] = xt[k] + xt[k + m];
            xn[k] += xn[k + m];
        }
    }
    double vb = xb[0], vt = xt[0];
    long long vn = xn[0];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        vb = vb + __shfl_down(vb, off);
        vt = vt + __shfl_down(vt, off);
        vn += __shfl_down(vn, off);
    }
    if (t != 0) return;
    ctrl_logic(s, p, ctrl, c0, vb, vt, vn, save, lg);
}

__global__ __launch_bounds__(kCtrlThreads) void k_controller(DevState s, DevParams p,
                                                             Ctrl* __restrict__ ctrl) {
    ctrl_body(s, p, ctrl);
}

// Tangents of the platen forces where the value controller records a curve
// point: per direction the block partials summed strided over the block's
// threads and by the tree through LDS and shuffles, stored beside the point.
// Runs inside the value controller, after its partial sums' loads and
// before it advances the step and curve index.
struct SensTangents {
    const DevParams& p;
    const SensState& t;
    __device__ __forceinline__ void operator()(const Ctrl& c0) const {
        __shared__ double tb[kCtrlThreads], tt[kCtrlThreads];
        if (p.kind == RF_TEST_NONE) return;
        const int64_t save = c0.stepsave;
        if (!sample_step(c0.step, p.dnsave) || save >= p.curve_cap) return;  // uniform
        const int tid = threadIdx.x;
        for (int k = 0; k < t.nl; ++k) {
            double ab = 0.0, at = 0.0;
            for (int b = tid; b < p.nblk; b += kCtrlThreads) {
                ab = ab + t.tpart_b[k * p.nblk + b];
                at = at + t.tpart_t[k * p.nblk + b];
            }
            tb[tid] = ab;
            tt[tid] = at;
            __syncthreads();
            for (int off = kCtrlThreads / 2; off >= 64; off >>= 1) {
                if (tid < off) {
                    tb[tid] = tb[tid] + tb[tid + off];
                    tt[tid] = tt[tid] + tt[tid + off];
                }
                __syncthreads();
            }
            if (tid < 64) {
                double vb = tb[tid], vt = tt[tid];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    vb = vb + __shfl_down(vb, off);
                    vt = vt + __shfl_down(vt, off);
                }
                if (tid == 0) {
                    t.tcurve[(save * kSensK + k) * 2] = vb * p.inv_area;
                    t.tcurve[(save * kSensK + k) * 2 + 1] = vt * p.inv_area;
                }
            }
            __syncthreads();
        }
    }
};

// Controller of a sensitivity step: the value controller with the
// tangents of a curve point.
__global__ __launch_bounds__(kCtrlThreads) void k_controller_sens(DevState s, DevParams p,
                                                                  SensState t,
                                                                  Ctrl* __restrict__ ctrl) {
    const VwLog lg{t.br_mode, 0, t.br_step0, t.br_steps, t.br_vw};
    ctrl_body(s, p, ctrl, &lg, SensTangents{p, t});
}

// One block per sample.
__global__ __launch_bounds__(kCtrlThreads) void k_controller_batch(const BatchEntry* __restrict__ e) {
    const BatchEntry& b = e[blockIdx.x];
    ctrl_body<2>(b.s, b.p, b.ctrl);  // 2: fewer registers
}

// ------------------------------------------------------ cell binning ------
// d_CalculateCellIndex (md_calculate_cell.cu:13-41) plus a grid check the
// reference only prints under pre_debugtest.
// Whether cell (fx, fy, fz) lies in 1..dims-2 along every axis, so that its
// whole 27-cell neighbourhood is in the grid.  A particle outside is flagged
// (kErrGrid) by k_cell_keys, and the kernels that scan a stencil skip it;
// the host then reports RF_ERR_GRID.
__device__ __forceinline__ bool in_grid(const DevParams& p, double fx, double fy, double fz) {
    return fx >= 1.0 && fx <= p.nx - 2 && fy >= 1.0 && fy <= p.ny - 2 && fz >= 1.0 &&
           fz <= p.nz - 2;
}

__global__ void k_cell_keys(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 r = s.pos[ctrl->cur][i];
    const double fx = floor(r.x * p.inv_cell), fy = floor(r.y * p.inv_cell),
                 fz = floor(r.z * p.inv_cell);
    if (!in_grid(p, fx, fy, fz)) {
        atomicOr(&ctrl->err, kErrGrid);
        s.ckey[i] = 0;
        return;
    }
    s.ckey[i] = (int32_t)fx + (int32_t)fy * p.nx + (int32_t)fz * p.nx * p.ny;
}

__global__ void k_zero_i32(int32_t* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_cell_count(DevState s, DevParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p.n) atomicAdd(&s.ccount[s.ckey[i]], 1);
}

// Exclusive scan of the cell counts (one block); also clears the fill cursors.
__global__ __launch_bounds__(1024) void k_cell_scan(DevState s, DevParams p) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (p.ncell + 1023) / 1024;
    const int64_t lo = t * per, hi = lo + per < p.ncell ? lo + per : p.ncell;
    int32_t acc = 0;
    for (int64_t q = lo; q < hi; ++q) acc += s.ccount[q];
    part[t] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - acc;
    for (int64_t q = lo; q < hi; ++q) {
        s.cstart[q] = run;
        run += s.ccount[q];
        s.cfill[q] = 0;
    }
}

__global__ void k_cell_scatter(DevState s, DevParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int32_t key = s.ckey[i];
    const int32_t at = s.cstart[key] + atomicAdd(&s.cfill[key], 1);
    s.ctmp[at] = (int32_t)i;
}

// Within a cell the particles are put in ascending order of the caller's
// ids: the stable (cell, id) order of cub::DeviceRadixSort::SortPairs over
// identity values (md_cell_distribution.cu:87).  Rank = number of smaller
// ids in the cell.
__global__ void k_cell_rank(DevState s, DevParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int32_t key = s.ckey[i];
    const int32_t st = s.cstart[key], ct = s.ccount[key];
    const int32_t oi = s.orig[i];
    int32_t rank = 0;
    for (int32_t q = 0; q < ct; ++q) rank += s.orig[s.ctmp[st + q]] < oi ? 1 : 0;
    s.sorted[st + rank] = (int32_t)i;
}

// ----------------------------------------------------- t = 0 set-up -------
// d_ConstructInteractionList + addlink (md_interaction_list_construct.cu:17-130).
__global__ void k_construct_rows(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[i];
    const double fx = floor(pi.x * p.inv_cell), fy = floor(pi.y * p.inv_cell),
                 fz = floor(pi.z * p.inv_cell);
    if (!in_grid(p, fx, fy, fz)) {  // flagged by k_cell_keys: an empty row
        s.occ[i] = 0u;
        s.bnd[i] = 0u;
        s.otnz[i] = 0u;
        return;
    }
    const int64_t nx = p.nx, nxy = (int64_t)p.nx * p.ny;
    const int64_t base = (int64_t)fx + (int64_t)fy * nx + (int64_t)fz * nxy;
    const int64_t sx = copysign(1.0, pi.x - (fx + 0.5) * p.cell) > 0.0 ? 1 : -1;
    const int64_t sy = copysign(1.0, pi.y - (fy + 0.5) * p.cell) > 0.0 ? 1 : -1;
    const int64_t sz = copysign(1.0, pi.z - (fz + 0.5) * p.cell) > 0.0 ? 1 : -1;
    const int64_t cells[8] = {base, base + sx, base + sy * nx, base + sz * nxy,
                              base + sx + sy * nx, base + sx + sz * nxy,
                              base + sy * nx + sz * nxy, base + sx + sy * nx + sz * nxy};
    int cnt = 0;
    for (int c = 0; c < 8; ++c) {
        const int32_t st = s.cstart[cells[c]], ct = s.ccount[cells[c]];
        for (int32_t q = 0; q < ct; ++q) {
            const int32_t j = s.sorted[st + q];
            const double4 pj = pos[j];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr < p.acut2 && j != (int32_t)i) {
                if (cnt >= p.S) {
                    atomicOr(&ctrl->err, kErrRowOverflow);
                    continue;
                }
                s.nbr[i * p.S + cnt] = j;
                ++cnt;
            }
        }
    }
    s.occ[i] = cnt >= 32 ? 0xffffffffu : ((1u << cnt) - 1u);
    s.bnd[i] = 0u;
    s.otnz[i] = 0u;
}

// d_ConstructBoundInteractions (md_boundinteractions_construct.cu:35-106):
// bonds numbered in ascending slot order into the bond table.
__global__ void k_construct_bonds(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[i];
    const uint32_t occ = s.occ[i];
    uint32_t bnd = 0u;
    int nb = 0;
    for (int k = 0; k < p.S; ++k) {
        if (!((occ >> k) & 1u)) continue;
        const int32_t j = s.nbr[i * p.S + k];
        const double4 pj = pos[j];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double rm2 = dx * dx + dy * dy + dz * dz;
        const double bd = pi.w + pj.w + p.bond_gap;
        if (rm2 < bd * bd && rm2 > 1e-18) {
            const int64_t qb = i * p.B + nb;
            s.bslot[qb] = (int8_t)k;
            s.bpart[qb] = j;
            BondRec rec;
            rec.pnx = 0.0; rec.pny = 0.0; rec.pnz = 0.0;
            rec.otx = 0.0; rec.oty = 0.0; rec.otz = 0.0;
            rec.mnx = 0.0; rec.mny = 0.0; rec.mnz = 0.0;
            rec.mtx = 0.0; rec.mty = 0.0; rec.mtz = 0.0;
            rec.invL0 = 1.0 / sqrt(rm2);
            rec.sc = 0.0; rec.tc = 0.0; rec.L0 = 1.0 / rec.invL0;
            s.brec[qb] = rec;
            bnd |= 1u << k;
            ++nb;
        }
    }
    s.bnd[i] = bnd;
    s.nbond[i] = nb;
    s.otnz[i] = 0u;
}

// The first d_DeleteFarLinks (md_calculationuniaxialcompression.cu:212).
__global__ void k_prune_initial(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[i];
    uint32_t occ = s.occ[i];
    uint32_t drop = 0u;
    for (int k = 0; k < p.S; ++k) {
        if (!((occ >> k) & 1u)) continue;
        const double4 pj = pos[s.nbr[i * p.S + k]];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double rr = dx * dx + dy * dy + dz * dz;
        if (rr > p.acut2 || rr < 1e-10) drop |= 1u << k;
    }
    s.occ[i] = occ & ~drop;
    s.bnd[i] &= ~drop;
    s.otnz[i] &= ~drop;
}

// Candidate lists: every j within acut + skin (cells of edge >= 2 acut) that
// is not a current row partner.  Particles at the hidden point neither have
// nor are candidates: they sit at one point outside the box (rho = 0 never
// links, d_ReConstructInteractionListbpm's 1e-10 < rr) and never move again.
// Eight lanes per particle share the cells that can reach acut + skin; each
// lane keeps its hits in LDS and a prefix over the eight counts places them,
// so the list order is fixed (lane by lane, cell by cell, sorted order in a
// cell).  The current partners are held in registers for the exclusion.
// The order of a list never changes a result: new pairs are ranked by
// stencil cell and id when they are appended.
__global__ __launch_bounds__(256) void k_build_candidates(DevState s, DevParams p, Ctrl* ctrl) {
    // each lane's first kMaxCand hits, entry k of thread t at [k * 256 + t]
    __shared__ int32_t hitbuf[kMaxCand * 256];
    const int l = threadIdx.x & (kLanes - 1);
    const int64_t i = (int64_t)blockIdx.x * (256 / kLanes) + (threadIdx.x >> 3);
    const bool valid = i < p.n;
    const int64_t ic = valid ? i : 0;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[ic];
    // particles at the hidden point, and particles outside the grid (flagged
    // by k_cell_keys), get no candidates
    const bool skip = !valid | ((pi.x == p.hid[0]) & (pi.y == p.hid[1]) & (pi.z == p.hid[2])) |
                      !in_grid(p, floor(pi.x * p.inv_cell), floor(pi.y * p.inv_cell),
                               floor(pi.z * p.inv_cell));
    const int64_t key = s.ckey[ic];
    const int64_t nx = p.nx, nxy = (int64_t)p.nx * p.ny;
    const int64_t C = p.C;
    const uint32_t occ = s.occ[ic];
    // the particle's current partners (-1 where a slot is free), held by
    // every lane of the group
    int32_t row[kMaxSlots];
    if (p.S == kMaxSlots) {  // 128-byte rows: eight 16-byte loads
        const int4* rp = reinterpret_cast<const int4*>(s.nbr + ic * kMaxSlots);
#pragma unroll
        for (int q = 0; q < kMaxSlots / 4; ++q) {
            const int4 v = rp[q];
            row[4 * q] = v.x;
            row[4 * q + 1] = v.y;
            row[4 * q + 2] = v.z;
            row[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kMaxSlots; ++k) row[k] = s.nbr[ic * p.S + (k < p.S ? k : p.S - 1)];
    }
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k)
        if (!((occ >> k) & 1u)) row[k] = -1;
    // bit u of the result: entry q + u of the cell is a candidate (four
    // entries per round, their loads issued together)
    auto hits4 = [&](int32_t st, int32_t ct, int32_t q, int32_t (&j)[4]) -> uint32_t {
        double4 pj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = q + u < ct ? s.sorted[st + q + u] : (int32_t)ic;
#pragma unroll
        for (int u = 0; u < 4; ++u) pj[u] = pos[j[u]];
        uint32_t m4 = 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double ex = pj[u].x - pi.x, ey = pj[u].y - pi.y, ez = pj[u].z - pi.z;
            const bool hid = (pj[u].x == p.hid[0]) & (pj[u].y == p.hid[1]) & (pj[u].z == p.hid[2]);
            const bool in = (j[u] != (int32_t)ic) & !hid & (ex * ex + ey * ey + ez * ez < p.cand_r2);
            m4 |= (uint32_t)in << u;
        }
        for (uint32_t b = m4; b; b &= b - 1u) {  // current partners hold a slot already
            const int u = __builtin_ctz(b);
            bool partner = false;
#pragma unroll
            for (int k = 0; k < kMaxSlots; ++k) partner |= row[k] == j[u];
            if (partner) m4 &= ~(1u << u);
        }
        return m4;
    };
    auto cell_of = [&](int c) {
        return key + (c % 3 - 1) + ((c / 3) % 3 - 1) * nx + (c / 9 - 1) * nxy;
    };
    // Cells that can hold a point within R = acut + skin (R <= cell edge):
    // along each axis the own cell, the lower neighbour when the particle
    // is closer than R to the cell's lower face, the upper one likewise
    // (with cell edge 2 acut and skin 0.08 acut, about 9 of the 27).
    const double R = sqrt(p.cand_r2) * (1.0 + 1e-9);
    const double fx = floor(pi.x * p.inv_cell), fy = floor(pi.y * p.inv_cell),
                 fz = floor(pi.z * p.inv_cell);
    const double ox = pi.x - fx * p.cell, oy = pi.y - fy * p.cell, oz = pi.z - fz * p.cell;
    const uint32_t ax = 2u | (uint32_t)(ox < R) | ((uint32_t)(p.cell - ox < R) << 2);
    const uint32_t ay = 2u | (uint32_t)(oy < R) | ((uint32_t)(p.cell - oy < R) << 2);
    const uint32_t az = 2u | (uint32_t)(oz < R) | ((uint32_t)(p.cell - oz < R) << 2);
    uint32_t need = 0u;  // bit c: stencil cell c (offsets c%3-1, (c/3)%3-1, c/9-1)
#pragma unroll
    for (int c = 0; c < 27; ++c)
        if (((ax >> (c % 3)) & (ay >> ((c / 3) % 3)) & (az >> (c / 9))) & 1u) need |= 1u << c;
    // one pass: hits counted and the first kMaxCand kept in LDS; the k-th
    // needed cell goes to lane k % 8, and the list is lane-major (each
    // lane's cells in order), as positions come from the lanes' prefix sum
    int cnt = 0;
    uint32_t m = skip ? 0u : need;
    for (int t = 0; t < l; ++t) m &= m - 1u;  // this lane starts at the l-th needed cell
    while (m) {
        const int c = __builtin_ctz(m);
#pragma unroll
        for (int t = 0; t < kLanes; ++t) m &= m - 1u;  // then every 8th
        const int64_t cc = cell_of(c);
        const int32_t st = s.cstart[cc], ct = s.ccount[cc];
        for (int32_t q = 0; q < ct; q += 4) {
            int32_t j[4];
            for (uint32_t b = hits4(st, ct, q, j); b; b &= b - 1u) {
                if (cnt < kMaxCand) hitbuf[cnt * 256 + threadIdx.x] = j[__builtin_ctz(b)];
                ++cnt;
            }
        }
    }
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < kLanes; off <<= 1) {
        const int t = __shfl_up(incl, off, kLanes);
        if (l >= off) incl += t;
    }
    const int total = __shfl(incl, kLanes - 1, kLanes);
    const int w0 = incl - cnt;
    for (int k = 0; k < cnt && k < kMaxCand && w0 + k < C; ++k) {
        s.cand[ic * C + w0 + k] = hitbuf[k * 256 + threadIdx.x];
        s.cslot[ic * C + w0 + k] = -1;
    }
    if (valid && l == 0) {
        if (total > C) atomicOr(&ctrl->err, kErrCandOverflow);
        s.ncand[i] = total < C ? total : (int32_t)C;
        s.rbuild[i] = rebuild_mark(p, pi);
    }
}

// Per-bond strengths drawn on the host, scattered into the bond records.
__global__ void k_set_strengths(BondRec* brec, const double* sc, const double* tc, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    brec[q].sc = sc[q];
    brec[q].tc = tc[q];
}

__global__ void k_set_stop(Ctrl* ctrl, int64_t stop_at) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->stop_at = stop_at;
        ctrl->halt = (ctrl->done || ctrl->err || ctrl->rebuild ||
                      (stop_at > 0 && ctrl->step >= stop_at)) ? 1 : 0;
    }
}

__global__ void k_clear_flags(Ctrl* ctrl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->rebuild = 0;
        ctrl->halt = (ctrl->done || ctrl->err ||
                      (ctrl->stop_at > 0 && ctrl->step >= ctrl->stop_at)) ? 1 : 0;
    }
}

// Slot statistics for the bytes model: occupied, bonds, contacts in overlap,
// candidates.
__global__ void k_counters(DevState s, DevParams p, Ctrl* ctrl, unsigned long long* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[i];
    uint32_t occ = s.occ[i], bnd = s.bnd[i];
    unsigned long long co = 0;
    for (int k = 0; k < p.S; ++k) {
        if (!((occ >> k) & 1u)) continue;
        const double4 pj = pos[s.nbr[i * p.S + k]];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double rr = dx * dx + dy * dy + dz * dz;
        if (rr > p.acut2 || rr < 1e-10) {  // pruned at the start of the next step
            occ &= ~(1u << k);
            bnd &= ~(1u << k);
            continue;
        }
        const double rs = pi.w + pj.w;
        if (!((bnd >> k) & 1u) && rr < rs * rs) ++co;
    }
    atomicAdd(&out[0], (unsigned long long)__popc(occ));
    atomicAdd(&out[1], (unsigned long long)__popc(bnd));
    atomicAdd(&out[2], co);
    atomicAdd(&out[3], (unsigned long long)s.ncand[i]);
}

// ----------------------------------------------------------- wrappers -----
static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_cell_keys(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_cell_sort(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(p.ncell, 256)), dim3(256), 0, st, s.ccount,
                       p.ncell);
    hipLaunchKernelGGL(k_cell_count, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p);
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, st, s, p);
    hipLaunchKernelGGL(k_cell_scatter, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p);
    hipLaunchKernelGGL(k_cell_rank, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p);
    (void)ctrl;
    return hipGetLastError();
}

hipError_t launch_construct_rows(const DevState& s, const DevParams& p, Ctrl* ctrl,
                                 hipStream_t st) {
    hipLaunchKernelGGL(k_construct_rows, dim3(grid_for(p.n, 128)), dim3(128), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_construct_bonds(const DevState& s, const DevParams& p, Ctrl* ctrl,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_construct_bonds, dim3(grid_for(p.n, 128)), dim3(128), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_prune_initial(const DevState& s, const DevParams& p, Ctrl* ctrl,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_prune_initial, dim3(grid_for(p.n, 128)), dim3(128), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_build_candidates(const DevState& s, const DevParams& p, Ctrl* ctrl,
                                   hipStream_t st) {
    hipLaunchKernelGGL(k_build_candidates, dim3(grid_for(p.n * kLanes, 256)), dim3(256), 0, st, s,
                       p, ctrl);
    return hipGetLastError();
}

hipError_t launch_set_strengths(BondRec* brec, const double* sc, const double* tc, int64_t n,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_set_strengths, dim3(grid_for(n, 256)), dim3(256), 0, st, brec, sc, tc, n);
    return hipGetLastError();
}

hipError_t launch_set_stop(Ctrl* ctrl, int64_t stop_at, hipStream_t st) {
    hipLaunchKernelGGL(k_set_stop, dim3(1), dim3(64), 0, st, ctrl, stop_at);
    return hipGetLastError();
}

hipError_t launch_clear_flags(Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, st, ctrl);
    return hipGetLastError();
}

hipError_t launch_step(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_step, dim3((unsigned)(p.nblk * kWavesPerGroup)), dim3(kBlock), 0, st, s,
                       p, ctrl);
    return hipGetLastError();
}

hipError_t launch_step_timed(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st,
                             hipEvent_t start, hipEvent_t stop) {
    hipExtLaunchKernelGGL(k_step, dim3((unsigned)(p.nblk * kWavesPerGroup)), dim3(kBlock), 0, st,
                          start, stop, 0, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_controller(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(kCtrlThreads), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_step_batch(const BatchEntry* e, const int32_t* blk_sample,
                             const int32_t* blk_local, int32_t nblk_total, hipStream_t st) {
    hipLaunchKernelGGL(k_step_batch, dim3((unsigned)(nblk_total * kWavesPerGroup)), dim3(kBlock),
                       0, st, e, blk_sample, blk_local);
    return hipGetLastError();
}

hipError_t launch_controller_sens(const DevState& s, const DevParams& p, const SensState& t,
                                  Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_controller_sens, dim3(1), dim3(kCtrlThreads), 0, st, s, p, t, ctrl);
    return hipGetLastError();
}

hipError_t launch_controller_batch(const BatchEntry* e, int32_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_controller_batch, dim3((unsigned)count), dim3(kCtrlThreads), 0, st, e);
    return hipGetLastError();
}

hipError_t launch_counters(const DevState& s, const DevParams& p, Ctrl* ctrl, int64_t* out,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_counters, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p, ctrl,
                       (unsigned long long*)out);
    return hipGetLastError();
}

}  // namespace rf
