// rf_kernels.hip -- gfx950 kernels of the bonded-particle DEM loading step.
//
// One 32-lane group per particle (two particles per wave64, eight per
// 256-thread block).  In one launch a group
//   1. prunes and appends its interaction row (slot lanes check the stored
//      partners, candidate lanes look for new ones),
//   2. evaluates its bonds (one lane per bond) and contacts (one lane per
//      slot) and reduces their contributions with a fixed 32-wide tree,
//   3. adds the platen forces, kick-drifts the particle, applies the cut
//      restriction and writes the next state buffer.
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
#include "rf_internal.h"

#include "../../include/rockfit_dem.h"

namespace rf {

#define RF_PI 3.1415926535897932384626433832795
constexpr double kDamp = -1.8257;          // -2 sqrt(5/6), md_calculateforcesdem3.cu:248
constexpr double kTwoThirds = 2.0 / 3.0;   // MCf_2d3

// Position of a stencil cell in the reference's visiting order, indexed by
// (dx | dy << 1 | dz << 2) where d = 1 means "the neighbour on the particle's
// side along that axis": own, x, y, z, xy, xz, yz, xyz.
__constant__ int kStencilPos[8] = {0, 1, 2, 4, 3, 5, 6, 7};

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

struct WallOut {
    double fnx, fny, fnz, ftx, fty, ftz, mtx, mty, mtz, max_, may, maz, ox, oy, oz;
};

// Particle-platen Hertz-Mindlin, dd_Calculate_fijn_pwHM_2 +
// dd_Calculate_fijtmijt_pwHM_2 (md_compression_hm.cu:20-154): R* = r_i,
// m* = m_i, previous tangential overlap (opx, opy, opz) including the
// reference's o.x*o.y rescale typo (:91).
__device__ __forceinline__ void wall_contact(const DevParams& p, double ri, double mi,
                                             const double4& vi, const double4& wi,
                                             double nwx, double nwy, double nwz,
                                             double wvx, double wvy, double wvz, double d,
                                             double opx, double opy, double opz, WallOut& o) {
    double c1 = sqrt(ri * fabs(ri - d));
    const double Kn = 2.0 * p.Ew * c1;
    const double Kt = 8.0 * p.Gw * c1;
    const double vx = wvx - vi.x + ri * (nwy * wi.z - nwz * wi.y);
    const double vy = wvy - vi.y + ri * (nwz * wi.x - nwx * wi.z);
    const double vz = wvz - vi.z + ri * (nwx * wi.y - nwy * wi.x);
    const double vn = nwx * vx + nwy * vy + nwz * vz;
    const double fnm = (d - ri) * Kn * kTwoThirds;
    c1 = kDamp * p.pA * vn * sqrt(Kn * mi);
    o.fnx = nwx * (fnm + c1); o.fny = nwy * (fnm + c1); o.fnz = nwz * (fnm + c1);
    const double vtx = vx - nwx * vn, vty = vy - nwy * vn, vtz = vz - nwz * vn;
    c1 = nwx * opx + nwy * opy + nwz * opz;
    double ox = opx - nwx * c1, oy = opy - nwy * c1, oz = opz - nwz * c1;
    c1 = ox * ox + oy * oy + oz * oz;
    if (c1 > 1e-18) {
        c1 = 1.0 / c1;
        c1 *= opx * opy + opy * opy + opz * opz;
        c1 = sqrt(c1);
        ox *= c1; oy *= c1; oz *= c1;
    }
    ox += vtx * p.dt; oy += vty * p.dt; oz += vtz * p.dt;
    double ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
    const double ftm = ftx * ftx + fty * fty + ftz * ftz;
    if (ftm > p.mu * p.mu * fnm * fnm) {
        c1 = p.mu * fabs(fnm) * (1.0 / sqrt(ftm));
        ftx *= c1; fty *= c1; ftz *= c1;
        c1 = 1.0 / Kt;
        ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
    } else {
        c1 = kDamp * p.pA * sqrt(Kt * mi);
        ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
    }
    o.ox = ox; o.oy = oy; o.oz = oz;
    o.ftx = ftx; o.fty = fty; o.ftz = ftz;
    o.mtx = (nwy * ftz - nwz * fty) * ri;
    o.mty = (nwz * ftx - nwx * ftz) * ri;
    o.mtz = (nwx * fty - nwy * ftx) * ri;
    const double wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
    if (wim > 1e-18) {
        c1 = -p.mur * fabs(fnm) * (1.0 / sqrt(wim));
        o.max_ = wi.x * c1; o.may = wi.y * c1; o.maz = wi.z * c1;
    } else {
        o.max_ = 0.0; o.may = 0.0; o.maz = 0.0;
    }
}

// ------------------------------------------------------------ step kernel --
// One 8-lane group per particle; lane l owns slots, candidates and bonds
// l + 8t (t = 0..3).  Per-particle sums: each lane adds its own entries in
// ascending order, then tree8.  Memory round trips: rows of the particle;
// slot-partner and candidate gathers; per bond iteration the partner's state
// and the 128-byte bond record; the few contacts that can overlap.
__global__ __launch_bounds__(kBlock) void k_step(DevState s, DevParams p, Ctrl* __restrict__ ctrl) {
    __shared__ double sh_b[kPerBlock];
    __shared__ double sh_t[kPerBlock];
    __shared__ int sh_nb[kPerBlock];
    if (ctrl->halt) return;  // uniform for the whole grid
    const int l = threadIdx.x & (kLanes - 1);
    const int grp = threadIdx.x / kLanes;
    const int64_t i = (int64_t)blockIdx.x * kPerBlock + grp;
    const bool valid = i < p.n;
    const int cur = ctrl->cur;
    const double4* __restrict__ pos = s.pos[cur];
    const double4* __restrict__ vel = s.vel[cur];
    const double4* __restrict__ angv = s.angv[cur];
    const double top = ctrl->top, bottom = ctrl->bottom, vw = ctrl->vw;
    const int S = p.S, C = p.C, B = p.B;
    const int64_t ic = valid ? i : 0;  // clamped index for the inactive groups' loads
    const double4 pi = pos[ic];
    const double4 vi = vel[ic];
    const double4 wi = angv[ic];
    const double ri = pi.w, mi = vi.w;
    uint32_t occ = valid ? s.occ[ic] : 0u;
    uint32_t bnd = valid ? s.bnd[ic] : 0u;
    uint32_t otnz = valid ? s.otnz[ic] : 0u;
    const int nc = valid ? s.ncand[ic] : 0;
    const int nbd = valid ? s.nbond[ic] : 0;

    // ---- 1a. prune: each lane re-checks the partners of its slots (d_DeleteFarLinks)
    int32_t jk[kPerLane];
    uint32_t prn_l = 0u, pc_l = 0u;  // this lane's pruned / possibly-touching slots
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int k = l + kLanes * t;
        jk[t] = (k < S) ? s.nbr[ic * S + k] : 0;
    }
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int k = l + kLanes * t;
        const bool had = k < S && ((occ >> k) & 1u);
        const double4 pj = pos[had ? jk[t] : (int32_t)ic];
        if (had) {
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr > p.acut2 || rr < 1e-10) prn_l |= 1u << k;
            const double rs = ri + pj.w;
            if (rr < rs * rs * 1.000001) pc_l |= 1u << k;
        }
    }
    const uint32_t prune = grp_or(prn_l);
    occ &= ~prune;
    bnd &= ~prune;
    otnz &= ~prune;

    // ---- 1b. append: candidates (d_ReConstructInteractionListbpm)
    const double fx = floor(pi.x * p.inv_cell), fy = floor(pi.y * p.inv_cell),
                 fz = floor(pi.z * p.inv_cell);
    const double sx = copysign(1.0, pi.x - (fx + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    const double sy = copysign(1.0, pi.y - (fy + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    const double sz = copysign(1.0, pi.z - (fz + 0.5) * p.cell) > 0.0 ? 1.0 : -1.0;
    uint64_t key[kPerLane];
    int32_t cj[kPerLane];
    uint32_t nw_l = 0u;   // bit t: candidate l + 8t is a new pair
    uint32_t npc = 0u;    // bit t: the new pair may touch
#pragma unroll
    for (int t = 0; t < kPerLane; ++t) {
        const int c = l + kLanes * t;
        const bool in = c < nc && c < C;
        const int cs = in ? (int)s.cslot[ic * C + c] : -1;
        cj[t] = in ? s.cand[ic * C + c] : (int32_t)ic;
        key[t] = 0;
        if (cs >= 0 && ((prune >> cs) & 1u)) s.cslot[ic * C + c] = -1;
        const bool ab = in && cs < 0;
        const double4 pj = pos[ab ? cj[t] : (int32_t)ic];
        if (ab) {
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr < p.acut2 && rr > 1e-10) {
                const double gx = floor(pj.x * p.inv_cell), gy = floor(pj.y * p.inv_cell),
                             gz = floor(pj.z * p.inv_cell);
                const int bx = gx == fx ? 0 : (gx == fx + sx ? 1 : -1);
                const int by = gy == fy ? 0 : (gy == fy + sy ? 1 : -1);
                const int bz = gz == fz ? 0 : (gz == fz + sz ? 1 : -1);
                if (bx >= 0 && by >= 0 && bz >= 0) {
                    key[t] = ((uint64_t)kStencilPos[bx | (by << 1) | (bz << 2)] << 32) |
                             (uint32_t)cj[t];
                    nw_l |= 1u << t;
                    const double rs = ri + pj.w;
                    if (rr < rs * rs * 1.000001) npc |= 1u << t;
                }
            }
        }
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
        const bool q0 = (p.quirks & RF_QUIRK_SLOT0_SENTINEL) && i == 0;
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
        // hand each filled slot's partner to its owner lane
#pragma unroll
        for (int t2 = 0; t2 < kPerLane; ++t2) {
            uint32_t m = nbt[t2];
            while (m) {
                const int L = __builtin_ctz(m);
                m &= m - 1u;
                const int rk = __shfl(rank[t2], L, kLanes);
                const int32_t j = __shfl(cj[t2], L, kLanes);
                const int sl = free_slot(freem, rk, q0);
                if (sl >= 0 && (sl & (kLanes - 1)) == l) {
                    const int tt = sl / kLanes;
#pragma unroll
                    for (int t = 0; t < kPerLane; ++t)
                        if (t == tt) jk[t] = j;
                }
            }
        }
        occ |= added;
    }
    const uint32_t pc = grp_or(pc_l);

    // ---- 2a. bonds: lane l owns bonds l + 8t (d_CalculateForcesDEM_22)
    double bfx = 0.0, bfy = 0.0, bfz = 0.0, bmx = 0.0, bmy = 0.0, bmz = 0.0;
    uint32_t brk_l = 0u;
#pragma unroll 1
    for (int t = 0; t < kPerLane; ++t) {
        const int b = l + kLanes * t;
        if (!__any(b < nbd)) break;  // wave-uniform
        const bool inb = b < nbd && b < B;
        const int64_t qb = ic * B + b;
        const int bks = inb ? (int)s.bslot[qb] : -1;
        const int32_t bjr = inb ? s.bpart[qb] : (int32_t)ic;
        const bool live = bks >= 0 && ((bnd >> bks) & 1u);
        const int32_t bj = live ? bjr : (int32_t)ic;
        const double4 pj = pos[bj];
        const double4 vj = vel[bj];
        const double4 wj = angv[bj];
        const double2* rp = reinterpret_cast<const double2*>(s.brec + (live ? qb : ic * B));
        const double2 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3], r4 = rp[4], r5 = rp[5],
                      r6 = rp[6], r7 = rp[7];
        if (!live) continue;
        const double rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
        const double rijm = sqrt(rx * rx + ry * ry + rz * rz);
        const double inv = 1.0 / rijm;
        const double nx = rx * inv, ny = ry * inv, nz = rz * inv;
        const double pnx = r0.x, pny = r0.y, pnz = r1.x;
        const double oo0 = r1.y, oo1 = r2.x, oo2 = r2.y;
        const double on0 = r3.x, on1 = r3.y, on2 = r4.x;
        const double ot0 = r4.y, ot1 = r5.x, ot2 = r5.y;
        const double lam = r6.x, bsc = r6.y, btc = r7.x;
        const double rj = pj.w;
        const double bb = p.b_ratio > 0.0 ? p.b_ratio * (ri < rj ? ri : rj) : p.b_abs;
        double cefn = p.E * (rijm - 1.0 / lam) * lam;
        double Sc = cefn;
        cefn *= bb * bb * RF_PI;
        const double tx = pny * nz - pnz * ny;
        const double ty = pnz * nx - pnx * nz;
        const double tz = pnx * ny - pny * nx;
        const double cw = ((wj.x + wi.x) * nx + (wj.y + wi.y) * ny + (wj.z + wi.z) * nz) * p.dt * 0.5;
        const double px = nx * cw, py = ny * cw, pz = nz * cw;
        const double M1x = 1.0 + tz * pz + ty * py, M1y = pz - tz - ty * px, M1z = -py - tz * px + ty;
        const double M2x = tz - pz - tx * py, M2y = tz * pz + 1.0 + tx * px, M2z = -tz * py + px - tx;
        const double M3x = -ty - tx * pz + py, M3y = -ty * pz + tx - px, M3z = ty * py + tx * px + 1.0;
        // torsion
        const double dwx = wj.x - wi.x, dwy = wj.y - wi.y, dwz = wj.z - wi.z;
        double c1 = nx * dwx + ny * dwy + nz * dwz;
        c1 *= p.dt * 0.5 * p.G * lam;
        const double mnx = M1x * on0 + M1y * on1 + M1z * on2 + nx * c1;
        const double mny = M2x * on0 + M2y * on1 + M2z * on2 + ny * c1;
        const double mnz = M3x * on0 + M3y * on1 + M3z * on2 + nz * c1;
        const double J4 = RF_PI * bb * bb * bb * bb;
        // bending
        const double c2 = p.dt * 0.25 * p.E * lam;
        const double c1b = nx * dwx + ny * dwy + nz * dwz;
        const double mtx = M1x * ot0 + M1y * ot1 + M1z * ot2 + (dwx - nx * c1b) * c2;
        const double mty = M2x * ot0 + M2y * ot1 + M2z * ot2 + (dwy - ny * c1b) * c2;
        const double mtz = M3x * ot0 + M3y * ot1 + M3z * ot2 + (dwz - nz * c1b) * c2;
        const double c2J = (p.quirks & RF_QUIRK_BEND_DOUBLE_SCALE) ? c2 * J4 : J4;
        // shear force and asymmetric moment
        const double hx = 0.5 * (wj.x + wi.x), hy = 0.5 * (wj.y + wi.y), hz = 0.5 * (wj.z + wi.z);
        const double vx = vj.x - vi.x - hy * rz + hz * ry;
        const double vy = vj.y - vi.y - hz * rx + hx * rz;
        const double vz = vj.z - vi.z - hx * ry + hy * rx;
        const double cn = nx * vx + ny * vy + nz * vz;
        const double ox = (M1x * oo0 + M1y * oo1 + M1z * oo2) + (vx - nx * cn) * p.dt;
        const double oy = (M2x * oo0 + M2y * oo1 + M2z * oo2) + (vy - ny * cn) * p.dt;
        const double oz = (M3x * oo0 + M3y * oo1 + M3z * oo2) + (vz - nz * cn) * p.dt;
        double cef = p.G * lam;
        double Mc = cef;
        cef *= bb * bb * RF_PI;
        bfx = bfx + (nx * cefn + ox * cef);
        bfy = bfy + (ny * cefn + oy * cef);
        bfz = bfz + (nz * cefn + oz * cef);
        bmx = bmx + ((mnx * J4 + mtx * c2J) + 0.5 * (ry * oz - rz * oy) * cef);
        bmy = bmy + ((mny * J4 + mty * c2J) + 0.5 * (rz * ox - rx * oz) * cef);
        bmz = bmz + ((mnz * J4 + mtz * c2J) + 0.5 * (rx * oy - ry * ox) * cef);
        Mc *= sqrt(ox * ox + oy * oy + oz * oz);
        Sc += sqrt(mtx * mtx + mty * mty + mtz * mtz) * 4.0 * bb;
        Mc += sqrt(mnx * mnx + mny * mny + mnz * mnz) * 2.0 * bb;
        if (Sc > bsc || Mc > btc) {
            brk_l |= 1u << bks;  // the entry is dead from now on
        } else {
            double2* wp = reinterpret_cast<double2*>(s.brec + qb);
            wp[0] = make_double2(nx, ny);
            wp[1] = make_double2(nz, ox);
            wp[2] = make_double2(oy, oz);
            wp[3] = make_double2(mnx, mny);
            wp[4] = make_double2(mnz, mtx);
            wp[5] = make_double2(mty, mtz);
        }
    }
    const double Fbx = tree8(bfx), Fby = tree8(bfy), Fbz = tree8(bfz);
    const double Mbx = tree8(bmx), Mby = tree8(bmy), Mbz = tree8(bmz);
    bnd &= ~grp_or(brk_l);

    // ---- 2b. contacts: lane l owns slots l + 8t (d_CalculateForcesDEM_21)
    const double wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
    const bool roll = wim > 1e-18;
    const uint32_t cmask = occ & ~bnd;
    double cfx = 0.0, cfy = 0.0, cfz = 0.0, cmx = 0.0, cmy = 0.0, cmz = 0.0;
    uint32_t ovl_l = 0u;
    if (__any((cmask & pc) != 0u)) {
#pragma unroll 1
        for (int t = 0; t < kPerLane; ++t) {
            const int k = l + kLanes * t;
            const bool chk = ((cmask & pc) >> k) & 1u;
            int32_t j = 0;
#pragma unroll
            for (int u = 0; u < kPerLane; ++u)
                if (u == t) j = jk[u];
            if (!__any(chk)) continue;
            if (!chk) continue;
            const int64_t qs = ic * S + k;
            const double4 pj = pos[j];
            const double rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
            const double rijm = sqrt(rx * rx + ry * ry + rz * rz);
            const double rj = pj.w;
            const double rsum = ri + rj;
            if (!(rijm < rsum)) continue;
            const double inv = 1.0 / rijm;
            const double nx = rx * inv, ny = ry * inv, nz = rz * inv;
            ovl_l |= 1u << k;
            const double4 vj = vel[j];
            const double4 wj = angv[j];
            const double4 oprev = ((otnz >> k) & 1u) ? s.ot[qs] : make_double4(0.0, 0.0, 0.0, 0.0);
            const double mj = vj.w;
            const double Rst = ri * rj / (ri + rj);
            const double meff = mi * mj / (mi + mj);
            double c1 = sqrt(Rst * fabs(rsum - rijm));
            const double Kn = 2.0 * p.hmE * c1;
            const double Kt = 8.0 * p.hmG * c1;
            const double wsx = rj * wj.x + ri * wi.x;
            const double wsy = rj * wj.y + ri * wi.y;
            const double wsz = rj * wj.z + ri * wi.z;
            const double vx = vj.x - vi.x + (ny * wsz - nz * wsy);
            const double vy = vj.y - vi.y + (nz * wsx - nx * wsz);
            const double vz = vj.z - vi.z + (nx * wsy - ny * wsx);
            const double vn = nx * vx + ny * vy + nz * vz;
            const double fn = (rijm - rsum) * Kn * kTwoThirds;
            const double fd = kDamp * p.pA * vn * sqrt(Kn * meff);
            const double vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
            const double o0 = oprev.x, o1 = oprev.y, o2 = oprev.z;
            c1 = nx * o0 + ny * o1 + nz * o2;
            double ox = o0 - nx * c1, oy = o1 - ny * c1, oz = o2 - nz * c1;
            c1 = ox * ox + oy * oy + oz * oz;
            if (c1 > 1e-18) {
                c1 = 1.0 / c1;
                c1 *= o0 * o0 + o1 * o1 + o2 * o2;
                c1 = sqrt(c1);
                ox *= c1; oy *= c1; oz *= c1;
            }
            ox += vtx * p.dt; oy += vty * p.dt; oz += vtz * p.dt;
            double ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
            const double ftm = ftx * ftx + fty * fty + ftz * ftz;
            if (ftm > p.mu * p.mu * fn * fn) {
                c1 = p.mu * fabs(fn) * (1.0 / sqrt(ftm));
                ftx *= c1; fty *= c1; ftz *= c1;
                c1 = 1.0 / Kt;
                ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
            } else {
                c1 = kDamp * p.pA * sqrt(Kt * meff);
                ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
            }
            s.ot[qs] = make_double4(ox, oy, oz, 0.0);
            cfx = cfx + (nx * (fn + fd) + ftx);
            cfy = cfy + (ny * (fn + fd) + fty);
            cfz = cfz + (nz * (fn + fd) + ftz);
            const double tqx = (ny * ftz - nz * fty) * ri;
            const double tqy = (nz * ftx - nx * ftz) * ri;
            const double tqz = (nx * fty - ny * ftx) * ri;
            if (roll) {
                const double cr = -p.mur * fabs(fn) * (1.0 / sqrt(wim));
                cmx = cmx + (tqx + wi.x * cr);
                cmy = cmy + (tqy + wi.y * cr);
                cmz = cmz + (tqz + wi.z * cr);
            } else {
                cmx = cmx + tqx; cmy = cmy + tqy; cmz = cmz + tqz;
            }
        }
    }
    otnz = (otnz & ~cmask) | grp_or(ovl_l);
    const double Fcx = tree8(cfx), Fcy = tree8(cfy), Fcz = tree8(cfz);
    double Mcx = tree8(cmx), Mcy = tree8(cmy), Mcz = tree8(cmz);
    if ((p.quirks & RF_QUIRK_ROLL_RESET) && !roll) {
        Mcx = 0.0; Mcy = 0.0; Mcz = 0.0;
    }

    // ---- 3. totals, platens, kick-drift, cut (lane 0 of the group)
    double part_b = -0.0, part_t = -0.0;
    if (valid && l == 0) {
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
        const double2 pk = s.pk[i];
        const double4 rb = s.rbuild[i];
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
        const double ex = nrx - rb.x, ey = nry - rb.y, ez = nrz - rb.z;
        if (out || ex * ex + ey * ey + ez * ez > p.skin_half2) ctrl->rebuild = 1;
        sh_b[grp] = part_b;
        sh_t[grp] = part_t;
        sh_nb[grp] = __popc(bnd);
    } else if (l == 0) {
        sh_b[grp] = -0.0;
        sh_t[grp] = -0.0;
        sh_nb[grp] = 0;
    }

    // platen partials over the block's 32 particles, tree v[t] += v[t+off]
    __syncthreads();
    if (threadIdx.x == 0) {
        double vb[kPerBlock], vt[kPerBlock];
        int nb = 0;
#pragma unroll
        for (int t = 0; t < kPerBlock; ++t) {
            vb[t] = sh_b[t];
            vt[t] = sh_t[t];
            nb += sh_nb[t];
        }
#pragma unroll
        for (int off = kPerBlock / 2; off >= 1; off >>= 1)
#pragma unroll
            for (int t = 0; t < off; ++t) {
                vb[t] = vb[t] + vb[t + off];
                vt[t] = vt[t] + vt[t + off];
            }
        s.part_b[blockIdx.x] = vb[0];
        s.part_t[blockIdx.x] = vt[0];
        s.part_bonds[blockIdx.x] = nb;
    }
}

// -------------------------------------------------------- controller ------
// SumForcesULoading / SumForcesBLoading (md_results.cu:56-145) and
// CompressU_S3 / CompressB_S3 (md_data_types.h:347-379, 462-494), on device.
// Platen-force sum: kCtrlThreads strided running sums of the block partials,
// then a kCtrlThreads-wide tree (the CPU restatement uses the same order).
__global__ __launch_bounds__(kCtrlThreads) void k_controller(DevState s, DevParams p,
                                                             Ctrl* __restrict__ ctrl) {
    __shared__ double sb[kCtrlThreads];
    __shared__ double stt[kCtrlThreads];
    __shared__ long long sn[kCtrlThreads];
    if (ctrl->halt) return;
    const int t = threadIdx.x;
    double ab = 0.0, at = 0.0;
    long long nb = 0;
#pragma unroll 4
    for (int b = t; b < p.nblk; b += kCtrlThreads) {
        ab = ab + s.part_b[b];
        at = at + s.part_t[b];
        nb += s.part_bonds[b];
    }
    sb[t] = ab;
    stt[t] = at;
    sn[t] = nb;
    __syncthreads();
    for (int off = kCtrlThreads / 2; off >= 1; off >>= 1) {
        if (t < off) {
            sb[t] = sb[t] + sb[t + off];
            stt[t] = stt[t] + stt[t + off];
            sn[t] += sn[t + off];
        }
        __syncthreads();
    }
    if (t != 0) return;
    Ctrl c = *ctrl;
    if (p.kind != RF_TEST_NONE) {
        c.sfl0 = sb[0];
        c.sfl1 = stt[0];
        const int64_t dn = p.dnsave > 0 ? p.dnsave : 1;
        if (c.step % dn == 0) {
            if (c.stepsave < p.curve_cap) {
                s.bcount[c.stepsave] = sn[0];
                s.curve[3 * c.stepsave] = (c.size_axis - c.size0) / c.size0;
                s.curve[3 * c.stepsave + 1] = c.sfl0 * p.inv_area;
                s.curve[3 * c.stepsave + 2] = c.sfl1 * p.inv_area;
            }
            c.last_bonds = sn[0];
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
    c.halt = (c.done || c.rebuild || c.err) ? 1 : 0;
    *ctrl = c;
}

// ------------------------------------------------------ cell binning ------
// d_CalculateCellIndex (md_calculate_cell.cu:13-41) plus a grid check the
// reference only prints under pre_debugtest.
__global__ void k_cell_keys(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 r = s.pos[ctrl->cur][i];
    const double fx = floor(r.x * p.inv_cell), fy = floor(r.y * p.inv_cell),
                 fz = floor(r.z * p.inv_cell);
    if (!(fx >= 1.0 && fx <= p.nx - 2 && fy >= 1.0 && fy <= p.ny - 2 && fz >= 1.0 &&
          fz <= p.nz - 2)) {
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

// Within a cell the ids are put in ascending order: the stable (cell, id)
// order of cub::DeviceRadixSort::SortPairs over identity values
// (md_cell_distribution.cu:87).  Rank = number of smaller ids in the cell.
__global__ void k_cell_rank(DevState s, DevParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int32_t key = s.ckey[i];
    const int32_t st = s.cstart[key], ct = s.ccount[key];
    int32_t rank = 0;
    for (int32_t q = 0; q < ct; ++q) rank += s.ctmp[st + q] < (int32_t)i ? 1 : 0;
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
            rec.sc = 0.0; rec.tc = 0.0; rec.pad = 0.0;
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

// Candidate lists: every j within acut + skin (27 cells of edge 2 acut), plus
// every current row partner; cslot links a candidate to its slot.
__global__ void k_build_candidates(DevState s, DevParams p, Ctrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4* pos = s.pos[ctrl->cur];
    const double4 pi = pos[i];
    const int64_t key = s.ckey[i];
    const int64_t nx = p.nx, nxy = (int64_t)p.nx * p.ny;
    const int64_t C = p.C;
    int cnt = 0;
    bool over = false;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int64_t cc = key + dx + dy * nx + dz * nxy;
                const int32_t st = s.cstart[cc], ct = s.ccount[cc];
                for (int32_t q = 0; q < ct; ++q) {
                    const int32_t j = s.sorted[st + q];
                    if (j == (int32_t)i) continue;
                    const double4 pj = pos[j];
                    const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
                    if (ex * ex + ey * ey + ez * ez < p.cand_r2) {
                        if (cnt >= C) {
                            over = true;
                            continue;
                        }
                        s.cand[i * C + cnt] = j;
                        s.cslot[i * C + cnt] = -1;
                        ++cnt;
                    }
                }
            }
    const uint32_t occ = s.occ[i];
    for (int k = 0; k < p.S; ++k) {
        if (!((occ >> k) & 1u)) continue;
        const int32_t j = s.nbr[i * p.S + k];
        int c = 0;
        while (c < cnt && s.cand[i * C + c] != j) ++c;
        if (c == cnt) {
            if (cnt >= C) {
                over = true;
                continue;
            }
            s.cand[i * C + cnt] = j;
            ++cnt;
        }
        s.cslot[i * C + c] = (int8_t)k;
    }
    if (over) atomicOr(&ctrl->err, kErrCandOverflow);
    s.ncand[i] = cnt;
    s.rbuild[i] = pi;
}

// Per-bond strengths drawn on the host, scattered into the bond records.
__global__ void k_set_strengths(BondRec* brec, const double* sc, const double* tc, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    brec[q].sc = sc[q];
    brec[q].tc = tc[q];
}

__global__ void k_clear_flags(Ctrl* ctrl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->rebuild = 0;
        ctrl->halt = (ctrl->done || ctrl->err) ? 1 : 0;
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
    hipLaunchKernelGGL(k_build_candidates, dim3(grid_for(p.n, 128)), dim3(128), 0, st, s, p,
                       ctrl);
    return hipGetLastError();
}

hipError_t launch_set_strengths(BondRec* brec, const double* sc, const double* tc, int64_t n,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_set_strengths, dim3(grid_for(n, 256)), dim3(256), 0, st, brec, sc, tc, n);
    return hipGetLastError();
}

hipError_t launch_clear_flags(Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, st, ctrl);
    return hipGetLastError();
}

hipError_t launch_step(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_step, dim3((unsigned)p.nblk), dim3(kBlock), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_controller(const DevState& s, const DevParams& p, Ctrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(kCtrlThreads), 0, st, s, p, ctrl);
    return hipGetLastError();
}

hipError_t launch_counters(const DevState& s, const DevParams& p, Ctrl* ctrl, int64_t* out,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_counters, dim3(grid_for(p.n, 256)), dim3(256), 0, st, s, p, ctrl,
                       (unsigned long long*)out);
    return hipGetLastError();
}

}  // namespace rf
