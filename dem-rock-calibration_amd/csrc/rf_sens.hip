// rf_sens.hip -- forward-mode (tangent-linear) sensitivities of the loading
// step with respect to material constants.
//
// Every quantity of the step carries its value and its derivative along a
// parameter direction (Dual<1>, rf_physics.h), seeded by the derivatives of
// the material constants (E, G, hm_E, hm_G, mu, mu_r, wall moduli, bond
// radius ratio) with respect to the chosen parameter.  As many adjacent
// lanes of a wave as there are directions (1..kSensK) work on one particle,
// lane k carrying direction k: each lane recomputes the value (the same IEEE
// operations in the same order, so all lanes hold the same values) and one
// tangent.  Against one lane carrying all tangents this trades repeated
// value arithmetic for a fraction of the registers and more lanes.  Lane 0
// of a particle writes the values, every lane its tangent slice; tangent
// arrays are laid out [entity][kSensK] so the lanes of a particle read and
// write one contiguous run.  Platen forces go out per particle and a second
// kernel adds them per 32 particles in the value controller's tree order;
// the controller of a sensitivity step (k_controller_sens) adds the
// tangents where a curve point is recorded.  Per-particle sums follow the
// step kernel's lane order, so the values are the step kernel's bit for bit.  The slot table is held
// fixed: pairs crossing acut carry no force either way, and a touching pair
// without a slot ends the window with an error.  The step size dt is a
// constant of the run.
//
// Replaces the finite-difference Jacobian of the reference's calibration
// (CalculateStepCG, mp_determineconjgradient.cu:84-142: NParam+1 forward
// runs) by one run carrying kSensK tangents.
#include <hip/hip_runtime.h>

#include "rf_internal.h"
#include "rf_physics.h"

namespace rf {

using D = Dual<1>;

namespace {

#ifndef RF_SENS_BLOCK
#define RF_SENS_BLOCK 256
#endif
constexpr int kSensBlock = RF_SENS_BLOCK;  // threads per workgroup (whole waves)

// Minimum waves per SIMD the register allocation must allow (1: no limit,
// the kernel keeps everything in registers; 2 halves the registers at the
// price of spills).
#ifndef RF_SENS_WAVES
#define RF_SENS_WAVES 1
#endif

__device__ __forceinline__ D mk(double v, double t) {
    D r;
    r.v = v;
    r.d[0] = t;
    return r;
}

// value v, tangent slice dir of particle i ([i][direction] layout)
__device__ __forceinline__ Vec4<D> load4(const double4& v, const double4* t, int64_t i, int dir) {
    const double4 q = t[i * kSensK + dir];
    Vec4<D> r;
    r.x = mk(v.x, q.x);
    r.y = mk(v.y, q.y);
    r.z = mk(v.z, q.z);
    r.w = mk(v.w, 0.0);
    return r;
}

// Decisions of one pair evaluation in a frozen-branch window: taken
// (MODE 0), taken and reported in `out` (1), or read from `in` (2), with the
// recorded spin held by a chattering rolling torque.
template <int MODE>
struct SensBr {
    uint32_t in = 0u, out = 0u;
    double wx = 0.0, wy = 0.0, wz = 0.0;
    __device__ __forceinline__ bool decide(int b, bool v) {
        const bool d = MODE == 2 ? ((in >> b) & 1u) != 0u : v;
        if (MODE == 1) out |= (uint32_t)d << b;
        return d;
    }
    __device__ __forceinline__ void spin(const Vec4<D>& w, D& x, D& y, D& z) {
        if (MODE == 2) {
            x = dconst<1>(wx);
            y = dconst<1>(wy);
            z = dconst<1>(wz);
        } else {
            x = frozen(w.x);
            y = frozen(w.y);
            z = frozen(w.z);
        }
    }
};

__device__ __forceinline__ MatT<D> mat_dual(const DevParams& p, const SensState& t, int dir) {
    MatT<D> m;
    const double v[11] = {p.E, p.G, p.hmE, p.hmG, p.mu, p.mur, p.pA, p.Ew, p.Gw, p.b_ratio, p.b_abs};
    D* f[11] = {&m.E, &m.G, &m.hmE, &m.hmG, &m.mu, &m.mur, &m.pA, &m.Ew, &m.Gw, &m.b_ratio, &m.b_abs};
#pragma unroll
    for (int u = 0; u < 11; ++u) *f[u] = mk(v[u], t.dmat[dir][u]);
    return m;
}

// Bond balancing.  At 20k spheres a window's kernel is a single round of
// waves, so it lasts as long as its slowest lane, and a lane walks all of its
// particle's bonds: the most bonded particle of a wave sets its time (12
// bonds against a mean of 6.2 in the gradient workload).  Each direction
// class of a wave (its lanes d, d + nl, ...) therefore shares the class's
// bonds out: a lane evaluates the first `target` = ceil(class bonds / class
// particles) positions of its own particle, and the positions past that go to
// lanes of the class with fewer bonds, in the rounds those have left; their
// contributions come back through LDS and are added by the owner afterwards,
// in the owner's order, so every sum keeps its order.
constexpr int kHelpCap = 32;  // positions handed out per class and wave
constexpr int kSensWaves = kSensBlock / 64;

// Inclusive scan over the particles of a direction class (lanes pw * nl + d).
__device__ __forceinline__ int class_scan(int v, int pw, int nl) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, (unsigned)(off * nl));
        if (pw >= off) v += u;
    }
    return v;
}

// Virtual lane visited idx-th: 0, 4, 2, 6, 1, 5, 3, 7.
__device__ __forceinline__ int visit_lane(int idx) {
    return ((idx & 1) << 2) | (idx & 2) | ((idx >> 2) & 1);
}

// Bonds of virtual lane l among n: l, l + 8, ... below n.
__device__ __forceinline__ int lane_bonds(int l, int n) {
    return n > l ? (n - 1 - l) / kLanes + 1 : 0;
}

// Bond at position q of a particle's visiting order.
__device__ __forceinline__ int seq_bond(int q, int n) {
    for (int idx = 0; idx < kLanes; ++idx) {
        const int l = visit_lane(idx);
        const int c = lane_bonds(l, n);
        if (q < c) return l + kLanes * q;
        q -= c;
    }
    return -1;
}

// The per-particle sum of the step kernel's order, fed one position at a
// time: each virtual lane adds its entries from 0.0 in ascending order, and
// the lane sums meet in the tree ((v0+v4)+(v2+v6)) + ((v1+v5)+(v3+v7)),
// combined as the lanes complete (visiting order 0,4,2,6,1,5,3,7).
struct TreeSum {
    D hold[6], p1[6], q1[6], acc[6], tot[6];
    int idx, rem, n;
    __device__ __forceinline__ void init(int n_) {
        n = n_;
        idx = 0;
        rem = lane_bonds(visit_lane(0), n);
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[u] = dconst<1>(0.0);
    }
    __device__ __forceinline__ void close() {
        if (!(idx & 1)) {
#pragma unroll
            for (int u = 0; u < 6; ++u) hold[u] = acc[u];
        } else if (!(idx & 2)) {
#pragma unroll
            for (int u = 0; u < 6; ++u) p1[u] = hold[u] + acc[u];
        } else if (idx == 3) {
#pragma unroll
            for (int u = 0; u < 6; ++u) q1[u] = p1[u] + (hold[u] + acc[u]);
        } else {
#pragma unroll
            for (int u = 0; u < 6; ++u) tot[u] = q1[u] + (p1[u] + (hold[u] + acc[u]));
        }
        ++idx;
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[u] = dconst<1>(0.0);
        rem = idx < kLanes ? lane_bonds(visit_lane(idx), n) : 0;
    }
    // the next position: a bond that is alive adds its contribution
    __device__ __forceinline__ void add(bool alive, const D (&c6)[6]) {
        while (rem == 0 && idx < kLanes) close();
        if (alive) {
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c6[u];
        }
        --rem;
    }
    __device__ __forceinline__ void finish() {
        while (idx < kLanes) close();
    }
};

}  // namespace

template <int MODE>
__global__ __launch_bounds__(kSensBlock, RF_SENS_WAVES) void k_sens_step(DevState s, DevParams p, SensState t,
                                                          Ctrl* __restrict__ ctrl) {
    if (ctrl->halt) return;
    // frozen-branch record of this step and particle (MODE 1 writes it, 2 reads it)
    const int64_t ls = ctrl->step - t.br_step0;
    if (MODE != 0 && (ls < 0 || ls >= t.br_steps)) {
        if (threadIdx.x == 0) atomicOr(&ctrl->err, kErrSensList);
        return;
    }
    const int nl = t.nl;
    const int ppw = 64 / nl;  // class lanes per wave (at most this many particles)
    const int lane = threadIdx.x & 63;
    const int dir = lane % nl, pw = lane / nl;
    const bool lead = dir == 0;
    // the wave's particles: a run of at most ppw from the table the run's
    // start laid out (runs in bond-dense regions are shorter)
    const int wv = threadIdx.x >> 6;
    const int64_t gw = (int64_t)blockIdx.x * kSensWaves + wv;
    if (gw >= t.nwaves) return;  // wave-uniform
    const int64_t wave_base = t.wbase[gw];
    const int wcount = (int)(t.wbase[gw + 1] - wave_base);
    const int64_t i = wave_base + pw;
    const bool valid = pw < wcount;
    // a class lane without a particle in this wave only takes bond positions
    // handed out by the others
    const bool helper = !valid && pw < ppw;
    const int cur = ctrl->cur, nxt = cur ^ 1;
    const double top = ctrl->top, bottom = ctrl->bottom, vw = ctrl->vw;
    const int S = p.S, C = p.C, B = p.B;
    const MatT<D> m = mat_dual(p, t, dir);
    D pb = dconst<1>(-0.0), pt = dconst<1>(-0.0);
    int nb_out = 0;

    // ---- bond balancing within each direction class (every lane of the wave)
    const int64_t ic = valid ? i : 0;  // the particle (0 for the other lanes' loads)
    const uint32_t bnd0 = valid ? s.bnd[ic] : 0u;
    const int nbd0 = valid ? s.nbond[ic] : 0;
    const int nown = nbd0 < B ? nbd0 : B;
    // the slots and partners of the first 16 bonds (rows of 32: 16-byte aligned)
    const uint4 bsl = *reinterpret_cast<const uint4*>(s.bslot + ic * B);
    int4 bpr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bpr[q] = reinterpret_cast<const int4*>(s.bpart + ic * B)[q];
    auto slot_partner = [&](int b, int& ks, int64_t& j) {
        if (b < 16) {
            const uint32_t w = (b >> 2) == 0 ? bsl.x : (b >> 2) == 1 ? bsl.y
                                                   : (b >> 2) == 2 ? bsl.z : bsl.w;
            ks = (int)(int8_t)(uint8_t)(w >> (8 * (b & 3)));
            const int4 e = (b >> 2) == 0 ? bpr[0] : (b >> 2) == 1 ? bpr[1]
                                                  : (b >> 2) == 2 ? bpr[2] : bpr[3];
            j = (b & 3) == 0 ? e.x : (b & 3) == 1 ? e.y : (b & 3) == 2 ? e.z : e.w;
        } else {
            ks = s.bslot[ic * B + b];
            j = s.bpart[ic * B + b];
        }
    };
    const uint32_t rec3_0 =
        MODE == 2 && valid ? t.br_bits[(ls * p.n + ic) * kBrWords + 3] : 0u;
    const int last = (ppw - 1) * nl + dir;  // the class's last lane
    const int ntot = __shfl(class_scan(nown, pw, nl), last);
    // every lane of the class walks at most `target` positions
    int target = (ntot + ppw - 1) / ppw;
    int excess = nown > target ? nown - target : 0;
    int freec = (valid || helper) && target > nown ? target - nown : 0;
    const int ex_in = class_scan(excess, pw, nl), fr_in = class_scan(freec, pw, nl);
    int nhelp = __shfl(ex_in, last);  // the same in every class (same particles)
    int ex_pre = ex_in - excess, fr_pre = fr_in - freec;
    if (nhelp > kHelpCap) {  // too many to hand out: every lane walks its own bonds
        int mx = nown;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) mx = max(mx, __shfl_xor(mx, off));
        target = mx;
        excess = freec = nhelp = ex_pre = fr_pre = 0;
    }
    __shared__ int32_t h_j[kSensWaves][kSensK][kHelpCap];
    __shared__ int8_t h_ks[kSensWaves][kSensK][kHelpCap];
    __shared__ uint8_t h_pw[kSensWaves][kSensK][kHelpCap];
    __shared__ uint8_t h_b[kSensWaves][kSensK][kHelpCap];
    __shared__ uint8_t h_flag[kSensWaves][kSensK][kHelpCap];  // 1: alive, 2: recorded break
    __shared__ uint8_t h_brk[kSensWaves][kSensK][kHelpCap];
    __shared__ double h_c6[kSensWaves][kSensK][kHelpCap][12];
    for (int k = 0; k < excess; ++k) {  // this particle's positions past `target`
        const int e = ex_pre + k;
        const int b = seq_bond(target + k, nown);
        int ks;
        int64_t j;
        slot_partner(b, ks, j);
        const bool alive = ks >= 0 && ((bnd0 >> ks) & 1u);
        h_j[wv][dir][e] = (int32_t)j;
        h_ks[wv][dir][e] = (int8_t)ks;
        h_pw[wv][dir][e] = (uint8_t)pw;
        h_b[wv][dir][e] = (uint8_t)b;
        h_flag[wv][dir][e] = (uint8_t)((alive ? 1u : 0u) |
                                       (alive && ((rec3_0 >> ks) & 1u) ? 2u : 0u));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

    if (valid || helper) {  // helpers: the bond walk only
        const double4* pos = s.pos[cur];
        const double4* vel = s.vel[cur];
        const double4* angv = s.angv[cur];
        Vec4<D> pi = load4(pos[ic], t.tpos[cur], ic, dir);
        Vec4<D> vi = load4(vel[ic], t.tvel[cur], ic, dir);
        Vec4<D> wi = load4(angv[ic], t.tang[cur], ic, dir);
        uint32_t occ = valid ? s.occ[ic] : 0u, bnd = bnd0, otnz = valid ? s.otnz[ic] : 0u;
        uint32_t* const lw = MODE != 0 ? t.br_bits + (ls * p.n + ic) * kBrWords : nullptr;
        uint32_t rec[kBrWords] = {0u, 0u, 0u, 0u, 0u};  // the decisions (taken or replayed)
        double4 wrec = make_double4(0.0, 0.0, 0.0, 0.0);
        if (MODE == 2) {
#pragma unroll
            for (int u = 0; u < kBrWords; ++u) rec[u] = lw[u];
            wrec = t.br_w[ls * p.n + ic];
        }
        // The slot table is held fixed.  Contact pairs that cross acut
        // (> r_i + r_j) carry no force, so leaving them in or out changes
        // nothing; the plain step catches up on them afterwards.  A pair that
        // touches without a slot, or coincides, or a bond stretched past acut
        // (which the plain step would drop without a break) ends the window.  The particle's
        // lanes share the checks (blocks of 8 slots or candidates, every
        // gather of a block issued at once) and also find which slots
        // overlap, for the contact sums below.
        const int base = lane - dir;  // the particle's first lane
        uint32_t ovl_l = 0u;          // overlapping occupied slots (this lane's blocks)
        bool bad = false;
        for (int j = dir; j < kMaxSlots / 8; j += nl) {
            const uint32_t blk = (occ >> (8 * j)) & 0xffu;
            if (!blk) continue;
            int32_t jn[8];
            double4 q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = 8 * j + u;
                jn[u] = s.nbr[ic * S + (k < S ? k : S - 1)];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = pos[((blk >> u) & 1u) ? jn[u] : (int32_t)ic];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (!((blk >> u) & 1u)) continue;
                const double dx = q[u].x - pi.x.v, dy = q[u].y - pi.y.v, dz = q[u].z - pi.z.v;
                const double rr = dx * dx + dy * dy + dz * dz;
                bad = bad || rr < 1e-10 || (((bnd >> (8 * j + u)) & 1u) && rr > p.acut2);
                if (::sqrt(rr) < pi.w.v + q[u].w) ovl_l |= 1u << (8 * j + u);
            }
        }
        const int ncd = !valid ? 0 : s.ncand[ic] < C ? s.ncand[ic] : C;
        for (int j = dir; 8 * j < ncd; j += nl) {
            int32_t jc[8];
            int cs[8];
            double4 q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = 8 * j + u;
                const int64_t qc = ic * C + (c < C ? c : C - 1);
                jc[u] = s.cand[qc];
                cs[u] = s.cslot[qc];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                q[u] = pos[(8 * j + u < ncd && cs[u] < 0) ? jc[u] : (int32_t)ic];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (!(8 * j + u < ncd && cs[u] < 0)) continue;
                const double dx = q[u].x - pi.x.v, dy = q[u].y - pi.y.v, dz = q[u].z - pi.z.v;
                const double rr = dx * dx + dy * dy + dz * dz;
                const double rs = pi.w.v + q[u].w;
                bad = bad || rr < rs * rs;
            }
        }
        uint32_t ovl_all = 0u;
        int bad_all = 0;
        for (int d = 0; d < nl; ++d) {
            ovl_all |= (uint32_t)__shfl((int)ovl_l, base + d);
            bad_all |= __shfl((int)bad, base + d);
        }
        if (bad_all && lead) atomicOr(&ctrl->err, kErrSensList);
        const D zero = dconst<1>(0.0);
        D F[3] = {zero, zero, zero};
        D M[3] = {zero, zero, zero};
        // Per-particle sums in the step kernel's order: entry e goes to lane
        // e % 8, each lane adds its entries in ascending order from 0.0, and
        // the 8 lane sums meet in the tree v[t] += v[t+off], off = 4, 2, 1,
        // i.e. ((v0+v4)+(v2+v6)) + ((v1+v5)+(v3+v7)).  Lanes are visited in
        // the order 0,4,2,6,1,5,3,7 and combined as they complete, so only
        // four partial sums are live.
        auto tree_sum = [&](auto&& lane_sum, D (&total)[6]) {
            D hold[6], p1[6], q1[6], acc[6];
#pragma unroll 1
            for (int idx = 0; idx < kLanes; ++idx) {
                const int l = ((idx & 1) << 2) | (idx & 2) | ((idx >> 2) & 1);  // 0,4,2,6,1,5,3,7
#pragma unroll
                for (int u = 0; u < 6; ++u) acc[u] = zero;
                lane_sum(l, acc);
                if (!(idx & 1)) {
#pragma unroll
                    for (int u = 0; u < 6; ++u) hold[u] = acc[u];
                } else if (!(idx & 2)) {
#pragma unroll
                    for (int u = 0; u < 6; ++u) p1[u] = hold[u] + acc[u];
                } else if (idx == 3) {
#pragma unroll
                    for (int u = 0; u < 6; ++u) q1[u] = p1[u] + (hold[u] + acc[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < 6; ++u) total[u] = q1[u] + (p1[u] + (hold[u] + acc[u]));
                }
            }
        };
        // Bonds in creation order, entry b of a particle in virtual lane b % 8
        // (visiting order above).  Round r evaluates this particle's r-th
        // position while r < target, then the class's handed-out positions
        // assigned to this lane (their particle's state read from memory).
        TreeSum ts;
        ts.init(nown);
        const int own_n = nown < target ? nown : target;
#pragma unroll 1
        for (int r = 0; r < target; ++r) {
            const bool own = r < own_n;
            const int e = fr_pre + (r - own_n);
            if (!own && !(r - own_n < freec && e < nhelp)) continue;
            int64_t ip, j;
            int b, ks;
            bool alive;
            uint32_t brk_in;
            if (own) {
                ip = ic;
                b = seq_bond(r, nown);
                slot_partner(b, ks, j);
                alive = ks >= 0 && ((bnd >> ks) & 1u);
                brk_in = alive ? (rec[3] >> ks) & 1u : 0u;
            } else {
                ip = wave_base + h_pw[wv][dir][e];
                b = h_b[wv][dir][e];
                ks = h_ks[wv][dir][e];
                j = h_j[wv][dir][e];
                const uint32_t f = h_flag[wv][dir][e];
                alive = (f & 1u) != 0u;
                brk_in = (f >> 1) & 1u;
            }
            D c6[6];
#pragma unroll
            for (int u = 0; u < 6; ++u) c6[u] = dconst<1>(0.0);
            bool brk = false;
            if (alive) {
                const int64_t qb = ip * B + b;
                Vec4<D> pq = pi, vq = vi, wq = wi;
                if (!own) {
                    pq = load4(pos[ip], t.tpos[cur], ip, dir);
                    vq = load4(vel[ip], t.tvel[cur], ip, dir);
                    wq = load4(angv[ip], t.tang[cur], ip, dir);
                }
                const Vec4<D> pj = load4(pos[j], t.tpos[cur], j, dir);
                const Vec4<D> vj = load4(vel[j], t.tvel[cur], j, dir);
                const Vec4<D> wj = load4(angv[j], t.tang[cur], j, dir);
                const BondRec rr = s.brec[qb];
                const double hv[12] = {rr.pnx, rr.pny, rr.pnz, rr.otx, rr.oty, rr.otz,
                                       rr.mnx, rr.mny, rr.mnz, rr.mtx, rr.mty, rr.mtz};
                double* tb = t.tbrec + (qb * kSensK + dir) * 12;
                D hist[12];
#pragma unroll
                for (int u = 0; u < 12; ++u) hist[u] = mk(hv[u], tb[u]);
                D h[12];
                SensBr<MODE> br;
                br.in = brk_in << kBrBreak;
                brk = bond_eval_t<D>(m, p, pq, vq, wq, pj, vj, wj, hist, rr.invL0, rr.L0, rr.sc,
                                     rr.tc, c6, h, br);
                if (!brk) {
                    if (lead) {
                        BondRec& w = s.brec[qb];
                        w.pnx = h[0].v; w.pny = h[1].v; w.pnz = h[2].v;
                        w.otx = h[3].v; w.oty = h[4].v; w.otz = h[5].v;
                        w.mnx = h[6].v; w.mny = h[7].v; w.mnz = h[8].v;
                        w.mtx = h[9].v; w.mty = h[10].v; w.mtz = h[11].v;
                    }
#pragma unroll
                    for (int u = 0; u < 12; ++u) tb[u] = h[u].d[0];
                }
            }
            if (own) {
                if (MODE == 1 && alive) rec[3] |= (uint32_t)brk << ks;
                ts.add(alive, c6);
                if (brk) bnd &= ~(1u << ks);
            } else {
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    h_c6[wv][dir][e][2 * u] = c6[u].v;
                    h_c6[wv][dir][e][2 * u + 1] = c6[u].d[0];
                }
                h_brk[wv][dir][e] = brk ? 1u : 0u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int k = 0; k < excess; ++k) {  // positions evaluated by other lanes
            const int e = ex_pre + k;
            const bool alive = (h_flag[wv][dir][e] & 1u) != 0u;
            const int ks = h_ks[wv][dir][e];
            D c6[6];
#pragma unroll
            for (int u = 0; u < 6; ++u)
                c6[u] = alive ? mk(h_c6[wv][dir][e][2 * u], h_c6[wv][dir][e][2 * u + 1])
                              : dconst<1>(0.0);
            const bool brk = alive && h_brk[wv][dir][e] != 0u;
            if (MODE == 1 && alive) rec[3] |= (uint32_t)brk << ks;
            ts.add(alive, c6);
            if (brk) bnd &= ~(1u << ks);
        }
        ts.finish();
        D bsum[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) bsum[u] = ts.tot[u];
        const D ri = pi.w, mi = vi.w;
        // contacts in slot order; slot k goes to lane k % 8
        const D wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
        const bool roll = MODE == 2 ? (rec[4] & kBfRoll) != 0u : wim.v > 1e-18;
        const uint32_t cmask = occ & ~bnd;
        // contacts that touch
        const uint32_t ovl = cmask & (MODE == 2 ? rec[0] : ovl_all);
        if (MODE == 1) {
            rec[0] = ovl;
            rec[4] |= roll ? kBfRoll : 0u;
            wrec = make_double4(wi.x.v, wi.y.v, wi.z.v, 0.0);
        }
        auto contact_lane = [&](int l, D (&acc)[6]) {
#pragma unroll 1
            for (int k = l; k < 32; k += kLanes) {
                if (!((ovl >> k) & 1u)) continue;
                const int64_t j = s.nbr[ic * S + k];
                const int64_t qs = ic * S + k;
                const double4 pjv = pos[j];
                const Vec4<D> pj = load4(pjv, t.tpos[cur], j, dir);
                const Vec4<D> vj = load4(vel[j], t.tvel[cur], j, dir);
                const Vec4<D> wj = load4(angv[j], t.tang[cur], j, dir);
                const bool has = (otnz >> k) & 1u;
                const double4 ov = s.ot[qs];
                const double4 tq = t.tot[qs * kSensK + dir];
                D op[3] = {mk(has ? ov.x : 0.0, has ? tq.x : 0.0), mk(has ? ov.y : 0.0, has ? tq.y : 0.0),
                           mk(has ? ov.z : 0.0, has ? tq.z : 0.0)};
                D c6[6], on[3];
                SensBr<MODE> br;
                br.in = (((rec[1] >> k) & 1u) << kBrSlip) | (((rec[2] >> k) & 1u) << kBrChat);
                br.wx = wrec.x;
                br.wy = wrec.y;
                br.wz = wrec.z;
                contact_eval_t<D>(m, p, pi, vi, wi, pj, vj, wj, op, wim, roll, s.pk[ic].y, c6, on,
                                  br);
                if (MODE == 1) {
                    rec[1] |= ((br.out >> kBrSlip) & 1u) << k;
                    rec[2] |= ((br.out >> kBrChat) & 1u) << k;
                }
                if (lead) s.ot[qs] = make_double4(on[0].v, on[1].v, on[2].v, 0.0);
                t.tot[qs * kSensK + dir] = make_double4(on[0].d[0], on[1].d[0], on[2].d[0], 0.0);
#pragma unroll
                for (int u = 0; u < 6; ++u) acc[u] = acc[u] + c6[u];
            }
        };
        D csum[6];
        if (ovl != 0u) {
            tree_sum(contact_lane, csum);
        } else {  // no touching pair: the sums of +0.0 are +0.0
#pragma unroll
            for (int u = 0; u < 6; ++u) csum[u] = zero;
        }
        otnz = (otnz & ~cmask) | ovl;
        if ((p.quirks & RF_QUIRK_ROLL_RESET) && !roll)
#pragma unroll
            for (int u = 3; u < 6; ++u) csum[u] = zero;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            F[u] = (F[u] + bsum[u]) + csum[u];
            M[u] = (M[u] + bsum[3 + u]) + csum[3 + u];
        }
        // platens
        if (valid && p.kind != RF_TEST_NONE) {
            D px_, py_, pz_, r2;
            if (p.kind == RF_TEST_BRAZILIAN) {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y; pz_ = pi.z - p.scope_c[2];
                r2 = px_ * px_ + pz_ * pz_;
            } else {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y - p.scope_c[1]; pz_ = pi.z;
                r2 = px_ * px_ + py_ * py_;
            }
            D s_b = dconst<1>(0.0), s_t = dconst<1>(0.0);
            const bool scope = MODE == 2 ? (rec[4] & kBfScope) != 0u
                                         : (r2.v < p.rr_scope && pz_.v < p.zcut);
            if (MODE == 1 && scope) rec[4] |= kBfScope;
            if (scope) {
                const int ax = p.axis;
                const D ppos = ax == 2 ? pz_ : py_;
                D o0 = dconst<1>(0.0), o1 = o0, o2 = o0;
                WallOutT<D> wo;
                SensBr<MODE> br;
                br.wx = wrec.x;
                br.wy = wrec.y;
                br.wz = wrec.z;
                const bool t_on = MODE == 2 ? (rec[4] & kBfTop) != 0u : top - ppos.v < ri.v;
                if (t_on) {
                    br.in = (((rec[4] & kBfTopSlip) ? 1u : 0u) << kBrSlip) |
                            (((rec[4] & kBfTopChat) ? 1u : 0u) << kBrChat) |
                            (((rec[4] & kBfRoll) ? 1u : 0u) << kBrRoll);
                    br.out = 0u;
                    const double nwy = ax == 1 ? 1.0 : 0.0, nwz = ax == 2 ? 1.0 : 0.0;
                    const double wvy = ax == 1 ? vw : 0.0, wvz = ax == 2 ? vw : 0.0;
                    wall_contact_t<D>(m, p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                      fabs(top - ppos), o0, o1, o2, s.pk[ic].y, wo, br);
                    if (MODE == 1)
                        rec[4] |= kBfTop | ((br.out >> kBrSlip) & 1u ? kBfTopSlip : 0u) |
                                  ((br.out >> kBrChat) & 1u ? kBfTopChat : 0u);
                    s_t += ax == 2 ? wo.fnz : wo.fny;
                    F[0] += wo.fnx + wo.ftx; F[1] += wo.fny + wo.fty; F[2] += wo.fnz + wo.ftz;
                    M[0] += wo.mtx + wo.max_; M[1] += wo.mty + wo.may; M[2] += wo.mtz + wo.maz;
                    o0 = wo.ox; o1 = wo.oy; o2 = wo.oz;
                }
                const bool b_on = MODE == 2 ? (rec[4] & kBfBot) != 0u : ppos.v - bottom < ri.v;
                if (b_on) {
                    br.in = (((rec[4] & kBfBotSlip) ? 1u : 0u) << kBrSlip) |
                            (((rec[4] & kBfBotChat) ? 1u : 0u) << kBrChat) |
                            (((rec[4] & kBfRoll) ? 1u : 0u) << kBrRoll);
                    br.out = 0u;
                    const double nwy = ax == 1 ? -1.0 : 0.0, nwz = ax == 2 ? -1.0 : 0.0;
                    const double wvy = ax == 1 ? -vw : 0.0, wvz = ax == 2 ? -vw : 0.0;
                    wall_contact_t<D>(m, p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                      fabs(bottom - ppos), o0, o1, o2, s.pk[ic].y, wo, br);
                    if (MODE == 1)
                        rec[4] |= kBfBot | ((br.out >> kBrSlip) & 1u ? kBfBotSlip : 0u) |
                                  ((br.out >> kBrChat) & 1u ? kBfBotChat : 0u);
                    s_b += ax == 2 ? wo.fnz : wo.fny;
                    F[0] += wo.fnx + wo.ftx; F[1] += wo.fny + wo.fty; F[2] += wo.fnz + wo.ftz;
                    M[0] += wo.mtx + wo.max_; M[1] += wo.mty + wo.may; M[2] += wo.mtz + wo.maz;
                }
            }
            pb = -s_b;
            pt = -s_t;
        }
        // kick-drift; the window must not lose particles
        const double2 pk = s.pk[ic];
        const D nvx = vi.x + F[0] * pk.x, nvy = vi.y + F[1] * pk.x, nvz = vi.z + F[2] * pk.x;
        const D nrx = pi.x + nvx * p.dt, nry = pi.y + nvy * p.dt, nrz = pi.z + nvz * p.dt;
        const D nwx = wi.x + M[0] * pk.y, nwy = wi.y + M[1] * pk.y, nwz = wi.z + M[2] * pk.y;
        if (valid) {
            t.tpos[nxt][ic * kSensK + dir] = make_double4(nrx.d[0], nry.d[0], nrz.d[0], 0.0);
            t.tvel[nxt][ic * kSensK + dir] = make_double4(nvx.d[0], nvy.d[0], nvz.d[0], 0.0);
            t.tang[nxt][ic * kSensK + dir] = make_double4(nwx.d[0], nwy.d[0], nwz.d[0], 0.0);
        }
        if (lead && valid) {
            const double dx = nrx.v - p.cen[0], dy = nry.v - p.cen[1], dz = nrz.v - p.cen[2];
            if (dx - p.half[0] > 0.0 || dx + p.half[0] < 0.0 || dy - p.half[1] > 0.0 ||
                dy + p.half[1] < 0.0 || dz - p.half[2] > 0.0 || dz + p.half[2] < 0.0)
                atomicOr(&ctrl->err, kErrSensList);
            s.pos[nxt][ic] = make_double4(nrx.v, nry.v, nrz.v, pi.w.v);
            s.vel[nxt][ic] = make_double4(nvx.v, nvy.v, nvz.v, vi.w.v);
            s.angv[nxt][ic] = make_double4(nwx.v, nwy.v, nwz.v, 0.0);
            if (p.store_fm) {
                s.F[ic] = make_double4(F[0].v, F[1].v, F[2].v, 0.0);
                s.M[ic] = make_double4(M[0].v, M[1].v, M[2].v, 0.0);
            }
            s.occ[ic] = occ;
            s.bnd[ic] = bnd;
            s.otnz[ic] = otnz;
            if (MODE == 1) {
#pragma unroll
                for (int u = 0; u < kBrWords; ++u) lw[u] = rec[u];
                t.br_w[ls * p.n + ic] = wrec;
            }
            if (rebuild_due(s.rbuild[ic], nrx.v, nry.v, nrz.v)) ctrl->rebuild = 1;
        }
        nb_out = __popc(bnd);
    }
    if (valid) {
        if (lead) {
            t.pp[2 * i] = pb.v;
            t.pp[2 * i + 1] = pt.v;
            t.pnb[i] = nb_out;
        }
        t.tpp[(i * kSensK + dir) * 2] = pb.d[0];
        t.tpp[(i * kSensK + dir) * 2 + 1] = pt.d[0];
    }
}

// Platen forces and bond counts per 32 particles, in the step kernel's
// order: each run of 8 particles (one wave of the step kernel) by the tree
// v[k] += v[k+off], off = 4, 2, 1, with -0.0 past the last particle, into
// the group's four entries that the controller combines; tangents per
// direction by the tree v[g] += v[g+off], off = 16..1.
__global__ __launch_bounds__(256) void k_sens_partials(DevParams p, SensState t, DevState s,
                                                       Ctrl* __restrict__ ctrl) {
    // the values are read together with the control block; a halted step
    // writes nothing
    const int32_t halt = ctrl->halt;
    const int64_t step = ctrl->step, save = ctrl->stepsave;
    const int tid = threadIdx.x, g = tid & 31;  // one group per 32 lanes
    const int64_t grp = (int64_t)blockIdx.x * 8 + (tid >> 5);
    const int64_t i = grp * 32 + g;
    const bool in = i < p.n;
    const bool head = g == 0 && grp < p.nblk;
    double vb = in ? t.pp[2 * i] : -0.0, vt = in ? t.pp[2 * i + 1] : -0.0;
    int nb = in ? t.pnb[i] : 0;
    if (halt) return;
#pragma unroll
    for (int off = kPerWave / 2; off >= 1; off >>= 1) {
        vb = vb + __shfl_down(vb, off, kPerWave);
        vt = vt + __shfl_down(vt, off, kPerWave);
        nb += __shfl_down(nb, off, kPerWave);
    }
    if (g % kPerWave == 0 && grp < p.nblk) {
        const int64_t q = grp * kWavesPerGroup + g / kPerWave;
        s.part_b[q] = vb;
        s.part_t[q] = vt;
        s.part_bonds[q] = nb;
        if (vb != 0.0 || vt != 0.0) atomicOr(&s.pmask[grp >> 5], 1u << (grp & 31));
    }
    // tangents only at the steps whose curve point the controller records
    const int64_t dn = p.dnsave > 0 ? p.dnsave : 1;
    if (p.kind == RF_TEST_NONE || step % dn != 0 || save >= p.curve_cap) return;
    for (int k = 0; k < t.nl; ++k) {
        double db = in ? t.tpp[(i * kSensK + k) * 2] : -0.0;
        double dt_ = in ? t.tpp[(i * kSensK + k) * 2 + 1] : -0.0;
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) {
            db = db + __shfl_down(db, off, 32);
            dt_ = dt_ + __shfl_down(dt_, off, 32);
        }
        if (head) {
            t.tpart_b[k * p.nblk + grp] = db;
            t.tpart_t[k * p.nblk + grp] = dt_;
        }
    }
}

hipError_t launch_sens_step(const DevState& s, const DevParams& p, const SensState& t, Ctrl* ctrl,
                            hipStream_t st) {
    if (p.B != kMaxBonds) return hipErrorInvalidValue;  // bond rows read 16 entries at once
    const dim3 grid((unsigned)((t.nwaves + kSensWaves - 1) / kSensWaves));
    if (t.br_mode == 1)
        hipLaunchKernelGGL(k_sens_step<1>, grid, dim3(kSensBlock), 0, st, s, p, t, ctrl);
    else if (t.br_mode == 2)
        hipLaunchKernelGGL(k_sens_step<2>, grid, dim3(kSensBlock), 0, st, s, p, t, ctrl);
    else
        hipLaunchKernelGGL(k_sens_step<0>, grid, dim3(kSensBlock), 0, st, s, p, t, ctrl);
    hipLaunchKernelGGL(k_sens_partials, dim3((unsigned)((p.nblk + 7) / 8)), dim3(256), 0, st, p, t,
                       s, ctrl);
    return hipGetLastError();
}

}  // namespace rf
