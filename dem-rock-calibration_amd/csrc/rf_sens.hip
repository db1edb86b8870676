// rf_sens.hip -- forward-mode (tangent-linear) sensitivities of the loading
// step with respect to material constants.
//
// The step is evaluated once in Dual<kSensK> arithmetic (rf_physics.h):
// every quantity carries its value and its derivative along kSensK
// parameter directions, seeded by the derivatives of the material constants
// (E, G, hm_E, hm_G, mu, mu_r, wall moduli, bond radius ratio) with respect
// to the chosen parameters.  One thread per particle; per-particle sums in
// the step kernel's lane order, so the values are the step kernel's bit for
// bit.  The slot table is held fixed: pairs crossing acut
// carry no force either way, and a touching pair without a slot ends the
// window with an error.  The step size dt is a constant of the run.
//
// Replaces the finite-difference Jacobian of the reference's calibration
// (CalculateStepCG, mp_determineconjgradient.cu:84-142: NParam+1 forward
// runs) by one run carrying kSensK tangents.
#include <hip/hip_runtime.h>

#include "rf_internal.h"
#include "rf_physics.h"

namespace rf {

using D = Dual<kSensK>;

namespace {

__device__ __forceinline__ D mk(double v, const double* t, int64_t stride) {
    D r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < kSensK; ++k) r.d[k] = t[k * stride];
    return r;
}

__device__ __forceinline__ Vec4<D> load4(const double4& v, const double4* t, int64_t i, int64_t n,
                                         bool tangent_w) {
    Vec4<D> r;
    r.x.v = v.x; r.y.v = v.y; r.z.v = v.z; r.w.v = v.w;
#pragma unroll
    for (int k = 0; k < kSensK; ++k) {
        const double4 q = t[k * n + i];
        r.x.d[k] = q.x; r.y.d[k] = q.y; r.z.d[k] = q.z;
        r.w.d[k] = tangent_w ? q.w : 0.0;
    }
    return r;
}

__device__ __forceinline__ void store4(double4* vv, double4* tt, int64_t i, int64_t n,
                                       const D& x, const D& y, const D& z, double w) {
    vv[i] = make_double4(x.v, y.v, z.v, w);
#pragma unroll
    for (int k = 0; k < kSensK; ++k) tt[k * n + i] = make_double4(x.d[k], y.d[k], z.d[k], 0.0);
}

__device__ __forceinline__ MatT<D> mat_dual(const DevParams& p, const SensState& t) {
    MatT<D> m;
    const double v[11] = {p.E, p.G, p.hmE, p.hmG, p.mu, p.mur, p.pA, p.Ew, p.Gw, p.b_ratio, p.b_abs};
    D* f[11] = {&m.E, &m.G, &m.hmE, &m.hmG, &m.mu, &m.mur, &m.pA, &m.Ew, &m.Gw, &m.b_ratio, &m.b_abs};
#pragma unroll
    for (int u = 0; u < 11; ++u) {
        f[u]->v = v[u];
#pragma unroll
        for (int k = 0; k < kSensK; ++k) f[u]->d[k] = t.dmat[k][u];
    }
    return m;
}

}  // namespace

__global__ __launch_bounds__(128) void k_sens_step(DevState s, DevParams p, SensState t,
                                                   Ctrl* __restrict__ ctrl) {
    __shared__ double sh_b[128], sh_t[128];
    __shared__ double th_b[kSensK][128], th_t[kSensK][128];
    __shared__ int sh_nb[128];
    if (ctrl->halt) return;
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    const bool valid = i < p.n;
    const int cur = ctrl->cur, nxt = cur ^ 1;
    const double top = ctrl->top, bottom = ctrl->bottom, vw = ctrl->vw;
    const int64_t n = p.n;
    const int S = p.S, C = p.C, B = p.B;
    const MatT<D> m = mat_dual(p, t);
    D pb = dconst<kSensK>(-0.0), pt = dconst<kSensK>(-0.0);
    int nb_out = 0;
    if (valid) {
        const double4* pos = s.pos[cur];
        const double4* vel = s.vel[cur];
        const double4* angv = s.angv[cur];
        const Vec4<D> pi = load4(pos[i], t.tpos[cur], i, n, false);
        const Vec4<D> vi = load4(vel[i], t.tvel[cur], i, n, false);
        const Vec4<D> wi = load4(angv[i], t.tang[cur], i, n, false);
        const D ri = pi.w, mi = vi.w;
        uint32_t occ = s.occ[i], bnd = s.bnd[i], otnz = s.otnz[i];
        // The slot table is held fixed.  Pairs that cross acut (> r_i + r_j)
        // carry no force, so leaving them in or out changes nothing; the
        // plain step catches up on them afterwards.  A pair that touches
        // without a slot, or coincides, ends the window.
        for (uint32_t q = occ; q; q &= q - 1u) {
            const int k = __builtin_ctz(q);
            const double4 pj = pos[s.nbr[i * S + k]];
            const double dx = pj.x - pi.x.v, dy = pj.y - pi.y.v, dz = pj.z - pi.z.v;
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr < 1e-10) atomicOr(&ctrl->err, kErrSensList);
        }
        for (int c = 0; c < s.ncand[i] && c < C; ++c) {
            if (s.cslot[i * C + c] >= 0) continue;
            const double4 pj = pos[s.cand[i * C + c]];
            const double dx = pj.x - pi.x.v, dy = pj.y - pi.y.v, dz = pj.z - pi.z.v;
            const double rr = dx * dx + dy * dy + dz * dz;
            const double rs = pi.w.v + pj.w;
            if (rr < rs * rs) atomicOr(&ctrl->err, kErrSensList);
        }
        const D zero = dconst<kSensK>(0.0);
        D F[3] = {zero, zero, zero};
        D M[3] = {zero, zero, zero};
        // Per-particle sums in the step kernel's order: entry e goes to lane
        // e % 8, each lane adds its entries in ascending order from 0.0, and
        // the 8 lane sums meet in the tree v[t] += v[t+off], off = 4, 2, 1.
        D acc[kLanes][6];
#pragma unroll
        for (int l = 0; l < kLanes; ++l)
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[l][u] = zero;
        auto lane_total = [&](D (&out)[6]) {
#pragma unroll
            for (int off = kLanes / 2; off >= 1; off >>= 1)
#pragma unroll
                for (int l = 0; l < off; ++l)
#pragma unroll
                    for (int u = 0; u < 6; ++u) acc[l][u] = acc[l][u] + acc[l + off][u];
#pragma unroll
            for (int u = 0; u < 6; ++u) out[u] = acc[0][u];
#pragma unroll
            for (int l = 0; l < kLanes; ++l)
#pragma unroll
                for (int u = 0; u < 6; ++u) acc[l][u] = zero;
        };
        // bonds in creation order
        const int nbd = s.nbond[i];
        for (int b = 0; b < nbd && b < B; ++b) {
            const int ks = s.bslot[i * B + b];
            if (ks < 0 || !((bnd >> ks) & 1u)) continue;
            const int64_t j = s.bpart[i * B + b];
            const int64_t qb = i * B + b;
            const Vec4<D> pj = load4(pos[j], t.tpos[cur], j, n, false);
            const Vec4<D> vj = load4(vel[j], t.tvel[cur], j, n, false);
            const Vec4<D> wj = load4(angv[j], t.tang[cur], j, n, false);
            const BondRec r = s.brec[qb];
            const double hv[12] = {r.pnx, r.pny, r.pnz, r.otx, r.oty, r.otz,
                                   r.mnx, r.mny, r.mnz, r.mtx, r.mty, r.mtz};
            D hist[12];
#pragma unroll
            for (int u = 0; u < 12; ++u) hist[u] = mk(hv[u], t.tbrec + qb * 12 + u, n * B * 12);
            D c6[6], h[12];
            const bool brk = bond_eval_t<D>(m, p, pi, vi, wi, pj, vj, wj, hist, r.invL0, r.L0, r.sc,
                                            r.tc, c6, h);
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[b & (kLanes - 1)][u] = acc[b & (kLanes - 1)][u] + c6[u];
            if (brk) {
                bnd &= ~(1u << ks);
            } else {
                BondRec& w = s.brec[qb];
                w.pnx = h[0].v; w.pny = h[1].v; w.pnz = h[2].v;
                w.otx = h[3].v; w.oty = h[4].v; w.otz = h[5].v;
                w.mnx = h[6].v; w.mny = h[7].v; w.mnz = h[8].v;
                w.mtx = h[9].v; w.mty = h[10].v; w.mtz = h[11].v;
#pragma unroll
                for (int u = 0; u < 12; ++u)
#pragma unroll
                    for (int k = 0; k < kSensK; ++k) t.tbrec[(k * n * B + qb) * 12 + u] = h[u].d[k];
            }
        }
        D bsum[6];
        lane_total(bsum);
        // contacts in slot order
        const D wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
        const bool roll = wim.v > 1e-18;
        const uint32_t cmask = occ & ~bnd;
        uint32_t ovl = 0u;
        for (uint32_t q = cmask; q; q &= q - 1u) {
            const int k = __builtin_ctz(q);
            const int64_t j = s.nbr[i * S + k];
            const int64_t qs = i * S + k;
            const double4 pjv = pos[j];
            const double dx = pjv.x - pi.x.v, dy = pjv.y - pi.y.v, dz = pjv.z - pi.z.v;
            if (!(::sqrt(dx * dx + dy * dy + dz * dz) < pi.w.v + pjv.w)) continue;
            ovl |= 1u << k;
            const Vec4<D> pj = load4(pjv, t.tpos[cur], j, n, false);
            const Vec4<D> vj = load4(vel[j], t.tvel[cur], j, n, false);
            const Vec4<D> wj = load4(angv[j], t.tang[cur], j, n, false);
            D op[3];
            const bool has = (otnz >> k) & 1u;
            const double4 ov = s.ot[qs];
            const double ovv[3] = {ov.x, ov.y, ov.z};
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                op[u].v = has ? ovv[u] : 0.0;
#pragma unroll
                for (int kk = 0; kk < kSensK; ++kk) {
                    const double4 tq = t.tot[(int64_t)kk * n * S + qs];
                    const double tv[3] = {tq.x, tq.y, tq.z};
                    op[u].d[kk] = has ? tv[u] : 0.0;
                }
            }
            D c6[6], on[3];
            contact_eval_t<D>(m, p, pi, vi, wi, pj, vj, wj, op, wim, roll, s.pk[i].y, c6, on);
            s.ot[qs] = make_double4(on[0].v, on[1].v, on[2].v, 0.0);
#pragma unroll
            for (int kk = 0; kk < kSensK; ++kk)
                t.tot[(int64_t)kk * n * S + qs] = make_double4(on[0].d[kk], on[1].d[kk], on[2].d[kk], 0.0);
#pragma unroll
            for (int u = 0; u < 6; ++u) acc[k & (kLanes - 1)][u] = acc[k & (kLanes - 1)][u] + c6[u];
        }
        otnz = (otnz & ~cmask) | ovl;
        D csum[6];
        lane_total(csum);
        if ((p.quirks & RF_QUIRK_ROLL_RESET) && !roll)
#pragma unroll
            for (int u = 3; u < 6; ++u) csum[u] = zero;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            F[u] = (F[u] + bsum[u]) + csum[u];
            M[u] = (M[u] + bsum[3 + u]) + csum[3 + u];
        }
        // platens
        if (p.kind != RF_TEST_NONE) {
            D px_, py_, pz_, r2;
            if (p.kind == RF_TEST_BRAZILIAN) {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y; pz_ = pi.z - p.scope_c[2];
                r2 = px_ * px_ + pz_ * pz_;
            } else {
                px_ = pi.x - p.scope_c[0]; py_ = pi.y - p.scope_c[1]; pz_ = pi.z;
                r2 = px_ * px_ + py_ * py_;
            }
            D s_b = dconst<kSensK>(0.0), s_t = dconst<kSensK>(0.0);
            if (r2.v < p.rr_scope && pz_.v < p.zcut) {
                const int ax = p.axis;
                const D ppos = ax == 2 ? pz_ : py_;
                D o0 = dconst<kSensK>(0.0), o1 = o0, o2 = o0;
                WallOutT<D> wo;
                if (top - ppos.v < ri.v) {
                    const double nwy = ax == 1 ? 1.0 : 0.0, nwz = ax == 2 ? 1.0 : 0.0;
                    const double wvy = ax == 1 ? vw : 0.0, wvz = ax == 2 ? vw : 0.0;
                    wall_contact_t<D>(m, p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                      fabs(top - ppos), o0, o1, o2, s.pk[i].y, wo);
                    s_t += ax == 2 ? wo.fnz : wo.fny;
                    F[0] += wo.fnx + wo.ftx; F[1] += wo.fny + wo.fty; F[2] += wo.fnz + wo.ftz;
                    M[0] += wo.mtx + wo.max_; M[1] += wo.mty + wo.may; M[2] += wo.mtz + wo.maz;
                    o0 = wo.ox; o1 = wo.oy; o2 = wo.oz;
                }
                if (ppos.v - bottom < ri.v) {
                    const double nwy = ax == 1 ? -1.0 : 0.0, nwz = ax == 2 ? -1.0 : 0.0;
                    const double wvy = ax == 1 ? -vw : 0.0, wvz = ax == 2 ? -vw : 0.0;
                    wall_contact_t<D>(m, p, ri, mi, vi, wi, 0.0, nwy, nwz, 0.0, wvy, wvz,
                                      fabs(bottom - ppos), o0, o1, o2, s.pk[i].y, wo);
                    s_b += ax == 2 ? wo.fnz : wo.fny;
                    F[0] += wo.fnx + wo.ftx; F[1] += wo.fny + wo.fty; F[2] += wo.fnz + wo.ftz;
                    M[0] += wo.mtx + wo.max_; M[1] += wo.mty + wo.may; M[2] += wo.mtz + wo.maz;
                }
            }
            pb = -s_b;
            pt = -s_t;
        }
        // kick-drift; the window must not lose particles
        const double2 pk = s.pk[i];
        const D nvx = vi.x + F[0] * pk.x, nvy = vi.y + F[1] * pk.x, nvz = vi.z + F[2] * pk.x;
        const D nrx = pi.x + nvx * p.dt, nry = pi.y + nvy * p.dt, nrz = pi.z + nvz * p.dt;
        const D nwx = wi.x + M[0] * pk.y, nwy = wi.y + M[1] * pk.y, nwz = wi.z + M[2] * pk.y;
        {
            const double dx = nrx.v - p.cen[0], dy = nry.v - p.cen[1], dz = nrz.v - p.cen[2];
            if (dx - p.half[0] > 0.0 || dx + p.half[0] < 0.0 || dy - p.half[1] > 0.0 ||
                dy + p.half[1] < 0.0 || dz - p.half[2] > 0.0 || dz + p.half[2] < 0.0)
                atomicOr(&ctrl->err, kErrSensList);
        }
        store4(s.pos[nxt], t.tpos[nxt], i, n, nrx, nry, nrz, pi.w.v);
        store4(s.vel[nxt], t.tvel[nxt], i, n, nvx, nvy, nvz, vi.w.v);
        store4(s.angv[nxt], t.tang[nxt], i, n, nwx, nwy, nwz, 0.0);
        if (p.store_fm) {
            s.F[i] = make_double4(F[0].v, F[1].v, F[2].v, 0.0);
            s.M[i] = make_double4(M[0].v, M[1].v, M[2].v, 0.0);
        }
        s.occ[i] = occ;
        s.bnd[i] = bnd;
        s.otnz[i] = otnz;
        const double4 rb = s.rbuild[i];
        const double ex = nrx.v - rb.x, ey = nry.v - rb.y, ez = nrz.v - rb.z;
        if (ex * ex + ey * ey + ez * ez > p.skin_half2) ctrl->rebuild = 1;
        nb_out = __popc(bnd);
    }
    // partials per 32 particles (the value controller's grouping)
    const int tid = threadIdx.x;
    sh_b[tid] = pb.v;
    sh_t[tid] = pt.v;
    sh_nb[tid] = nb_out;
#pragma unroll
    for (int k = 0; k < kSensK; ++k) {
        th_b[k][tid] = pb.d[k];
        th_t[k][tid] = pt.d[k];
    }
    __syncthreads();
    const int g = tid & 31;
    for (int off = 16; off >= 1; off >>= 1) {
        if (g < off) {
            sh_b[tid] = sh_b[tid] + sh_b[tid + off];
            sh_t[tid] = sh_t[tid] + sh_t[tid + off];
            sh_nb[tid] += sh_nb[tid + off];
#pragma unroll
            for (int k = 0; k < kSensK; ++k) {
                th_b[k][tid] = th_b[k][tid] + th_b[k][tid + off];
                th_t[k][tid] = th_t[k][tid] + th_t[k][tid + off];
            }
        }
        __syncthreads();
    }
    const int64_t grp = (int64_t)blockIdx.x * 4 + (tid >> 5);
    if (g == 0 && grp < p.nblk) {
        s.part_b[grp] = sh_b[tid];
        s.part_t[grp] = sh_t[tid];
        s.part_bonds[grp] = sh_nb[tid];
#pragma unroll
        for (int k = 0; k < kSensK; ++k) {
            t.tpart_b[k * p.nblk + grp] = th_b[k][tid];
            t.tpart_t[k * p.nblk + grp] = th_t[k][tid];
        }
    }
}

// Tangents of the platen forces; records the curve's tangents where the
// value controller (launched next) records a curve point.
__global__ __launch_bounds__(256) void k_sens_ctrl(DevParams p, SensState t, Ctrl* __restrict__ ctrl) {
    __shared__ double sb[kSensK][256], st[kSensK][256];
    if (ctrl->halt) return;
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSensK; ++k) {
        double ab = 0.0, at = 0.0;
        for (int b = tid; b < p.nblk; b += 256) {
            ab = ab + t.tpart_b[k * p.nblk + b];
            at = at + t.tpart_t[k * p.nblk + b];
        }
        sb[k][tid] = ab;
        st[k][tid] = at;
    }
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off)
#pragma unroll
            for (int k = 0; k < kSensK; ++k) {
                sb[k][tid] += sb[k][tid + off];
                st[k][tid] += st[k][tid + off];
            }
        __syncthreads();
    }
    if (tid != 0 || p.kind == RF_TEST_NONE) return;
    const int64_t dn = p.dnsave > 0 ? p.dnsave : 1;
    const int64_t step = ctrl->step, save = ctrl->stepsave;
    if (step % dn == 0 && save < p.curve_cap)
#pragma unroll
        for (int k = 0; k < kSensK; ++k) {
            t.tcurve[(save * kSensK + k) * 2] = sb[k][0] * p.inv_area;
            t.tcurve[(save * kSensK + k) * 2 + 1] = st[k][0] * p.inv_area;
        }
}

hipError_t launch_sens_step(const DevState& s, const DevParams& p, const SensState& t, Ctrl* ctrl,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_sens_step, dim3((unsigned)((p.n + 127) / 128)), dim3(128), 0, st, s, p, t,
                       ctrl);
    hipLaunchKernelGGL(k_sens_ctrl, dim3(1), dim3(256), 0, st, p, t, ctrl);
    return hipGetLastError();
}

}  // namespace rf
