// rf_physics.h -- pair physics of the loading step, written once for any
// scalar type: `double` (the step kernel) and `Dual<K>` (value plus K
// tangents, the sensitivity kernel).  For Dual, the value part of every
// operation is the same IEEE double operation in the same order, so a dual
// computation carries exactly the values of the double one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_internal.h"

namespace rf {

#define RF_PI 3.1415926535897932384626433832795
constexpr double kDamp = -1.8257;          // -2 sqrt(5/6), md_calculateforcesdem3.cu:248
constexpr double kTwoThirds = 2.0 / 3.0;   // MCf_2d3

using ::fabs;
using ::sqrt;

// ----------------------------------------------------------- dual numbers --
template <int K>
struct Dual {
    double v;
    double d[K];
};

__device__ __forceinline__ double val(double x) { return x; }
template <int K>
__device__ __forceinline__ double val(const Dual<K>& x) { return x.v; }

template <int K>
__device__ __forceinline__ Dual<K> dconst(double v) {
    Dual<K> r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = 0.0;
    return r;
}

#define RF_DUAL_BIN(OP, VEXPR, DEXPR)                                                      \
    template <int K>                                                                       \
    __device__ __forceinline__ Dual<K> operator OP(const Dual<K>& a, const Dual<K>& b) {   \
        Dual<K> r;                                                                         \
        r.v = VEXPR;                                                                       \
        _Pragma("unroll") for (int k = 0; k < K; ++k) r.d[k] = DEXPR;                      \
        return r;                                                                          \
    }
RF_DUAL_BIN(+, a.v + b.v, a.d[k] + b.d[k])
RF_DUAL_BIN(-, a.v - b.v, a.d[k] - b.d[k])
RF_DUAL_BIN(*, a.v * b.v, a.d[k] * b.v + a.v * b.d[k])
#undef RF_DUAL_BIN

template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) {
    Dual<K> r = a;
    r.v = a.v + b;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b) {
    Dual<K> r = b;
    r.v = a + b.v;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) {
    Dual<K> r = a;
    r.v = a.v - b;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a - b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a / b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = -r.v * b.d[k] / b.v;
    return r;
}
template <int K, class U>
__device__ __forceinline__ Dual<K>& operator+=(Dual<K>& a, const U& b) { return a = a + b; }
template <int K, class U>
__device__ __forceinline__ Dual<K>& operator-=(Dual<K>& a, const U& b) { return a = a - b; }
template <int K, class U>
__device__ __forceinline__ Dual<K>& operator*=(Dual<K>& a, const U& b) { return a = a * b; }

template <int K>
__device__ __forceinline__ Dual<K> sqrt(const Dual<K>& a) {
    Dual<K> r;
    r.v = ::sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> fabs(const Dual<K>& a) {
    return a.v < 0.0 ? -a : a;
}

// The value of x without its derivative.  Used for the direction w/|w| of
// the rolling-friction torque -mu_r |F_n| w/|w| when it chatters (below):
// its derivative with respect to w is ~ 1/|w| and acts in the tangent
// dynamics as a damping far beyond what the explicit step can carry (the
// values themselves flip around w = 0, bounded).  The torque then keeps its
// derivatives through mu_r and F_n only.
__device__ __forceinline__ double frozen(double x) { return x; }
template <int K>
__device__ __forceinline__ Dual<K> frozen(const Dual<K>& x) { return dconst<K>(x.v); }

// Rolling friction in its chatter regime: one step of this torque alone
// would reverse w (mu_r |F_n| dt/I >= |w|), so w flips around zero from
// step to step; explicit steps of the tangent lose stability at twice that.
// Only there is the direction frozen.  The test only picks which derivative
// to carry; values are the same.
template <class T>
__device__ __forceinline__ bool rolling_chatter(const T& mur, const T& fn, const T& wim,
                                                double dtI) {
    return val(mur) * ::fabs(val(fn)) * dtI >= ::sqrt(val(wim));
}

__device__ __forceinline__ double zero_like(double) { return 0.0; }
template <int K>
__device__ __forceinline__ Dual<K> zero_like(const Dual<K>&) { return dconst<K>(0.0); }

template <class T>
struct Vec4 {
    T x, y, z, w;
};

// Discrete decisions of the pair physics, by bit: a bond breaks, a contact
// slips (Coulomb cap), rolling friction is in its chatter regime, a particle
// spins at all (|W|^2 > 1e-18).  The plain step takes them (Decide); a
// frozen-branch sensitivity run takes and reports them, or replays those of
// another run, with the spin that a chattering rolling torque holds fixed.
constexpr int kBrBreak = 0, kBrSlip = 1, kBrChat = 2, kBrRoll = 3;

struct Decide {
    __device__ __forceinline__ bool decide(int, bool v) { return v; }
    template <class T>
    __device__ __forceinline__ void spin(const Vec4<T>& w, T& x, T& y, T& z) {
        x = frozen(w.x);
        y = frozen(w.y);
        z = frozen(w.z);
    }
};

__device__ __forceinline__ Vec4<double> v4(const double4& a) { return {a.x, a.y, a.z, a.w}; }

// Material constants gradients flow through (DevParams holds their values).
template <class T>
struct MatT {
    T E, G, hmE, hmG, mu, mur, pA, Ew, Gw, b_ratio, b_abs;
};

__device__ __forceinline__ MatT<double> mat_of(const DevParams& p) {
    return {p.E, p.G, p.hmE, p.hmG, p.mu, p.mur, p.pA, p.Ew, p.Gw, p.b_ratio, p.b_abs};
}

// ------------------------------------------------------------------ bond --
// One parallel bond seen from particle i (d_CalculateForcesDEM_22 and the
// dd_*_2 helpers, md_calculateforcesdem3.cu:20-200, 427-486): rotates the
// stored history hist = (prev n, o_t, m_n, m_t) into the current frame,
// increments it, and returns the force and moment contributions c, the new
// history h (same order) and whether the bond breaks.
template <class T, class BR = Decide>
__device__ __forceinline__ bool bond_eval_t(const MatT<T>& m, const DevParams& p,
                                            const Vec4<T>& pi, const Vec4<T>& vi,
                                            const Vec4<T>& wi, const Vec4<T>& pj,
                                            const Vec4<T>& vj, const Vec4<T>& wj,
                                            const T (&hist)[12], double lam, double L0,
                                            double bsc, double btc, T (&c)[6], T (&h)[12],
                                            BR&& br = BR()) {
    const T ri = pi.w, rj = pj.w;
    const T rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
    const T rijm = sqrt(rx * rx + ry * ry + rz * rz);
    const T inv = 1.0 / rijm;
    const T nx = rx * inv, ny = ry * inv, nz = rz * inv;
    const T pnx = hist[0], pny = hist[1], pnz = hist[2];
    const T oo0 = hist[3], oo1 = hist[4], oo2 = hist[5];
    const T on0 = hist[6], on1 = hist[7], on2 = hist[8];
    const T ot0 = hist[9], ot1 = hist[10], ot2 = hist[11];
    const T bb = val(m.b_ratio) > 0.0 ? m.b_ratio * (val(ri) < val(rj) ? ri : rj) : m.b_abs;
    T cefn = m.E * (rijm - L0) * lam;
    T Sc = cefn;
    cefn *= bb * bb * RF_PI;
    const T tx = pny * nz - pnz * ny;
    const T ty = pnz * nx - pnx * nz;
    const T tz = pnx * ny - pny * nx;
    const T cw = ((wj.x + wi.x) * nx + (wj.y + wi.y) * ny + (wj.z + wi.z) * nz) * p.dt * 0.5;
    const T px = nx * cw, py = ny * cw, pz = nz * cw;
    const T M1x = 1.0 + tz * pz + ty * py, M1y = pz - tz - ty * px, M1z = -py - tz * px + ty;
    const T M2x = tz - pz - tx * py, M2y = tz * pz + 1.0 + tx * px, M2z = -tz * py + px - tx;
    const T M3x = -ty - tx * pz + py, M3y = -ty * pz + tx - px, M3z = ty * py + tx * px + 1.0;
    // torsion
    const T dwx = wj.x - wi.x, dwy = wj.y - wi.y, dwz = wj.z - wi.z;
    T c1 = nx * dwx + ny * dwy + nz * dwz;
    c1 *= p.dt * 0.5 * m.G * lam;
    const T mnx = M1x * on0 + M1y * on1 + M1z * on2 + nx * c1;
    const T mny = M2x * on0 + M2y * on1 + M2z * on2 + ny * c1;
    const T mnz = M3x * on0 + M3y * on1 + M3z * on2 + nz * c1;
    const T J4 = RF_PI * bb * bb * bb * bb;
    // bending
    const T c2 = p.dt * 0.25 * m.E * lam;
    const T c1b = nx * dwx + ny * dwy + nz * dwz;
    const T mtx = M1x * ot0 + M1y * ot1 + M1z * ot2 + (dwx - nx * c1b) * c2;
    const T mty = M2x * ot0 + M2y * ot1 + M2z * ot2 + (dwy - ny * c1b) * c2;
    const T mtz = M3x * ot0 + M3y * ot1 + M3z * ot2 + (dwz - nz * c1b) * c2;
    const T c2J = (p.quirks & RF_QUIRK_BEND_DOUBLE_SCALE) ? c2 * J4 : J4;
    // shear force and asymmetric moment
    const T hx = 0.5 * (wj.x + wi.x), hy = 0.5 * (wj.y + wi.y), hz = 0.5 * (wj.z + wi.z);
    const T vx = vj.x - vi.x - hy * rz + hz * ry;
    const T vy = vj.y - vi.y - hz * rx + hx * rz;
    const T vz = vj.z - vi.z - hx * ry + hy * rx;
    const T cn = nx * vx + ny * vy + nz * vz;
    const T ox = (M1x * oo0 + M1y * oo1 + M1z * oo2) + (vx - nx * cn) * p.dt;
    const T oy = (M2x * oo0 + M2y * oo1 + M2z * oo2) + (vy - ny * cn) * p.dt;
    const T oz = (M3x * oo0 + M3y * oo1 + M3z * oo2) + (vz - nz * cn) * p.dt;
    T cef = m.G * lam;
    T Mc = cef;
    cef *= bb * bb * RF_PI;
    c[0] = nx * cefn + ox * cef;
    c[1] = ny * cefn + oy * cef;
    c[2] = nz * cefn + oz * cef;
    c[3] = (mnx * J4 + mtx * c2J) + 0.5 * (ry * oz - rz * oy) * cef;
    c[4] = (mny * J4 + mty * c2J) + 0.5 * (rz * ox - rx * oz) * cef;
    c[5] = (mnz * J4 + mtz * c2J) + 0.5 * (rx * oy - ry * ox) * cef;
    // Breakage: Sc = Sc0 + |m_t| 4b > sigma_c or Mc = (G/L0)|o_t| + |m_n| 2b > tau_c.
    // The three norms are first taken in single precision (relative error
    // below 2e-7); when both sums stay below their strengths by more than
    // 1e-6 of the norm terms the exact FP64 test cannot succeed either, and
    // is skipped.  Otherwise the exact test decides.
    const T os2 = ox * ox + oy * oy + oz * oz;
    const T mt2 = mtx * mtx + mty * mty + mtz * mtz;
    const T mn2 = mnx * mnx + mny * mny + mnz * mnz;
    const double b0 = val(bb);
    const double tS = (double)__builtin_amdgcn_sqrtf((float)val(mt2)) * 4.0 * b0;
    const double tM1 = val(Mc) * (double)__builtin_amdgcn_sqrtf((float)val(os2));
    const double tM2 = (double)__builtin_amdgcn_sqrtf((float)val(mn2)) * 2.0 * b0;
    const double Se = val(Sc) + tS, Me = tM1 + tM2;
    const bool certain = (Se + 1e-6 * tS + 1e-14 * (::fabs(Se) + b0) < bsc) &
                         (Me + 1e-6 * Me + 1e-14 * (Me + b0) < btc);
    bool brk = false;
    if (!certain) {
        Mc *= sqrt(os2);
        Sc += sqrt(mt2) * 4.0 * bb;
        Mc += sqrt(mn2) * 2.0 * bb;
        brk = val(Sc) > bsc || val(Mc) > btc;
    }
    brk = br.decide(kBrBreak, brk);
    h[0] = nx; h[1] = ny; h[2] = nz;
    h[3] = ox; h[4] = oy; h[5] = oz;
    h[6] = mnx; h[7] = mny; h[8] = mnz;
    h[9] = mtx; h[10] = mty; h[11] = mtz;
    return brk;
}

// --------------------------------------------------------------- contact --
// One Hertz-Mindlin contact seen from particle i (d_CalculateForcesDEM_21 and
// dd_*ppHM_2, md_calculateforcesdem3.cu:202-425) for a pair known to overlap
// (rijm < ri + rj).  Returns the force and moment contributions and the new
// tangential overlap.
template <class T, class BR = Decide>
__device__ __forceinline__ void contact_eval_t(const MatT<T>& m, const DevParams& p,
                                               const Vec4<T>& pi, const Vec4<T>& vi,
                                               const Vec4<T>& wi, const Vec4<T>& pj,
                                               const Vec4<T>& vj, const Vec4<T>& wj,
                                               const T (&oprev)[3], const T& wim, bool roll,
                                               double dtI, T (&c)[6], T (&onew)[3],
                                               BR&& br = BR()) {
    const T ri = pi.w, mi = vi.w, rj = pj.w, mj = vj.w;
    const T rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
    const T rijm = sqrt(rx * rx + ry * ry + rz * rz);
    const T rsum = ri + rj;
    const T inv = 1.0 / rijm;
    const T nx = rx * inv, ny = ry * inv, nz = rz * inv;
    const T Rst = ri * rj / (ri + rj);
    const T meff = mi * mj / (mi + mj);
    T c1 = sqrt(Rst * fabs(rsum - rijm));
    const T Kn = 2.0 * m.hmE * c1;
    const T Kt = 8.0 * m.hmG * c1;
    const T wsx = rj * wj.x + ri * wi.x;
    const T wsy = rj * wj.y + ri * wi.y;
    const T wsz = rj * wj.z + ri * wi.z;
    const T vx = vj.x - vi.x + (ny * wsz - nz * wsy);
    const T vy = vj.y - vi.y + (nz * wsx - nx * wsz);
    const T vz = vj.z - vi.z + (nx * wsy - ny * wsx);
    const T vn = nx * vx + ny * vy + nz * vz;
    const T fn = (rijm - rsum) * Kn * kTwoThirds;
    const T fd = kDamp * m.pA * vn * sqrt(Kn * meff);
    const T vtx = vx - nx * vn, vty = vy - ny * vn, vtz = vz - nz * vn;
    const T o0 = oprev[0], o1 = oprev[1], o2 = oprev[2];
    c1 = nx * o0 + ny * o1 + nz * o2;
    T ox = o0 - nx * c1, oy = o1 - ny * c1, oz = o2 - nz * c1;
    c1 = ox * ox + oy * oy + oz * oz;
    if (val(c1) > 1e-18) {
        c1 = 1.0 / c1;
        c1 *= o0 * o0 + o1 * o1 + o2 * o2;
        c1 = sqrt(c1);
        ox *= c1; oy *= c1; oz *= c1;
    }
    ox += vtx * p.dt; oy += vty * p.dt; oz += vtz * p.dt;
    T ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
    const T ftm = ftx * ftx + fty * fty + ftz * ftz;
    if (br.decide(kBrSlip, val(ftm) > val(m.mu * m.mu * fn * fn))) {
        c1 = m.mu * fabs(fn) * (1.0 / sqrt(ftm));
        ftx *= c1; fty *= c1; ftz *= c1;
        c1 = 1.0 / Kt;
        ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
    } else {
        c1 = kDamp * m.pA * sqrt(Kt * meff);
        ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
    }
    onew[0] = ox; onew[1] = oy; onew[2] = oz;
    c[0] = nx * (fn + fd) + ftx;
    c[1] = ny * (fn + fd) + fty;
    c[2] = nz * (fn + fd) + ftz;
    const T tqx = (ny * ftz - nz * fty) * ri;
    const T tqy = (nz * ftx - nx * ftz) * ri;
    const T tqz = (nx * fty - ny * ftx) * ri;
    if (roll) {
        if (br.decide(kBrChat, rolling_chatter(m.mur, fn, wim, dtI))) {
            T fx, fy, fz;
            br.spin(wi, fx, fy, fz);
            const T fw = fx * fx + fy * fy + fz * fz;
            const T cr = -m.mur * fabs(fn) * (1.0 / sqrt(fw));
            c[3] = tqx + fx * cr;
            c[4] = tqy + fy * cr;
            c[5] = tqz + fz * cr;
        } else {
            const T cr = -m.mur * fabs(fn) * (1.0 / sqrt(wim));
            c[3] = tqx + wi.x * cr; c[4] = tqy + wi.y * cr; c[5] = tqz + wi.z * cr;
        }
    } else {
        c[3] = tqx; c[4] = tqy; c[5] = tqz;
    }
}

// ------------------------------------------------------------------ wall --
template <class T>
struct WallOutT {
    T fnx, fny, fnz, ftx, fty, ftz, mtx, mty, mtz, max_, may, maz, ox, oy, oz;
};

// Particle-platen Hertz-Mindlin, dd_Calculate_fijn_pwHM_2 +
// dd_Calculate_fijtmijt_pwHM_2 (md_compression_hm.cu:20-154): R* = r_i,
// m* = m_i, previous tangential overlap (opx, opy, opz) including the
// reference's o.x*o.y rescale typo (:91).
template <class T, class BR = Decide>
__device__ __forceinline__ void wall_contact_t(const MatT<T>& m, const DevParams& p, const T& ri,
                                               const T& mi, const Vec4<T>& vi, const Vec4<T>& wi,
                                               double nwx, double nwy, double nwz, double wvx,
                                               double wvy, double wvz, const T& d, const T& opx,
                                               const T& opy, const T& opz, double dtI,
                                               WallOutT<T>& o, BR&& br = BR()) {
    T c1 = sqrt(ri * fabs(ri - d));
    const T Kn = 2.0 * m.Ew * c1;
    const T Kt = 8.0 * m.Gw * c1;
    const T vx = wvx - vi.x + ri * (nwy * wi.z - nwz * wi.y);
    const T vy = wvy - vi.y + ri * (nwz * wi.x - nwx * wi.z);
    const T vz = wvz - vi.z + ri * (nwx * wi.y - nwy * wi.x);
    const T vn = nwx * vx + nwy * vy + nwz * vz;
    const T fnm = (d - ri) * Kn * kTwoThirds;
    c1 = kDamp * m.pA * vn * sqrt(Kn * mi);
    o.fnx = nwx * (fnm + c1); o.fny = nwy * (fnm + c1); o.fnz = nwz * (fnm + c1);
    const T vtx = vx - nwx * vn, vty = vy - nwy * vn, vtz = vz - nwz * vn;
    c1 = nwx * opx + nwy * opy + nwz * opz;
    T ox = opx - nwx * c1, oy = opy - nwy * c1, oz = opz - nwz * c1;
    c1 = ox * ox + oy * oy + oz * oz;
    if (val(c1) > 1e-18) {
        c1 = 1.0 / c1;
        c1 *= opx * opy + opy * opy + opz * opz;
        c1 = sqrt(c1);
        ox *= c1; oy *= c1; oz *= c1;
    }
    ox += vtx * p.dt; oy += vty * p.dt; oz += vtz * p.dt;
    T ftx = ox * Kt, fty = oy * Kt, ftz = oz * Kt;
    const T ftm = ftx * ftx + fty * fty + ftz * ftz;
    if (br.decide(kBrSlip, val(ftm) > val(m.mu * m.mu * fnm * fnm))) {
        c1 = m.mu * fabs(fnm) * (1.0 / sqrt(ftm));
        ftx *= c1; fty *= c1; ftz *= c1;
        c1 = 1.0 / Kt;
        ox = ftx * c1; oy = fty * c1; oz = ftz * c1;
    } else {
        c1 = kDamp * m.pA * sqrt(Kt * mi);
        ftx += vtx * c1; fty += vty * c1; ftz += vtz * c1;
    }
    o.ox = ox; o.oy = oy; o.oz = oz;
    o.ftx = ftx; o.fty = fty; o.ftz = ftz;
    o.mtx = (nwy * ftz - nwz * fty) * ri;
    o.mty = (nwz * ftx - nwx * ftz) * ri;
    o.mtz = (nwx * fty - nwy * ftx) * ri;
    const T wim = wi.x * wi.x + wi.y * wi.y + wi.z * wi.z;
    if (br.decide(kBrRoll, val(wim) > 1e-18)) {
        if (br.decide(kBrChat, rolling_chatter(m.mur, fnm, wim, dtI))) {
            T fx, fy, fz;
            br.spin(wi, fx, fy, fz);
            const T fw = fx * fx + fy * fy + fz * fz;
            c1 = -m.mur * fabs(fnm) * (1.0 / sqrt(fw));
            o.max_ = fx * c1; o.may = fy * c1; o.maz = fz * c1;
        } else {
            c1 = -m.mur * fabs(fnm) * (1.0 / sqrt(wim));
            o.max_ = wi.x * c1; o.may = wi.y * c1; o.maz = wi.z * c1;
        }
    } else {
        o.max_ = zero_like(ri); o.may = zero_like(ri); o.maz = zero_like(ri);
    }
}

}  // namespace rf
