// rf_fire.hip -- FIRE relaxation of a packing on the GPU (sample preparation).
//
// Replaces the relaxation loop of h_CreateCylinderSample
// (md_createcylindersample.cu:353-420).  Per step, three launches on one
// stream with the FIRE controller on the device:
//   k_fire_force  d_CalculateForces (md_calculateforces.cu:20-67), the
//                 particle's F.V for d_FdotVEntire (md_firerelaxation.cu:375+)
//                 and its worst overlap; block partials
//   k_fire_ctrl   the host loop's power sum and FIRE update
//                 (md_createcylindersample.cu:367-397, SetFIREData
//                 md_firerelaxation.cu:72-86), overlap check and loop condition
//                 (:353, :403-417)
//   k_fire_move   d_CalculateDecrementsHalfStepFIRE (md_firerelaxation.cu:58-70),
//                 d_CalculateIncrementsFIRE (:18-56),
//                 d_CylinderRestrictionZ (md_cylinderrestriction.cu:69-113),
//                 d_ReduceParticleNumber (md_reduceparticlenumber.cu:18-39)
// Neighbour lists (all j within 2 r_max + skin, ascending j) are rebuilt
// when a particle has moved half the skin; they never change a result.
// Order contract: the force on i sums contributing j in ascending j; power
// per particle (Fx Vx + Fy Vy) + Fz Vz, blocks of 256 by the tree
// v[t] += v[t+off] (off 128..1), block partials by 1024 strided running sums
// and a 1024-wide tree; worst pair = smallest (rho - a)/a, ties to the
// smallest i (then j).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/rockfit_dem.h"
#include "rf_internal.h"

namespace {

constexpr int kFB = 256;       // particles per block (force, move)
constexpr int kFC = 1024;      // controller threads
constexpr int kFList = 64;     // neighbour-list capacity per particle
constexpr double kPiF = 3.1415926535897932384626433832795;

struct FireParams {
    int64_t n;
    int32_t nblk, K;
    double D, dt0, dtmax, dtmin, grow, shrink, ashrink, alpha0;
    int64_t ndelay, npmax, max_steps, cap;
    int32_t container, noinertia, check_every, delete_after, law, pad_;
    double cen[3], R0, H0, hide_z, hidden[3], gap_tol;
    double list_r2, skin_half2;
    double lo[3], inv_cell;
    int32_t nc[3];
    int64_t ncell;
};

struct FireCtrl {
    double dt, alpha, min_gap, power;
    int64_t step, npos, nneg, removed, del_j;
    int32_t halt, move, neg, rebuild, err, reason, pad0, pad1;
};

struct FireDev {
    double4* pos;    // x, y, z, r
    double4* vel;    // vx, vy, vz, m
    double4* frc;    // Fx, Fy, Fz, 0
    double4* rb;     // positions at the last list build
    uint8_t* alive;
    int32_t* nb;     // [i*K + q], ascending
    int32_t* nn;
    int32_t* key;
    int32_t* cnt;
    int32_t* start;
    int32_t* fill;
    int32_t* sorted;
    double* pw;      // block power partials
    double* gmin;    // block min gap
    int32_t* gi;     // particle of the block's worst pair
    int32_t* gj;     // its partner
};

__device__ __forceinline__ int32_t fire_cell(const FireParams& p, double x, double y, double z) {
    const double c[3] = {x, y, z};
    int32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double f = floor((c[d] - p.lo[d]) * p.inv_cell);
        f = f < 0.0 ? 0.0 : (f > (double)(p.nc[d] - 1) ? (double)(p.nc[d] - 1) : f);
        q[d] = (int32_t)f;
    }
    return (q[2] * p.nc[1] + q[1]) * p.nc[0] + q[0];
}

// ---------------------------------------------------------- list building --
__global__ void k_fire_keys(FireDev d, FireParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 q = d.pos[i];
    const int32_t c = d.alive[i] ? fire_cell(p, q.x, q.y, q.z) : -1;
    d.key[i] = c;
    if (c >= 0) atomicAdd(&d.cnt[c], 1);
}

__global__ void k_fire_zero(int32_t* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// Exclusive scan of the cell counts by one block: each thread scans a
// contiguous chunk, then the chunk totals are scanned in LDS.
__global__ void k_fire_scan(FireDev d, FireParams p) {
    __shared__ int32_t tot[kFC];
    const int64_t per = (p.ncell + kFC - 1) / kFC;
    const int64_t b = threadIdx.x * per, e = std::min<int64_t>(b + per, p.ncell);
    int32_t s = 0;
    for (int64_t c = b; c < e; ++c) s += d.cnt[c];
    tot[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kFC; off <<= 1) {
        const int32_t v = threadIdx.x >= off ? tot[threadIdx.x - off] : 0;
        __syncthreads();
        tot[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t run = tot[threadIdx.x] - s;
    for (int64_t c = b; c < e; ++c) {
        d.start[c] = run;
        d.fill[c] = run;
        run += d.cnt[c];
    }
}

__global__ void k_fire_scatter(FireDev d, FireParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int32_t c = d.key[i];
    if (c < 0) return;
    d.sorted[atomicAdd(&d.fill[c], 1)] = (int32_t)i;
}

__global__ void k_fire_lists(FireDev d, FireParams p, FireCtrl* ctrl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 pi = d.pos[i];
    d.rb[i] = pi;
    int32_t m = 0;
    if (d.alive[i]) {
        const int32_t c = d.key[i];
        const int32_t cx = c % p.nc[0], cy = (c / p.nc[0]) % p.nc[1], cz = c / (p.nc[0] * p.nc[1]);
        int32_t* row = d.nb + i * p.K;
        bool over = false;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int32_t qx = cx + dx, qy = cy + dy, qz = cz + dz;
                    if (qx < 0 || qy < 0 || qz < 0 || qx >= p.nc[0] || qy >= p.nc[1] || qz >= p.nc[2])
                        continue;
                    const int32_t cc = (qz * p.nc[1] + qy) * p.nc[0] + qx;
                    const int32_t s0 = d.start[cc], s1 = s0 + d.cnt[cc];
                    for (int32_t q = s0; q < s1; ++q) {
                        const int32_t j = d.sorted[q];
                        if (j == (int32_t)i) continue;
                        const double4 pj = d.pos[j];
                        const double ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
                        if (!(ex * ex + ey * ey + ez * ez < p.list_r2)) continue;
                        if (m == p.K) {
                            over = true;
                            continue;
                        }
                        // insertion into the ascending row
                        int32_t w = m++;
                        while (w > 0 && row[w - 1] > j) {
                            row[w] = row[w - 1];
                            --w;
                        }
                        row[w] = j;
                    }
                }
        if (over) atomicOr(&ctrl->err, 1);
    }
    d.nn[i] = m;
}

// ------------------------------------------------------------ one step ----
__global__ __launch_bounds__(kFB) void k_fire_force(FireDev d, FireParams p, FireCtrl* ctrl) {
    __shared__ double s_p[kFB];
    __shared__ double s_g[kFB];
    __shared__ int32_t s_i[kFB];
    __shared__ int32_t s_j[kFB];
    if (ctrl->halt || ctrl->rebuild) return;  // uniform
    const int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0, pw = 0.0, g = 0.0;
    int32_t gj = -1;
    if (i < p.n && d.alive[i]) {
        const double4 pi = d.pos[i];
        const int32_t m = d.nn[i];
        const int32_t* row = d.nb + i * p.K;
        for (int32_t q = 0; q < m; ++q) {
            const int32_t j = row[q];
            const double4 pj = d.pos[j];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double rr = dx * dx + dy * dy + dz * dz;
            const double a = pi.w + pj.w;
            const double aa = a * a;
            if (rr < aa) {
                double fm;
                if (p.law == RF_FIRE_LAW_HARMONIC) {
                    const double rho = sqrt(rr);
                    fm = -p.D * ((a - rho) / rho);
                } else {
                    const double q1 = aa * (1.0 / rr);
                    fm = -p.D * q1;
                }
                fx += fm * dx;
                fy += fm * dy;
                fz += fm * dz;
                const double gg = (sqrt(rr) - a) / a;
                if (gg < g) {
                    g = gg;
                    gj = j;
                }
            }
        }
        const double4 v = d.vel[i];
        pw = fx * v.x + fy * v.y + fz * v.z;
    }
    if (i < p.n) d.frc[i] = make_double4(fx, fy, fz, 0.0);
    s_p[threadIdx.x] = pw;
    s_g[threadIdx.x] = g;
    s_i[threadIdx.x] = (int32_t)i;
    s_j[threadIdx.x] = gj;
    __syncthreads();
    for (int off = kFB / 2; off >= 1; off >>= 1) {
        if (threadIdx.x < off) {
            const int t = threadIdx.x;
            s_p[t] = s_p[t] + s_p[t + off];
            // worst gap; equal gaps go to the lower particle
            if (s_g[t + off] < s_g[t] || (s_g[t + off] == s_g[t] && s_i[t + off] < s_i[t])) {
                s_g[t] = s_g[t + off];
                s_i[t] = s_i[t + off];
                s_j[t] = s_j[t + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        d.pw[blockIdx.x] = s_p[0];
        d.gmin[blockIdx.x] = s_g[0];
        d.gi[blockIdx.x] = s_i[0];
        d.gj[blockIdx.x] = s_j[0];
    }
}

__global__ __launch_bounds__(kFC) void k_fire_ctrl(FireDev d, FireParams p, FireCtrl* ctrl) {
    __shared__ double s_p[kFC];
    __shared__ double s_g[kFC];
    __shared__ int32_t s_i[kFC];
    __shared__ int32_t s_j[kFC];
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = (ctrl->halt || ctrl->rebuild) ? 1 : 0;
    __syncthreads();
    if (s_stop) {
        if (threadIdx.x == 0) ctrl->move = 0;
        return;
    }
    double acc = 0.0, g = 0.0;
    int32_t gi = 0x7fffffff, gj = -1;
    for (int64_t b = threadIdx.x; b < p.nblk; b += kFC) {
        acc = acc + d.pw[b];
        if (d.gmin[b] < g || (d.gmin[b] == g && d.gi[b] < gi)) {
            g = d.gmin[b];
            gi = d.gi[b];
            gj = d.gj[b];
        }
    }
    s_p[threadIdx.x] = acc;
    s_g[threadIdx.x] = g;
    s_i[threadIdx.x] = gi;
    s_j[threadIdx.x] = gj;
    __syncthreads();
    for (int off = kFC / 2; off >= 1; off >>= 1) {
        if (threadIdx.x < off) {
            const int t = threadIdx.x;
            s_p[t] = s_p[t] + s_p[t + off];
            if (s_g[t + off] < s_g[t] || (s_g[t + off] == s_g[t] && s_i[t + off] < s_i[t])) {
                s_g[t] = s_g[t + off];
                s_i[t] = s_i[t + off];
                s_j[t] = s_j[t + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    FireCtrl c = *ctrl;
    const double P = s_p[0];
    c.power = P;
    int neg = 0;
    if (P > 0) {
        ++c.npos;
        c.nneg = 0;
        if (c.npos > p.ndelay) {
            c.dt = (c.dt * p.grow < p.dtmax) ? c.dt * p.grow : p.dtmax;
            c.alpha *= p.ashrink;
        }
    } else {
        c.npos = 0;
        ++c.nneg;
        if (c.nneg > p.npmax && c.min_gap >= p.gap_tol) {
            c.halt = 1;
            c.reason = 2;
            c.move = 0;
            *ctrl = c;
            return;
        }
        if (c.step > p.ndelay) {
            c.dt = (c.dt * p.shrink > p.dtmin) ? c.dt * p.shrink : p.dtmin;
            c.alpha = p.alpha0;
        }
        neg = 1;
    }
    c.neg = neg;
    c.move = 1;
    c.del_j = -1;
    if (p.check_every > 0 && c.step % p.check_every == 0) {
        c.min_gap = s_g[0];
        if (p.delete_after >= 0 && c.step > p.delete_after && c.min_gap < p.gap_tol && s_j[0] >= 0) {
            c.del_j = s_j[0];
            ++c.removed;
        }
    }
    ++c.step;
    if (!(c.step < p.max_steps || c.min_gap < p.gap_tol)) {
        c.halt = 1;
        c.reason = 1;
    } else if (c.step >= p.cap) {
        c.halt = 1;
        c.reason = 3;
    }
    *ctrl = c;
}

__global__ __launch_bounds__(kFB) void k_fire_move(FireDev d, FireParams p, FireCtrl* ctrl) {
    if (!ctrl->move) return;  // uniform
    const int64_t i = (int64_t)blockIdx.x * kFB + threadIdx.x;
    if (i >= p.n || !d.alive[i]) return;
    const double dt = ctrl->dt, alpha = ctrl->alpha;
    const int neg = ctrl->neg;
    const int64_t del = ctrl->del_j;
    double4 r = d.pos[i];
    double4 v = d.vel[i];
    const double4 F = d.frc[i];
    if (neg) {
        const double h = 0.5 * dt;
        r.x -= v.x * h;
        r.y -= v.y * h;
        r.z -= v.z * h;
        v.x = 0.0; v.y = 0.0; v.z = 0.0;
    }
    double vx, vy, vz;
    if (p.noinertia) {
        const double dtm = p.dt0 / v.w;
        vx = F.x * dtm; vy = F.y * dtm; vz = F.z * dtm;
    } else {
        const double dtm = dt / v.w;
        vx = v.x + F.x * dtm; vy = v.y + F.y * dtm; vz = v.z + F.z * dtm;
    }
    const double vv = vx * vx + vy * vy + vz * vz;
    const double f = F.x * F.x + F.y * F.y + F.z * F.z;
    const double vdf = f > 1e-12 ? alpha * sqrt(vv * (1.0 / f)) : alpha * sqrt(vv) * 1e6;
    vx = (1.0 - alpha) * vx + vdf * F.x;
    vy = (1.0 - alpha) * vy + vdf * F.y;
    vz = (1.0 - alpha) * vz + vdf * F.z;
    r.x += vx * dt;
    r.y += vy * dt;
    r.z += vz * dt;
    if (p.container == RF_CONTAINER_CYLINDER_Z || p.container == RF_CONTAINER_CYLINDER_Z_ABSORB) {
        const bool absorb = p.container == RF_CONTAINER_CYLINDER_Z_ABSORB;
        const double ex = r.x - p.cen[0], ey = r.y - p.cen[1], ez = r.z - p.cen[2];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        const double rr = ex * ex + ey * ey;
        if (ez < p.hide_z - p.cen[2]) {
            if (rr > p.R0 * p.R0) {
                const double q = 1.0 / sqrt(rr);
                sx = ex * (p.R0 * q - 1.0);
                sy = ey * (p.R0 * q - 1.0);
                if (absorb) {
                    const double ux = ex * q, uy = ey * q;
                    const double vr = vx * ux + vy * uy;
                    if (vr > 0.0) {
                        vx -= vr * ux;
                        vy -= vr * uy;
                    }
                } else {
                    vx *= -1.0;
                    vy *= -1.0;
                }
            }
            if (ez > p.H0) {
                sz = p.H0 - ez;
                vz = absorb ? (vz > 0.0 ? 0.0 : vz) : -vz;
            }
            if (ez < -p.H0) {
                sz = -p.H0 - ez;
                vz = absorb ? (vz < 0.0 ? 0.0 : vz) : -vz;
            }
            r.x += sx;
            r.y += sy;
            r.z += sz;
        }
    }
    if (i == del) {  // d_ReduceParticleNumber: after this step's move
        r.x = p.hidden[0]; r.y = p.hidden[1]; r.z = p.hidden[2];
        vx = 0.0; vy = 0.0; vz = 0.0;
        d.frc[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        d.alive[i] = 0;
    }
    d.pos[i] = r;
    d.vel[i] = make_double4(vx, vy, vz, v.w);
    const double4 b = d.rb[i];
    const double ex = r.x - b.x, ey = r.y - b.y, ez = r.z - b.z;
    if (i == del || ex * ex + ey * ey + ez * ez > p.skin_half2) ctrl->rebuild = 1;
}

__global__ void k_fire_clear(FireCtrl* ctrl) {
    if (threadIdx.x == 0) ctrl->rebuild = 0;
}

// Sample in, sample out for the static-memory entry point: SoA host copies
// staged in device memory, packed into / unpacked from the particle state.
// The mass is the same expression, in the same order, as rf_fire_relax's
// host-side one.
__global__ void k_fire_pack(FireDev d, const double* x, const double* y, const double* z,
                            const double* r, const double* v, int has_v, double density,
                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ri = r[i];
    const double m = density * (4.0 / 3.0) * kPiF * ri * ri * ri;
    d.pos[i] = make_double4(x[i], y[i], z[i], ri);
    d.vel[i] = has_v ? make_double4(v[i], v[n + i], v[2 * n + i], m) : make_double4(0.0, 0.0, 0.0, m);
    d.alive[i] = 1;
}

__global__ void k_fire_unpack(FireDev d, double* x, double* y, double* z, double* v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 q = d.pos[i], u = d.vel[i];
    x[i] = q.x;
    y[i] = q.y;
    z[i] = q.z;
    v[i] = u.x;
    v[n + i] = u.y;
    v[2 * n + i] = u.z;
}

unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define FIRE_TRY(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return rf::report(RF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

struct FireMem {
    std::vector<void*> ptrs;
    ~FireMem() {
        for (void* q : ptrs) hipFree(q);
    }
    template <class T>
    hipError_t get(T** out, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        ptrs.push_back(q);
        *out = (T*)q;
        return hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
    }
};

rf_status build_lists(FireDev& d, const FireParams& p, FireCtrl* ctrl, hipStream_t st) {
    hipLaunchKernelGGL(k_fire_zero, dim3(grid_of(p.ncell, 256)), dim3(256), 0, st, d.cnt, p.ncell);
    hipLaunchKernelGGL(k_fire_keys, dim3(grid_of(p.n, 256)), dim3(256), 0, st, d, p);
    hipLaunchKernelGGL(k_fire_scan, dim3(1), dim3(kFC), 0, st, d, p);
    hipLaunchKernelGGL(k_fire_scatter, dim3(grid_of(p.n, 256)), dim3(256), 0, st, d, p);
    hipLaunchKernelGGL(k_fire_lists, dim3(grid_of(p.n, 128)), dim3(128), 0, st, d, p, ctrl);
    hipLaunchKernelGGL(k_fire_clear, dim3(1), dim3(64), 0, st, ctrl);
    FIRE_TRY(hipGetLastError());
    return RF_OK;
}

// Checks the arguments and fills the run's constants from cfg and the
// sample (host arrays).
rf_status fire_setup(int64_t n, const double* x, const double* y, const double* z, const double* r,
                     const rf_fire_cfg* cfg, FireParams& p) {
    if (n <= 0 || !x || !y || !z || !r || !cfg)
        return rf::report(RF_ERR_ARG, "rf_fire_relax: bad arguments");
    if (n >= (int64_t)1 << 31) return rf::report(RF_ERR_ARG, "rf_fire_relax: n too large");
    if (cfg->law != RF_FIRE_LAW_REFERENCE && cfg->law != RF_FIRE_LAW_HARMONIC)
        return rf::report(RF_ERR_ARG, "rf_fire_relax: unknown force law %d", (int)cfg->law);
    if (cfg->container < RF_CONTAINER_NONE || cfg->container > RF_CONTAINER_CYLINDER_Z_ABSORB)
        return rf::report(RF_ERR_ARG, "rf_fire_relax: unknown container %d", (int)cfg->container);
    if (!(cfg->dt0 > 0.0) || !(cfg->density > 0.0) || cfg->max_steps < 0)
        return rf::report(RF_ERR_ARG, "rf_fire_relax: dt0, density must be > 0, max_steps >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return rf::report(RF_ERR_NODEVICE, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return rf::report(RF_ERR_ARG, "bad device");
    FIRE_TRY(hipSetDevice(cfg->device));

    double rmax = 0.0;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i) {
        if (!(r[i] > 0.0)) return rf::report(RF_ERR_ARG, "radius %lld not positive", (long long)i);
        rmax = std::max(rmax, r[i]);
        const double q[3] = {x[i], y[i], z[i]};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], q[k]);
            hi[k] = std::max(hi[k], q[k]);
        }
    }
    if (cfg->container == RF_CONTAINER_CYLINDER_Z || cfg->container == RF_CONTAINER_CYLINDER_Z_ABSORB) {
        const double ext[3] = {cfg->radius, cfg->radius, cfg->half_height};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], cfg->center[k] - ext[k]);
            hi[k] = std::max(hi[k], cfg->center[k] + ext[k]);
        }
    }
    const double skin = cfg->skin >= 0.0 ? cfg->skin : 0.25 * rmax;
    const double cut = 2.0 * rmax + skin;

    p = FireParams{};
    p.n = n;
    p.nblk = (int32_t)grid_of(n, kFB);
    p.K = kFList;
    p.D = cfg->D;
    p.dt0 = cfg->dt0;
    p.dtmax = cfg->dtmax_factor * cfg->dt0;
    p.dtmin = cfg->dtmin_factor * cfg->dt0;
    p.grow = cfg->dtgrow;
    p.shrink = cfg->dtshrink;
    p.ashrink = cfg->alpha_shrink;
    p.alpha0 = cfg->alpha0;
    p.ndelay = cfg->ndelay;
    p.npmax = cfg->np_negative_max;
    p.max_steps = cfg->max_steps;
    p.cap = cfg->hard_cap > 0 ? cfg->hard_cap : 10 * cfg->max_steps;
    p.container = cfg->container;
    p.noinertia = (cfg->quirks & RF_QUIRK_FIRE_NO_INERTIA) ? 1 : 0;
    p.law = cfg->law;
    p.check_every = cfg->check_every;
    p.delete_after = cfg->delete_after;
    for (int k = 0; k < 3; ++k) {
        p.cen[k] = cfg->center[k];
        p.hidden[k] = cfg->hidden[k];
    }
    p.R0 = cfg->radius;
    p.H0 = cfg->half_height;
    p.hide_z = cfg->hide_z;
    p.gap_tol = cfg->gap_tol;
    p.list_r2 = cut * cut;
    p.skin_half2 = 0.25 * skin * skin;
    // Grid over the initial extent plus two cells; particles beyond it are
    // clamped into the border cells, which keeps every pair within `cut` in
    // neighbouring cells.  At most 256 cells a side.
    double cell = cut;
    for (int k = 0; k < 3; ++k) cell = std::max(cell, (hi[k] - lo[k] + 4.0 * cut) / 256.0);
    p.inv_cell = 1.0 / cell;
    p.ncell = 1;
    for (int k = 0; k < 3; ++k) {
        p.lo[k] = lo[k] - 2.0 * cell;
        p.nc[k] = (int32_t)std::floor((hi[k] + 2.0 * cell - p.lo[k]) / cell) + 1;
        p.ncell *= p.nc[k];
    }
    return RF_OK;
}

// The FIRE controller's starting state.
FireCtrl fire_ctrl0(const rf_fire_cfg* cfg, const FireParams& p) {
    FireCtrl hc{};
    hc.dt = cfg->dt0;
    hc.alpha = cfg->alpha0;
    hc.min_gap = 0.0;
    hc.del_j = -1;
    if (!(0 < p.max_steps || hc.min_gap < p.gap_tol)) {
        hc.halt = 1;
        hc.reason = 1;
    } else if (p.cap <= 0) {
        hc.halt = 1;
        hc.reason = 3;
    }
    return hc;
}

// The relaxation loop on stream `st`: 256 steps per host wait, lists
// rebuilt when the device asks.  `hc` holds the uploaded starting state and
// receives the final one.
rf_status fire_loop(FireDev& d, const FireParams& p, FireCtrl* ctrl, FireCtrl& hc, hipStream_t st) {
    const unsigned gb = (unsigned)p.nblk;
    rf_status rs = build_lists(d, p, ctrl, st);
    if (rs) return rs;
    int64_t last_step = -1;
    while (!hc.halt) {
        for (int s = 0; s < 256; ++s) {
            hipLaunchKernelGGL(k_fire_force, dim3(gb), dim3(kFB), 0, st, d, p, ctrl);
            hipLaunchKernelGGL(k_fire_ctrl, dim3(1), dim3(kFC), 0, st, d, p, ctrl);
            hipLaunchKernelGGL(k_fire_move, dim3(gb), dim3(kFB), 0, st, d, p, ctrl);
        }
        FIRE_TRY(hipGetLastError());
        FIRE_TRY(hipMemcpyAsync(&hc, ctrl, sizeof hc, hipMemcpyDeviceToHost, st));
        FIRE_TRY(hipStreamSynchronize(st));
        if (hc.err) return rf::report(RF_ERR_OVERFLOW, "FIRE neighbour list longer than %d", kFList);
        if (hc.halt) break;
        if (hc.rebuild) {
            rs = build_lists(d, p, ctrl, st);
            if (rs) return rs;
        } else if (hc.step == last_step) {
            return rf::report(RF_ERR_STATE, "FIRE loop halted without a reason");
        }
        last_step = hc.step;
    }
    return RF_OK;
}

void fire_result(const FireCtrl& hc, rf_fire_result* out) {
    out->steps = hc.step;
    out->np_positive = hc.npos;
    out->np_negative = hc.nneg;
    out->removed = hc.removed;
    out->dt = hc.dt;
    out->alpha = hc.alpha;
    out->min_gap = hc.min_gap;
    out->power = hc.power;
    out->stop_reason = hc.reason;
}

}  // namespace

extern "C" {

rf_status rf_fire_defaults(rf_fire_cfg* c) {
    if (!c) return rf::report(RF_ERR_ARG, "null cfg");
    *c = rf_fire_cfg{};
    c->alpha0 = 0.1;          // md_createcylindersample.cu:349
    c->dtgrow = 1.1;          // SetFIREData, md_firerelaxation.cu:72-86
    c->dtshrink = 0.5;
    c->alpha_shrink = 0.99;
    c->dtmax_factor = 100.0;
    c->dtmin_factor = 0.001;
    c->ndelay = 20;
    c->np_negative_max = 30000;
    c->gap_tol = -1e300;
    c->check_every = 1;
    c->delete_after = -1;
    c->quirks = RF_QUIRK_FIRE_NO_INERTIA;
    c->skin = -1.0;
    return RF_OK;
}

rf_status rf_fire_relax(int64_t n, double* x, double* y, double* z, const double* r,
                        const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out) {
    if (out) *out = rf_fire_result{};
    if (!out) return rf::report(RF_ERR_ARG, "rf_fire_relax: bad arguments");
    FireParams p;
    rf_status rs = fire_setup(n, x, y, z, r, cfg, p);
    if (rs) return rs;

    FireMem mem;
    FireDev d{};
    FireCtrl* ctrl = nullptr;
    FIRE_TRY(mem.get(&d.pos, n));
    FIRE_TRY(mem.get(&d.vel, n));
    FIRE_TRY(mem.get(&d.frc, n));
    FIRE_TRY(mem.get(&d.rb, n));
    FIRE_TRY(mem.get(&d.alive, n));
    FIRE_TRY(mem.get(&d.nb, n * kFList));
    FIRE_TRY(mem.get(&d.nn, n));
    FIRE_TRY(mem.get(&d.key, n));
    FIRE_TRY(mem.get(&d.cnt, p.ncell));
    FIRE_TRY(mem.get(&d.start, p.ncell));
    FIRE_TRY(mem.get(&d.fill, p.ncell));
    FIRE_TRY(mem.get(&d.sorted, n));
    FIRE_TRY(mem.get(&d.pw, p.nblk));
    FIRE_TRY(mem.get(&d.gmin, p.nblk));
    FIRE_TRY(mem.get(&d.gi, p.nblk));
    FIRE_TRY(mem.get(&d.gj, p.nblk));
    FIRE_TRY(mem.get(&ctrl, 1));

    {
        std::vector<double4> hp(n), hv(n);
        std::vector<uint8_t> ha(n, 1);
        for (int64_t i = 0; i < n; ++i) {
            const double m = cfg->density * (4.0 / 3.0) * kPiF * r[i] * r[i] * r[i];
            hp[i] = make_double4(x[i], y[i], z[i], r[i]);
            hv[i] = velocity ? make_double4(velocity[i], velocity[n + i], velocity[2 * n + i], m)
                             : make_double4(0.0, 0.0, 0.0, m);
        }
        FIRE_TRY(hipMemcpy(d.pos, hp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        FIRE_TRY(hipMemcpy(d.vel, hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        FIRE_TRY(hipMemcpy(d.alive, ha.data(), n, hipMemcpyHostToDevice));
    }
    FireCtrl hc = fire_ctrl0(cfg, p);
    FIRE_TRY(hipMemcpy(ctrl, &hc, sizeof hc, hipMemcpyHostToDevice));

    hipStream_t st = nullptr;
    FIRE_TRY(hipStreamCreate(&st));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            if (s) hipStreamDestroy(s);
        }
    } sg{st};
    rs = fire_loop(d, p, ctrl, hc, st);
    if (rs) return rs;
    {
        std::vector<double4> hp(n), hv(n);
        FIRE_TRY(hipMemcpy(hp.data(), d.pos, n * sizeof(double4), hipMemcpyDeviceToHost));
        FIRE_TRY(hipMemcpy(hv.data(), d.vel, n * sizeof(double4), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            x[i] = hp[i].x;
            y[i] = hp[i].y;
            z[i] = hp[i].z;
            if (velocity) {
                velocity[i] = hv[i].x;
                velocity[n + i] = hv[i].y;
                velocity[2 * n + i] = hv[i].z;
            }
        }
    }
    fire_result(hc, out);
    return RF_OK;
}

}  // extern "C"

// ------------------------------------------------- static working memory --
// rf_fire_relax_noalloc's working memory: one array per GPU, loaded with the
// library's code object and held until the process ends.  Host code never
// uses the name `scratch` as an address: each device's address comes from
// hipGetSymbolAddress on that device's first call and is kept per device.
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];  // one call at a time per GPU
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch` (the device must be current;
// the caller holds g_scratch_mu[dev]).
rf_status scratch_base(int dev, unsigned char** out) {
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return rf::report(RF_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                              hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return RF_OK;
}

// Consecutive 256-byte-aligned pieces of one region; with base == nullptr
// it only measures.
struct Carve {
    unsigned char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* q = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
        return q;
    }
};

// The run's arrays (and the SoA staging of the host copies) in one region.
void fire_carve(Carve& m, const FireParams& p, FireDev& d, FireCtrl** ctrl, double** stage) {
    const size_t n = (size_t)p.n;
    d.pos = m.take<double4>(n);
    d.vel = m.take<double4>(n);
    d.frc = m.take<double4>(n);
    d.rb = m.take<double4>(n);
    d.alive = m.take<uint8_t>(n);
    d.nb = m.take<int32_t>(n * kFList);
    d.nn = m.take<int32_t>(n);
    d.key = m.take<int32_t>(n);
    d.cnt = m.take<int32_t>((size_t)p.ncell);
    d.start = m.take<int32_t>((size_t)p.ncell);
    d.fill = m.take<int32_t>((size_t)p.ncell);
    d.sorted = m.take<int32_t>(n);
    d.pw = m.take<double>((size_t)p.nblk);
    d.gmin = m.take<double>((size_t)p.nblk);
    d.gi = m.take<int32_t>((size_t)p.nblk);
    d.gj = m.take<int32_t>((size_t)p.nblk);
    *ctrl = m.take<FireCtrl>(1);
    *stage = m.take<double>(7 * n);  // x, y, z, r, velocity (3 n)
}

}  // namespace

extern "C" {

rf_status rf_fire_relax_noalloc(int64_t n, double* x, double* y, double* z, const double* r,
                                const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out) {
    if (out) *out = rf_fire_result{};
    if (!out) return rf::report(RF_ERR_ARG, "rf_fire_relax_noalloc: bad arguments");
    FireParams p;
    rf_status rs = fire_setup(n, x, y, z, r, cfg, p);  // also makes cfg->device current
    if (rs) return rs;
    if (cfg->device >= kMaxDevices) return rf::report(RF_ERR_ARG, "device %d too high", cfg->device);
    FireDev d{};
    FireCtrl* ctrl = nullptr;
    double* stage = nullptr;
    Carve probe{nullptr};
    fire_carve(probe, p, d, &ctrl, &stage);
    if (probe.off > kScratchBytes)
        return rf::report(RF_ERR_ARG,
                          "rf_fire_relax_noalloc: %lld particles (%lld grid cells) need %zu bytes "
                          "of working memory; the static array holds %zu",
                          (long long)n, (long long)p.ncell, probe.off, kScratchBytes);
    std::lock_guard<std::mutex> lock(g_scratch_mu[cfg->device]);
    unsigned char* base = nullptr;
    rs = scratch_base(cfg->device, &base);
    if (rs) return rs;
    Carve m{base};
    fire_carve(m, p, d, &ctrl, &stage);
    double *sx = stage, *sy = stage + n, *sz = stage + 2 * n, *sr = stage + 3 * n,
           *sv = stage + 4 * n;
    const hipStream_t st = nullptr;  // the null stream: no stream to create
    FIRE_TRY(hipMemcpy(sx, x, n * sizeof(double), hipMemcpyHostToDevice));
    FIRE_TRY(hipMemcpy(sy, y, n * sizeof(double), hipMemcpyHostToDevice));
    FIRE_TRY(hipMemcpy(sz, z, n * sizeof(double), hipMemcpyHostToDevice));
    FIRE_TRY(hipMemcpy(sr, r, n * sizeof(double), hipMemcpyHostToDevice));
    if (velocity) FIRE_TRY(hipMemcpy(sv, velocity, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fire_pack, dim3(grid_of(n, 256)), dim3(256), 0, st, d, sx, sy, sz, sr, sv,
                       velocity ? 1 : 0, cfg->density, n);
    FIRE_TRY(hipGetLastError());
    FireCtrl hc = fire_ctrl0(cfg, p);
    FIRE_TRY(hipMemcpy(ctrl, &hc, sizeof hc, hipMemcpyHostToDevice));
    rs = fire_loop(d, p, ctrl, hc, st);
    if (rs) return rs;
    hipLaunchKernelGGL(k_fire_unpack, dim3(grid_of(n, 256)), dim3(256), 0, st, d, sx, sy, sz, sv, n);
    FIRE_TRY(hipGetLastError());
    FIRE_TRY(hipMemcpy(x, sx, n * sizeof(double), hipMemcpyDeviceToHost));
    FIRE_TRY(hipMemcpy(y, sy, n * sizeof(double), hipMemcpyDeviceToHost));
    FIRE_TRY(hipMemcpy(z, sz, n * sizeof(double), hipMemcpyDeviceToHost));
    if (velocity) FIRE_TRY(hipMemcpy(velocity, sv, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    fire_result(hc, out);
    return RF_OK;
}

}  // extern "C"
