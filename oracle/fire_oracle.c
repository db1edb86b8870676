/*
 * fire_oracle.c -- CPU FP64 restatement of the reference's FIRE relaxation of
 * a packing (sample preparation).
 *
 * TEST INFRASTRUCTURE ONLY (see dem_oracle.h): tests/ and the CPU baseline
 * use it as the checker; the product never links it.
 *
 * Follows, per step:
 *   d_CalculateForces          md_calculateforces.cu:20-67  (pair force
 *                              -D a^2/rho^2 (R_j - R_i) for rho^2 < a^2;
 *                              a = r_i + r_j generalises 2 p_r)
 *   d_FdotVEntire + host sum   md_firerelaxation.cu:375+, md_createcylindersample.cu:367-373
 *   FIRE controller            md_createcylindersample.cu:374-397
 *   d_CalculateDecrementsHalfStepFIRE  md_firerelaxation.cu:58-70
 *   d_CalculateIncrementsFIRE  md_firerelaxation.cu:18-56 (A8 as a quirk)
 *   d_CylinderRestrictionZ     md_cylinderrestriction.cu:69-113
 *   overlap check + d_ReduceParticleNumber  md_createcylindersample.cu:403-417,
 *                              md_reduceparticlenumber.cu:18-39
 *   loop condition             md_createcylindersample.cu:353
 *   constants                  SetFIREData md_firerelaxation.cu:72-86, :349
 *
 * Order contract shared with the device code (no code is shared):
 *   - the force on i sums its partners j in ascending j (contributing pairs
 *     only), fs += fm * d, starting from 0.0;
 *   - power p_i = (Fx Vx + Fy Vy) + Fz Vz; blocks of 256 consecutive particles
 *     are added by the tree v[t] += v[t+off], off = 128..1 (0.0 padding); the
 *     block partials by 1024 strided running sums from 0.0 and a 1024-wide
 *     tree;
 *   - the worst pair is the smallest (rho - a)/a, ties to the smallest i then
 *     the smallest j.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dem_oracle.h"

#define FIRE_BLOCK 256
#define FIRE_CTRL 1024
#define ORA_PI 3.1415926535897932384626433832795

int ora_fire_defaults(rf_fire_cfg* c) {
    memset(c, 0, sizeof(*c));
    c->alpha0 = 0.1;
    c->dtgrow = 1.1;
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
    return 0;
}

typedef struct {
    int64_t n;
    double *x, *y, *z, *r, *vx, *vy, *vz, *fx, *fy, *fz, *m;
    unsigned char* alive;
    /* cell grid of the current positions */
    double lo[3], cell;
    int64_t nc[3];
    int64_t* head;  /* first particle of a cell, -1 = none */
    int64_t* next;
    double* gap;    /* per particle: min (rho - a)/a over its overlapping partners */
    int64_t* gapj;
} fire_state;

static int64_t cell_of(const fire_state* s, double x, double y, double z, int64_t* cx) {
    const double p[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        int64_t q = (int64_t)floor((p[d] - s->lo[d]) / s->cell);
        if (q < 0) q = 0;
        if (q >= s->nc[d]) q = s->nc[d] - 1;
        cx[d] = q;
    }
    return (cx[2] * s->nc[1] + cx[1]) * s->nc[0] + cx[0];
}

static int build_cells(fire_state* s) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmax = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        if (!s->alive[i]) continue;
        const double p[3] = {s->x[i], s->y[i], s->z[i]};
        for (int d = 0; d < 3; ++d) {
            if (p[d] < lo[d]) lo[d] = p[d];
            if (p[d] > hi[d]) hi[d] = p[d];
        }
        if (s->r[i] > rmax) rmax = s->r[i];
    }
    /* cells of at least the largest pair distance 2 r_max, at most 256 a side */
    double cell = 2.0 * rmax > 0.0 ? 2.0 * rmax : 1.0;
    for (int d = 0; d < 3; ++d) {
        if (hi[d] < lo[d]) { lo[d] = 0.0; hi[d] = 0.0; }
        if ((hi[d] - lo[d]) / 255.0 > cell) cell = (hi[d] - lo[d]) / 255.0;
    }
    s->cell = cell;
    int64_t tot = 1;
    for (int d = 0; d < 3; ++d) {
        s->lo[d] = lo[d];
        const int64_t k = (int64_t)floor((hi[d] - lo[d]) / s->cell) + 1;
        s->nc[d] = k;
        tot *= k;
    }
    free(s->head);
    s->head = (int64_t*)malloc(sizeof(int64_t) * tot);
    if (!s->head) return -1;
    for (int64_t c = 0; c < tot; ++c) s->head[c] = -1;
    for (int64_t i = s->n - 1; i >= 0; --i) {
        if (!s->alive[i]) continue;
        int64_t cx[3];
        const int64_t c = cell_of(s, s->x[i], s->y[i], s->z[i], cx);
        s->next[i] = s->head[c];
        s->head[c] = i;
    }
    return 0;
}

static int cmp_i64(const void* a, const void* b) {
    const int64_t u = *(const int64_t*)a, v = *(const int64_t*)b;
    return (u > v) - (u < v);
}

/* d_CalculateForces for every live particle, plus its overlap gap. */
static int forces(fire_state* s, double D, int law, int nthreads) {
    if (build_cells(s)) return -1;
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        int64_t cap = 256, *cand = (int64_t*)malloc(sizeof(int64_t) * cap);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < s->n; ++i) {
            s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0;
            s->gap[i] = 0.0; s->gapj[i] = -1;
            if (!s->alive[i] || !cand) continue;
            int64_t cx[3], nc = 0;
            cell_of(s, s->x[i], s->y[i], s->z[i], cx);
            for (int64_t dz = -1; dz <= 1; ++dz)
                for (int64_t dy = -1; dy <= 1; ++dy)
                    for (int64_t dx = -1; dx <= 1; ++dx) {
                        const int64_t qx = cx[0] + dx, qy = cx[1] + dy, qz = cx[2] + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx >= s->nc[0] || qy >= s->nc[1] ||
                            qz >= s->nc[2])
                            continue;
                        for (int64_t j = s->head[(qz * s->nc[1] + qy) * s->nc[0] + qx]; j >= 0;
                             j = s->next[j]) {
                            if (j == i) continue;
                            if (nc == cap) {
                                cap *= 2;
                                int64_t* t = (int64_t*)realloc(cand, sizeof(int64_t) * cap);
                                if (!t) { err = 1; break; }
                                cand = t;
                            }
                            cand[nc++] = j;
                        }
                    }
            qsort(cand, (size_t)nc, sizeof(int64_t), cmp_i64);
            double fsx = 0.0, fsy = 0.0, fsz = 0.0, g = 0.0;
            int64_t gj = -1;
            for (int64_t q = 0; q < nc; ++q) {
                const int64_t j = cand[q];
                const double dx = s->x[j] - s->x[i], dy = s->y[j] - s->y[i], dz = s->z[j] - s->z[i];
                const double rr = dx * dx + dy * dy + dz * dz;
                const double a = s->r[i] + s->r[j];
                const double aa = a * a;
                if (rr < aa) {
                    double fm;
                    if (law == RF_FIRE_LAW_HARMONIC) { /* build option, see rockfit_dem.h */
                        const double rho = sqrt(rr);
                        fm = -D * ((a - rho) / rho);
                    } else {
                        const double q1 = aa * (1.0 / rr);
                        fm = -D * q1;
                    }
                    fsx += fm * dx;
                    fsy += fm * dy;
                    fsz += fm * dz;
                    const double gg = (sqrt(rr) - a) / a;
                    if (gg < g) { g = gg; gj = j; }
                }
            }
            s->fx[i] = fsx; s->fy[i] = fsy; s->fz[i] = fsz;
            s->gap[i] = g; s->gapj[i] = gj;
        }
        free(cand);
    }
    return err ? -1 : 0;
}

static double power_sum(const fire_state* s) {
    const int64_t nb = (s->n + FIRE_BLOCK - 1) / FIRE_BLOCK;
    double acc[FIRE_CTRL];
    for (int t = 0; t < FIRE_CTRL; ++t) acc[t] = 0.0;
    for (int64_t b = 0; b < nb; ++b) {
        double v[FIRE_BLOCK];
        for (int t = 0; t < FIRE_BLOCK; ++t) {
            const int64_t i = b * FIRE_BLOCK + t;
            v[t] = 0.0;
            if (i < s->n && s->alive[i])
                v[t] = s->fx[i] * s->vx[i] + s->fy[i] * s->vy[i] + s->fz[i] * s->vz[i];
        }
        for (int off = FIRE_BLOCK / 2; off >= 1; off >>= 1)
            for (int t = 0; t < off; ++t) v[t] = v[t] + v[t + off];
        acc[b % FIRE_CTRL] = acc[b % FIRE_CTRL] + v[0];
    }
    for (int off = FIRE_CTRL / 2; off >= 1; off >>= 1)
        for (int t = 0; t < off; ++t) acc[t] = acc[t] + acc[t + off];
    return acc[0];
}

int ora_fire_relax(int64_t n, double* x, double* y, double* z, const double* r,
                   const rf_fire_cfg* cfg, double* velocity, rf_fire_result* out,
                   int32_t nthreads) {
    if (n <= 0 || !x || !y || !z || !r || !cfg || !out) return -1;
    if (nthreads < 1) nthreads = 1;
    fire_state s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.x = x; s.y = y; s.z = z;
    s.r = (double*)r;
    double* buf = (double*)calloc((size_t)n * 8, sizeof(double));
    s.alive = (unsigned char*)malloc((size_t)n);
    s.next = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    s.gapj = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    if (!buf || !s.alive || !s.next || !s.gapj) return -1;
    s.vx = buf; s.vy = buf + n; s.vz = buf + 2 * n;
    s.fx = buf + 3 * n; s.fy = buf + 4 * n; s.fz = buf + 5 * n;
    s.m = buf + 6 * n; s.gap = buf + 7 * n;
    for (int64_t i = 0; i < n; ++i) {
        s.alive[i] = 1;
        s.m[i] = cfg->density * (4.0 / 3.0) * ORA_PI * r[i] * r[i] * r[i];
        if (velocity) { s.vx[i] = velocity[i]; s.vy[i] = velocity[n + i]; s.vz[i] = velocity[2 * n + i]; }
    }
    const double dt0 = cfg->dt0, dtmax = cfg->dtmax_factor * dt0, dtmin = cfg->dtmin_factor * dt0;
    double dt = dt0, alpha = cfg->alpha0, min_gap = 0.0, P = 0.0;
    int64_t npos = 0, nneg = 0, removed = 0, step = 0;
    const int64_t cap = cfg->hard_cap > 0 ? cfg->hard_cap : 10 * cfg->max_steps;
    const int noinertia = (cfg->quirks & RF_QUIRK_FIRE_NO_INERTIA) != 0;
    int reason = 0, rc = 0;
    const double cx = cfg->center[0], cy = cfg->center[1], cz = cfg->center[2];
    const double R0 = cfg->radius, H0 = cfg->half_height;
    while (1) {
        if (!(step < cfg->max_steps || min_gap < cfg->gap_tol)) { reason = 1; break; }
        if (step >= cap) { reason = 3; break; }
        if (forces(&s, cfg->D, cfg->law, nthreads)) { rc = -1; break; }
        P = power_sum(&s);
        int neg = 0;
        if (P > 0) {
            ++npos;
            nneg = 0;
            if (npos > cfg->ndelay) {
                dt = (dt * cfg->dtgrow < dtmax) ? dt * cfg->dtgrow : dtmax;
                alpha *= cfg->alpha_shrink;
            }
        } else {
            npos = 0;
            ++nneg;
            if (nneg > cfg->np_negative_max && min_gap >= cfg->gap_tol) { reason = 2; break; }
            if (step > cfg->ndelay) {
                dt = (dt * cfg->dtshrink > dtmin) ? dt * cfg->dtshrink : dtmin;
                alpha = cfg->alpha0;
            }
            neg = 1;
        }
        const double hdt = 0.5 * dt;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            if (!s.alive[i]) continue;
            if (neg) {
                x[i] -= s.vx[i] * hdt; y[i] -= s.vy[i] * hdt; z[i] -= s.vz[i] * hdt;
                s.vx[i] = 0.0; s.vy[i] = 0.0; s.vz[i] = 0.0;
            }
            const double Fx = s.fx[i], Fy = s.fy[i], Fz = s.fz[i];
            double vx, vy, vz;
            if (noinertia) {
                const double dtm = dt0 / s.m[i];
                vx = Fx * dtm; vy = Fy * dtm; vz = Fz * dtm;
            } else {
                const double dtm = dt / s.m[i];
                vx = s.vx[i] + Fx * dtm; vy = s.vy[i] + Fy * dtm; vz = s.vz[i] + Fz * dtm;
            }
            const double vv = vx * vx + vy * vy + vz * vz;
            const double f = Fx * Fx + Fy * Fy + Fz * Fz;
            const double vdf = f > 1e-12 ? alpha * sqrt(vv * (1.0 / f)) : alpha * sqrt(vv) * 1e6;
            vx = (1.0 - alpha) * vx + vdf * Fx;
            vy = (1.0 - alpha) * vy + vdf * Fy;
            vz = (1.0 - alpha) * vz + vdf * Fz;
            s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
            x[i] += vx * dt; y[i] += vy * dt; z[i] += vz * dt;
            if (cfg->container == RF_CONTAINER_CYLINDER_Z ||
                cfg->container == RF_CONTAINER_CYLINDER_Z_ABSORB) {
                /* absorbing variant (a build option, no reference counterpart):
                 * remove the outward normal velocity instead of reversing */
                const int absorb = cfg->container == RF_CONTAINER_CYLINDER_Z_ABSORB;
                const double ex = x[i] - cx, ey = y[i] - cy, ez = z[i] - cz;
                double sx = 0.0, sy = 0.0, sz = 0.0;
                const double rr = ex * ex + ey * ey;
                if (ez < cfg->hide_z - cz) {
                    if (rr > R0 * R0) {
                        const double q = 1.0 / sqrt(rr);
                        sx = ex * (R0 * q - 1.0);
                        sy = ey * (R0 * q - 1.0);
                        if (absorb) {
                            const double ux = ex * q, uy = ey * q;
                            const double vr = s.vx[i] * ux + s.vy[i] * uy;
                            if (vr > 0.0) {
                                s.vx[i] -= vr * ux;
                                s.vy[i] -= vr * uy;
                            }
                        } else {
                            s.vx[i] *= -1.0;
                            s.vy[i] *= -1.0;
                        }
                    }
                    if (ez > H0) {
                        sz = H0 - ez;
                        if (!absorb) s.vz[i] *= -1.0;
                        else if (s.vz[i] > 0.0) s.vz[i] = 0.0;
                    }
                    if (ez < -H0) {
                        sz = -H0 - ez;
                        if (!absorb) s.vz[i] *= -1.0;
                        else if (s.vz[i] < 0.0) s.vz[i] = 0.0;
                    }
                    x[i] += sx; y[i] += sy; z[i] += sz;
                }
            }
        }
        if (cfg->check_every > 0 && step % cfg->check_every == 0) {
            double g = 0.0;
            int64_t wi = -1, wj = -1;
            for (int64_t i = 0; i < n; ++i)
                if (s.alive[i] && s.gap[i] < g) { g = s.gap[i]; wi = i; wj = s.gapj[i]; }
            (void)wi;
            min_gap = g;
            if (cfg->delete_after >= 0 && step > cfg->delete_after && min_gap < cfg->gap_tol &&
                wj >= 0) {
                x[wj] = cfg->hidden[0]; y[wj] = cfg->hidden[1]; z[wj] = cfg->hidden[2];
                s.vx[wj] = 0.0; s.vy[wj] = 0.0; s.vz[wj] = 0.0;
                s.fx[wj] = 0.0; s.fy[wj] = 0.0; s.fz[wj] = 0.0;
                s.alive[wj] = 0;
                ++removed;
            }
        }
        ++step;
    }
    if (velocity)
        for (int64_t i = 0; i < n; ++i) {
            velocity[i] = s.vx[i]; velocity[n + i] = s.vy[i]; velocity[2 * n + i] = s.vz[i];
        }
    out->steps = step;
    out->np_positive = npos;
    out->np_negative = nneg;
    out->removed = removed;
    out->dt = dt;
    out->alpha = alpha;
    out->min_gap = min_gap;
    out->power = P;
    out->stop_reason = reason;
    out->pad_ = 0;
    free(buf); free(s.alive); free(s.next); free(s.gapj); free(s.head);
    return rc;
}
