// rf_runtime.hip -- host side of the C-ABI (include/rockfit_dem.h).
//
// Owns all device state of one sample, runs the t = 0 set-up
// (md_calculationuniaxialcompression.cu:155-212), the loading loop in chunks
// of graph-free launches with the controller on device (no per-step host
// sync, unlike md_results.cu:58), the candidate-list rebuilds, and the
// post-processing of the stress-strain curve (md_results.cu:147-187).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rockfit_dem.h"
#include "rf_internal.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace rf {
rf_status report(rf_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}
}  // namespace rf

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;  // MC_pi
rf_status (*const fail)(rf_status, const char*, ...) = rf::report;

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(RF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Per-bond strength: lognormal with the given mean and coefficient of
// variation, keyed by the unordered pair so both directed copies agree.
// Constant (the reference's single m_Ec / m_Gc) when cv == 0.
double bond_strength(uint64_t seed, int64_t lo, int64_t hi, int stream, double mean, double cv) {
    if (!(cv > 0.0)) return mean;
    uint64_t base = seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(lo + 1)) ^
                    (0xC2B2AE3D27D4EB4FULL * (uint64_t)(hi + 1)) ^
                    (0x165667B19E3779F9ULL * (uint64_t)(stream + 1));
    uint64_t h1 = mix64(base);
    uint64_t h2 = mix64(h1 ^ 0xD6E8FEB86659FD93ULL);
    double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    double s2 = std::log(1.0 + cv * cv);
    double s = std::sqrt(s2);
    return mean * std::exp(s * z - 0.5 * s2);
}

struct Buf {
    void* p = nullptr;
};

}  // namespace

struct rf_ctx {
    int64_t n = 0;
    int32_t S = 32, C = 32, B = 32;
    double skin = 0.0;      // current candidate-list skin
    double skin_set = 0.0;  // the skin asked for (rf_options.skin or its default)
    rf_material mat{};
    rf_derived der{};
    rf_box box{};
    rf_options opt{};
    bool own_stream = true;  // created by rf_create (else the caller's, rf_options.stream)
    rf::DevState d{};
    rf::DevParams p{};
    rf::Ctrl* ctrl = nullptr;
    rf::Ctrl hctrl{};
    // pinned staging of the control block's copies to the host (a copy into
    // pageable memory goes through a driver staging buffer), and the two
    // events that bracket an rf_step call; made on first use
    rf::Ctrl* hpin = nullptr;
    hipEvent_t wall_a = nullptr, wall_b = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    bool built = false;
    bool has_test = false;
    rf_platen_cfg cfg{};
    int64_t rebuilds = 0;
    int64_t* dcount = nullptr;
    std::vector<void*> allocs;
    int32_t chunk = 256;
    int32_t timing = 0;  // bracket every timing-th step launch with events (0: off)
    // Full chunks of (step, controller) launch pairs run as captured graphs,
    // replayed while the constants they captured are unchanged.
    bool use_graph = true;
    // sensitivity runs
    bool sens_on = false;
    int32_t sens_np = 0;
    int64_t sens_curve0 = 0;
    rf::SensState sens{};
    // Captured chunks by length: the loop shortens its chunks while list
    // rebuilds come often (a rebuild halts the rest of a chunk, whose
    // launches then do nothing) and lengthens them again after full chunks.
    // A graph bakes in the kernels' by-value arguments: the constants and
    // every device pointer, so it is re-captured when either changes (a
    // reallocated curve buffer, for one).
    struct Graph {
        int64_t steps = 0;
        hipGraphExec_t exec = nullptr;
        rf::DevParams p{};
        rf::DevState d{};
    };
    std::vector<Graph> graphs;
    int32_t chunk_eff = 256;
    std::vector<hipEvent_t> ev;
    double last_kernel_ms = 0.0, last_total_ms = 0.0;
    int64_t last_steps = 0, last_timed = 0;
    double ld = 0.0, destruction = 0.0;
    // Storage order: orig[p] is the caller's id of stored particle p, inv its
    // inverse (RF_ORDER_SPATIAL sorts by space at creation).
    std::vector<int32_t> orig, inv;
    // frozen-branch record of this context (rf_sens_branches)
    uint32_t* br_bits = nullptr;
    double4* br_w = nullptr;
    double* br_vw = nullptr;
    int64_t br_cap = 0, br_recorded = 0;
    bool br_own = false;

    template <class T>
    rf_status alloc(T** out, size_t count) {
        void* ptr = nullptr;
        hipError_t e = hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess)
            return fail(RF_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                        hipGetErrorString(e));
        hipMemset(ptr, 0, std::max<size_t>(count, 1) * sizeof(T));
        allocs.push_back(ptr);
        *out = (T*)ptr;
        return RF_OK;
    }
};

extern "C" {

const char* rf_last_error(void) { return g_err.c_str(); }

int32_t rf_version(void) { return RF_ABI_VERSION; }

// potential_data::CalculateParameters (md_data_types.h:155-187) with the
// smallest particle setting both time scales.
rf_status rf_derive(const rf_material* m, double r_min, double r_max, rf_derived* d) {
    if (!m || !d || !(r_min > 0.0) || r_max < r_min) return fail(RF_ERR_ARG, "rf_derive: bad args");
    d->r_min = r_min;
    d->r_max = r_max;
    d->G_bond = 0.5 * m->E_bond / (1.0 + m->nu);
    d->G_contact = 0.5 * m->E_contact / (1.0 + m->nu);
    d->hm_E = 0.5 * m->E_contact / (1.0 - m->nu * m->nu);
    d->hm_G = 0.5 * d->G_contact / (2.0 - m->nu);
    const double le = std::log(m->restitution);
    d->p_A = le / std::sqrt(kPi * kPi + le * le);
    const double r = r_min;
    d->T_rayleigh = kPi * r * std::sqrt(2.0 * m->density * (1.0 + m->nu)) /
                    (std::sqrt(m->E_contact) * (0.163 * m->nu + 0.8766));
    const double b = m->bond_radius_ratio > 0.0 ? m->bond_radius_ratio * r : m->bond_radius;
    const double bS = kPi * b * b;
    const double pV = 4.0 * (1.0 / 3.0) * kPi * r * r * r;
    const double pm = m->density * pV;
    const double t0 = m->E_bond * bS / (2.0 * r);
    const double t1 = d->G_bond * bS / (2.0 * r);
    d->T_bond = 2.0 * std::sqrt(pm / (t0 > t1 ? t0 : t1));
    if (m->dt > 0.0)
        d->dt = m->dt;
    else
        d->dt = (d->T_rayleigh < d->T_bond) ? 0.05 * d->T_rayleigh : 0.05 * d->T_bond;
    return RF_OK;
}

// sample_data (md_data_types.h:87-124) and cell_data (:42-55) in FP64.
rf_status rf_box_reference(const double space[3], double cell, rf_box* b) {
    if (!space || !b || !(cell > 0.0)) return fail(RF_ERR_ARG, "rf_box_reference: bad args");
    const double A = 1.01 * cell;
    double B[3];
    for (int k = 0; k < 3; ++k) {
        B[k] = A + space[k];
        b->center[k] = 0.5 * (A + B[k]);
        b->half[k] = 0.5 * space[k];
    }
    b->hidden[0] = A + 0.5 * space[0];
    b->hidden[1] = A + 0.5 * space[1];
    b->hidden[2] = (std::floor(B[2] / cell) + 3 - 0.5) * cell;
    b->cell = cell;
    const double inv = 1.0 / cell;
    b->dims[0] = (int32_t)std::floor(B[0] * inv) + 1;
    b->dims[1] = (int32_t)std::floor(B[1] * inv) + 1;
    b->dims[2] = (int32_t)std::floor(B[2] * inv) + 4;
    b->pad_ = 0;
    return RF_OK;
}

}  // extern "C"

namespace {

// Morton code of three 21-bit cell coordinates: bit 3b holds bit b of x,
// bit 3b + 1 of y, bit 3b + 2 of z.
uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t k = 0;
    for (int b = 20; b >= 0; --b)
        k = (k << 3) | ((uint64_t)((z >> b) & 1u) << 2) | ((uint64_t)((y >> b) & 1u) << 1) |
            (uint64_t)((x >> b) & 1u);
    return k;
}

uint32_t quarter_cell(double x, double s) {
    const double q = std::floor(x * s);
    if (!(q > 0.0)) return 0u;
    return q >= 2097151.0 ? 2097151u : (uint32_t)q;
}

// Spatial storage order: particles sorted by the Morton code of their
// quarter-cell coordinates (floor(4 x / cell)), ties by id.
std::vector<int32_t> spatial_order(const rf_sample* s, double cell) {
    const int64_t n = s->n;
    const double sc = 4.0 * (1.0 / cell);
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int64_t i = 0; i < n; ++i)
        key[i] = {morton3(quarter_cell(s->x[i], sc), quarter_cell(s->y[i], sc),
                          quarter_cell(s->z[i], sc)),
                  (int32_t)i};
    std::sort(key.begin(), key.end());
    std::vector<int32_t> ord(n);
    for (int64_t p = 0; p < n; ++p) ord[p] = key[p].second;
    return ord;
}

}  // namespace

extern "C" {

rf_status rf_create(const rf_sample* s, const rf_material* mat, const rf_box* box,
                    const rf_options* opt, rf_ctx** out) {
    if (!s || !mat || !box || !out || s->n <= 0 || !s->x || !s->y || !s->z || !s->r)
        return fail(RF_ERR_ARG, "rf_create: null or empty input");
    *out = nullptr;
    rf_options o{};
    o.slots = 32;
    o.cand_cap = rf::kMaxCand;
    o.skin = -1.0;
    o.device = 0;
    o.block = rf::kBlock;
    if (opt) {
        if (opt->slots > 0) o.slots = opt->slots;
        if (opt->cand_cap > 0) o.cand_cap = opt->cand_cap;
        if (opt->skin >= 0.0) o.skin = opt->skin;
        o.device = opt->device;
        o.stream = opt->stream;
        o.order = opt->order;
    }
    if (o.order != RF_ORDER_SPATIAL && o.order != RF_ORDER_GIVEN)
        return fail(RF_ERR_ARG, "unknown particle order %d", o.order);
    if (s->n >= ((int64_t)1 << 31)) return fail(RF_ERR_ARG, "rf_create: n too large");
    if (o.slots > rf::kMaxSlots)
        return fail(RF_ERR_ARG, "slots=%d exceeds the %d lanes of a particle group", o.slots,
                    rf::kMaxSlots);
    if (o.cand_cap > rf::kMaxCand)
        return fail(RF_ERR_ARG, "cand_cap=%d exceeds %d", o.cand_cap, rf::kMaxCand);
    if (!(mat->acut > 0.0) || !(box->cell >= 2.0 * mat->acut * 0.999999999))
        return fail(RF_ERR_ARG, "cell edge %.6g must be >= 2*acut (acut %.6g)", box->cell,
                    mat->acut);
    if (o.skin < 0.0) o.skin = 0.08 * mat->acut;
    if (mat->acut + o.skin > box->cell)
        return fail(RF_ERR_ARG, "acut + skin must not exceed the cell edge");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(RF_ERR_NODEVICE, "no HIP device available");
    if (o.device < 0 || o.device >= ndev)
        return fail(RF_ERR_ARG, "device %d out of range (%d devices)", o.device, ndev);
    HIP_TRY(hipSetDevice(o.device));

    rf_ctx* c = new rf_ctx();
    c->n = s->n;
    c->S = o.slots;
    c->C = o.cand_cap;
    c->B = rf::kMaxSlots;
    c->mat = *mat;
    c->skin = o.skin;
    c->skin_set = o.skin;
    c->box = *box;
    c->opt = o;
    c->device = o.device;
    double rmin = s->r[0], rmax = s->r[0];
    for (int64_t i = 0; i < s->n; ++i) {
        if (!(s->r[i] > 0.0)) {
            delete c;
            return fail(RF_ERR_ARG, "radius of particle %lld is not positive", (long long)i);
        }
        rmin = std::min(rmin, s->r[i]);
        rmax = std::max(rmax, s->r[i]);
    }
    rf_derive(mat, rmin, rmax, &c->der);
    const int64_t n = c->n;
    auto bail = [&](rf_status st) {
        for (void* ptr : c->allocs) hipFree(ptr);
        if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
        delete c;
        return st;
    };
    if (o.stream) {
        c->stream = (hipStream_t)o.stream;
        c->own_stream = false;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess)
            return bail(fail(RF_ERR_HIP, "hipStreamCreate"));
        c->own_stream = true;
    }
    rf::DevState& d = c->d;
    const int64_t nS = n * c->S, nC = n * c->C, nB = n * c->B;
    rf_status st = RF_OK;
#define ALLOC(ptr, count)                      \
    if ((st = c->alloc(&(ptr), (count))) != RF_OK) return bail(st);
    for (int b = 0; b < 2; ++b) {
        ALLOC(d.pos[b], n);
        ALLOC(d.vel[b], n);
        ALLOC(d.angv[b], n);
    }
    ALLOC(d.pk, n);
    ALLOC(d.F, n);
    ALLOC(d.M, n);
    ALLOC(d.rbuild, n);
    ALLOC(d.nbr, nS);
    ALLOC(d.ot, nS);
    ALLOC(d.occ, n);
    ALLOC(d.bnd, n);
    ALLOC(d.otnz, n);
    ALLOC(d.nbond, n);
    ALLOC(d.bslot, nB);
    ALLOC(d.bpart, nB);
    ALLOC(d.brec, nB);
    ALLOC(d.cand, nC);
    ALLOC(d.cslot, nC);
    ALLOC(d.ncand, n);
    // Working grid: the reference's dims plus one cell layer beyond the top
    // x and y faces.  The reference grid ends inside the last cell that holds
    // the space box (Nr = floor(B/a) + 1), so a particle near B would need a
    // stencil cell outside it; with the extra layer every point of the space
    // box (and the hidden point) lies in cells 1..dims-2.  Cell membership,
    // and so every list order, is unchanged.
    const int32_t gx = box->dims[0] + 1, gy = box->dims[1] + 1, gz = box->dims[2];
    const int64_t ncell = (int64_t)gx * gy * gz;
    ALLOC(d.orig, n);
    ALLOC(d.ckey, n);
    ALLOC(d.ccount, ncell);
    ALLOC(d.cstart, ncell);
    ALLOC(d.cfill, ncell);
    ALLOC(d.ctmp, n);
    ALLOC(d.sorted, n);
    const int64_t nblk = (n + rf::kGroup - 1) / rf::kGroup;
    ALLOC(d.part_b, nblk * rf::kWavesPerGroup);
    ALLOC(d.part_t, nblk * rf::kWavesPerGroup);
    ALLOC(d.part_bonds, nblk * rf::kWavesPerGroup);
    ALLOC(d.pmask, (nblk + 31) / 32);
    ALLOC(c->ctrl, 1);
    ALLOC(c->dcount, 4);
#undef ALLOC

    // particle state: (x, y, z, r), (0, 0, 0, m), (0, 0, 0, 0); per-particle
    // dt/m and dt/I as md_data_types.h:159-163, 185-186.
    // stored particle p is the caller's particle orig[p]
    if (o.order == RF_ORDER_SPATIAL) {
        c->orig = spatial_order(s, box->cell);
    } else {
        c->orig.resize(n);
        for (int64_t i = 0; i < n; ++i) c->orig[i] = (int32_t)i;
    }
    c->inv.resize(n);
    for (int64_t q = 0; q < n; ++q) c->inv[c->orig[q]] = (int32_t)q;
    std::vector<double4> hp(n), hv(n), hw(n);
    std::vector<double2> hk(n);
    for (int64_t q = 0; q < n; ++q) {
        const int64_t i = c->orig[q];
        const double r = s->r[i];
        const double pV = 4.0 * (1.0 / 3.0) * kPi * r * r * r;
        const double pm = mat->density * pV;
        const double pI = 0.4 * pm * r * r;
        hp[q] = make_double4(s->x[i], s->y[i], s->z[i], r);
        hv[q] = make_double4(0.0, 0.0, 0.0, pm);
        hw[q] = make_double4(0.0, 0.0, 0.0, 0.0);
        hk[q] = make_double2((1.0 / pm) * c->der.dt, c->der.dt / pI);
    }
    if (hipMemcpy(d.orig, c->orig.data(), n * sizeof(int32_t), hipMemcpyHostToDevice))
        return bail(fail(RF_ERR_HIP, "upload of particle ids failed"));
    for (int b = 0; b < 2; ++b) {
        if (hipMemcpy(d.pos[b], hp.data(), n * sizeof(double4), hipMemcpyHostToDevice) ||
            hipMemcpy(d.vel[b], hv.data(), n * sizeof(double4), hipMemcpyHostToDevice) ||
            hipMemcpy(d.angv[b], hw.data(), n * sizeof(double4), hipMemcpyHostToDevice))
            return bail(fail(RF_ERR_HIP, "upload of particle state failed"));
    }
    if (hipMemcpy(d.pk, hk.data(), n * sizeof(double2), hipMemcpyHostToDevice))
        return bail(fail(RF_ERR_HIP, "upload of particle constants failed"));

    rf::DevParams& p = c->p;
    p.n = n;
    p.S = c->S;
    p.C = c->C;
    p.B = c->B;
    p.nblk = (int32_t)nblk;
    p.dt = c->der.dt;
    p.E = mat->E_bond;
    p.G = c->der.G_bond;
    p.hmE = c->der.hm_E;
    p.hmG = c->der.hm_G;
    p.pA = c->der.p_A;
    p.mu = mat->friction;
    p.mur = mat->rolling_friction;
    p.acut2 = mat->acut * mat->acut;
    p.cell = box->cell;
    p.inv_cell = 1.0 / box->cell;
    p.cand_r2 = (mat->acut + o.skin) * (mat->acut + o.skin);
    p.skin_half2 = 0.25 * o.skin * o.skin;
    p.b_ratio = mat->bond_radius_ratio;
    p.b_abs = mat->bond_radius;
    p.bond_gap = mat->bond_gap;
    p.quirks = mat->quirks;
    p.nx = gx;
    p.ny = gy;
    p.nz = gz;
    p.ncell = ncell;
    for (int k = 0; k < 3; ++k) {
        p.cen[k] = box->center[k];
        p.half[k] = box->half[k];
        p.hid[k] = box->hidden[k];
    }
    p.kind = RF_TEST_NONE;
    p.axis = 2;
    p.dnsave = 1;
    p.store_fm = 1;
    p.curve_cap = 0;
    p.zero_idx = c->inv[0];
    std::memset(&c->hctrl, 0, sizeof(c->hctrl));
    if (hipMemcpy(c->ctrl, &c->hctrl, sizeof(rf::Ctrl), hipMemcpyHostToDevice))
        return bail(fail(RF_ERR_HIP, "upload of control block failed"));
    *out = c;
    return RF_OK;
}

rf_status rf_destroy(rf_ctx* c) {
    if (!c) return RF_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev) hipEventDestroy(e);
    if (c->wall_a) hipEventDestroy(c->wall_a);
    if (c->wall_b) hipEventDestroy(c->wall_b);
    if (c->hpin) hipHostFree(c->hpin);
    for (auto& g : c->graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    for (void* ptr : c->allocs) hipFree(ptr);
    if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return RF_OK;
}

int64_t rf_num_particles(const rf_ctx* c) { return c ? c->n : 0; }
int32_t rf_num_slots(const rf_ctx* c) { return c ? c->S : 0; }

}  // extern "C"

namespace {

rf_status ensure_pinned(rf_ctx* c) {
    if (!c->hpin) HIP_TRY(hipHostMalloc((void**)&c->hpin, sizeof(rf::Ctrl), hipHostMallocDefault));
    return RF_OK;
}

rf_status sync_ctrl(rf_ctx* c) {
    rf_status st = ensure_pinned(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(c->hpin, c->ctrl, sizeof(rf::Ctrl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->hctrl = *c->hpin;
    return RF_OK;
}

rf_status device_error(rf_ctx* c) {
    const int e = c->hctrl.err;
    if (e & rf::kErrGrid) return fail(RF_ERR_GRID, "a particle left the padded cell grid");
    if (e & rf::kErrRowOverflow)
        return fail(RF_ERR_OVERFLOW, "interaction row overflow (more than %d partners)", c->S);
    if (e & rf::kErrCandOverflow)
        return fail(RF_ERR_OVERFLOW, "candidate list overflow (more than %d within acut+skin)",
                    c->C);
    return fail(RF_ERR_STATE, "device error flags 0x%x", e);
}

// Ends a capture begun on `st` whatever happened inside it, instantiates the
// graph and uploads it to the device (otherwise its first launch uploads it:
// about 100 us for a 40-node graph, all of it inside the first rf_step that
// replays it); `launch_err` is the first error a launch inside it returned.
rf_status end_capture(hipStream_t st, hipError_t launch_err, hipGraphExec_t* exec) {
    *exec = nullptr;
    hipGraph_t gr = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &gr);
    hipError_t e = launch_err != hipSuccess ? launch_err : ee;
    if (e == hipSuccess) e = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
    if (gr) hipGraphDestroy(gr);
    if (e == hipSuccess) {
        e = hipGraphUpload(*exec, st);
        if (e != hipSuccess) hipGraphExecDestroy(*exec);
    }
    if (e != hipSuccess) {
        *exec = nullptr;
        return fail(RF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    return RF_OK;
}

// Launch `k` (step, controller) pairs as one captured graph.  The kernels
// read the step's buffers and halt flag from the device control block, so the
// same graph serves every chunk of its length; it is captured again when the
// arguments passed by value (DevParams, and the DevState pointers) change.
rf_status ensure_graph(rf_ctx* c, int64_t k, rf_ctx::Graph** out) {
    rf_ctx::Graph* g = nullptr;
    for (auto& q : c->graphs)
        if (q.steps == k) g = &q;
    if (!g) {
        c->graphs.push_back(rf_ctx::Graph{});
        g = &c->graphs.back();
        g->steps = k;
    }
    if (!g->exec || std::memcmp(&g->p, &c->p, sizeof(rf::DevParams)) != 0 ||
        std::memcmp(&g->d, &c->d, sizeof(rf::DevState)) != 0) {
        if (g->exec) {
            HIP_TRY(hipGraphExecDestroy(g->exec));
            g->exec = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int64_t s = 0; s < k; ++s) {
            const hipError_t e1 = rf::launch_step(c->d, c->p, c->ctrl, c->stream);
            const hipError_t e2 = rf::launch_controller(c->d, c->p, c->ctrl, c->stream);
            if (le == hipSuccess) le = e1 != hipSuccess ? e1 : e2;
        }
        rf_status st = end_capture(c->stream, le, &g->exec);
        if (st) return st;
        g->p = c->p;
        g->d = c->d;
    }
    *out = g;
    return RF_OK;
}

// A chunk of k steps captured once as a graph (per chunk length, re-captured
// when the constants it baked in change) and replayed.
rf_status graph_chunk(rf_ctx* c, int64_t k) {
    rf_ctx::Graph* g = nullptr;
    rf_status st = ensure_graph(c, k, &g);
    if (st) return st;
    HIP_TRY(hipGraphLaunch(g->exec, c->stream));
    return RF_OK;
}

// k loading steps of one sample on its stream, without waiting: a captured
// graph (one per chunk length) or, for a single step, the two launches.
rf_status launch_chunk(rf_ctx* c, int64_t k) {
    if (c->use_graph && k > 1) return graph_chunk(c, k);
    for (int64_t s = 0; s < k; ++s) {
        HIP_TRY(rf::launch_step(c->d, c->p, c->ctrl, c->stream));
        HIP_TRY(rf::launch_controller(c->d, c->p, c->ctrl, c->stream));
    }
    return RF_OK;
}

// Next chunk length: after a chunk cut short by a list rebuild, the power of
// two at or above the steps it ran (at least 8); after a full chunk, twice
// the length, up to the set chunk.
void adapt_chunk(rf_ctx* c, int64_t k, int64_t executed, bool rebuilt) {
    if (rebuilt && executed < k) {
        int64_t q = 8;
        while (q < executed) q <<= 1;
        c->chunk_eff = (int32_t)std::min<int64_t>(q, c->chunk);
    } else if (executed == k && k == c->chunk_eff) {
        c->chunk_eff = (int32_t)std::min<int64_t>(2 * (int64_t)c->chunk_eff, c->chunk);
    }
}

void set_skin(rf_ctx* c, double skin) {
    c->skin = skin;
    c->p.cand_r2 = (c->mat.acut + c->skin) * (c->mat.acut + c->skin);
    c->p.skin_half2 = 0.25 * c->skin * c->skin;
}

// Cell sort of the current positions and fresh candidate lists.  A list
// longer than the candidate capacity halves the skin and builds again (the
// skin only decides how often lists are rebuilt, never a result); below
// acut/64 the overflow is an error.  A narrowed skin is doubled back toward
// its set value at the next rebuild, so one crowded moment does not leave
// the lists rebuilding every few steps for the rest of the run.
rf_status rebuild(rf_ctx* c) {
    if (c->skin < c->skin_set) set_skin(c, std::min(c->skin_set, 2.0 * c->skin));
    for (;;) {
        HIP_TRY(rf::launch_cell_keys(c->d, c->p, c->ctrl, c->stream));
        HIP_TRY(rf::launch_cell_sort(c->d, c->p, c->ctrl, c->stream));
        HIP_TRY(rf::launch_build_candidates(c->d, c->p, c->ctrl, c->stream));
        HIP_TRY(rf::launch_clear_flags(c->ctrl, c->stream));
        c->rebuilds++;
        rf_status st = sync_ctrl(c);
        if (st) return st;
        if (c->hctrl.err != rf::kErrCandOverflow || c->skin < c->mat.acut / 64.0) return RF_OK;
        set_skin(c, 0.5 * c->skin);
        const int32_t zero = 0;
        HIP_TRY(hipMemcpyAsync(&c->ctrl->err, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(rf::launch_clear_flags(c->ctrl, c->stream));
    }
}

}  // namespace

extern "C" {

rf_status rf_build_bonds(rf_ctx* c) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    rf::DevState& d = c->d;
    HIP_TRY(rf::launch_cell_keys(d, c->p, c->ctrl, c->stream));
    HIP_TRY(rf::launch_cell_sort(d, c->p, c->ctrl, c->stream));
    HIP_TRY(rf::launch_construct_rows(d, c->p, c->ctrl, c->stream));
    HIP_TRY(rf::launch_construct_bonds(d, c->p, c->ctrl, c->stream));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    if (c->hctrl.err) return device_error(c);
    // per-bond strengths on the host (same draw as every other host tool)
    const int64_t n = c->n, B = c->B;
    std::vector<int32_t> nb(n), part(n * B);
    HIP_TRY(hipMemcpy(nb.data(), d.nbond, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(part.data(), d.bpart, n * B * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> sc(n * B, 0.0), tc(n * B, 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < nb[i]; ++b) {
            const int64_t oi = c->orig[i], oj = c->orig[part[i * B + b]];
            const int64_t lo = std::min(oi, oj), hi = std::max(oi, oj);
            sc[i * B + b] = bond_strength(c->mat.strength_seed, lo, hi, 0, c->mat.sigma_c,
                                          c->mat.sigma_c_cv);
            tc[i * B + b] = bond_strength(c->mat.strength_seed, lo, hi, 1, c->mat.tau_c,
                                          c->mat.tau_c_cv);
        }
    {
        double* dsc = nullptr;
        double* dtc = nullptr;
        HIP_TRY(hipMalloc(&dsc, n * B * sizeof(double)));
        HIP_TRY(hipMalloc(&dtc, n * B * sizeof(double)));
        HIP_TRY(hipMemcpy(dsc, sc.data(), n * B * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dtc, tc.data(), n * B * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(rf::launch_set_strengths(d.brec, dsc, dtc, n * B, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipFree(dsc));
        HIP_TRY(hipFree(dtc));
    }
    HIP_TRY(rf::launch_prune_initial(d, c->p, c->ctrl, c->stream));
    st = rebuild(c);
    if (st) return st;
    st = sync_ctrl(c);
    if (st) return st;
    if (c->hctrl.err) return device_error(c);
    c->built = true;
    return RF_OK;
}

// compression_data::CalculateUParameters / CalculateBParameters
// (md_data_types.h:212-242, 380-409) and ResultsUInit (md_results.cu:16-29).
rf_status rf_set_test(rf_ctx* c, const rf_platen_cfg* cfg) {
    if (!c || !cfg) return fail(RF_ERR_ARG, "null argument");
    if (cfg->kind < RF_TEST_NONE || cfg->kind > RF_TEST_BRAZILIAN)
        return fail(RF_ERR_ARG, "unknown test kind %d", cfg->kind);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->cfg = *cfg;
    c->has_test = true;
    rf::DevParams& p = c->p;
    rf::Ctrl& h = c->hctrl;
    const rf::Ctrl old = h;
    std::memset(&h, 0, sizeof(h));
    h.cur = old.cur;
    p.kind = cfg->kind;
    p.dnsave = cfg->dnsave > 0 ? cfg->dnsave : 1;
    p.max_steps = cfg->max_steps;
    if (cfg->kind != RF_TEST_NONE) {
        const int ax = cfg->kind == RF_TEST_BRAZILIAN ? 1 : 2;
        p.axis = ax;
        const double sz = cfg->size[ax];
        h.size_axis = sz;
        h.size0 = sz;
        h.bottom = cfg->center[ax] - 0.5 * sz;
        h.top = cfg->center[ax] + 0.5 * sz;
        for (int k = 0; k < 3; ++k) p.scope_c[k] = cfg->center[k];
        const double hx = 0.5 * cfg->size[0];
        double area;
        if (cfg->kind == RF_TEST_BRAZILIAN) {
            const double hz = 0.5 * cfg->size[2];
            const double Rs = (hx > hz) ? hx : 1.05 * hz;
            p.rr_scope = Rs * Rs;
            area = 0.5 * 2 * kPi * hx * cfg->size[2];
            p.zero_stress = 0.5 * cfg->min_stress;
        } else {
            const double Rs = 1.5 * hx;
            p.rr_scope = Rs * Rs;
            area = kPi * (hx + cfg->r_ref) * (hx + cfg->r_ref);
            p.zero_stress = 0.33 * cfg->min_stress;
        }
        p.inv_area = 1.0 / area;
        p.zcut = c->box.center[2] + c->box.half[2];
        p.Ew = cfg->wall_modulus_factor * c->der.hm_E;
        p.Gw = cfg->wall_modulus_factor * c->der.hm_G;
        p.v_load = cfg->v_load;
        p.min_stress = cfg->min_stress;
        p.max_strain = cfg->max_strain;
        p.cond_fracture = cfg->cond_fracture;
        int64_t cap = cfg->curve_cap;
        if (cap <= 0) {
            const int64_t dn = p.dnsave;
            const double est = 1.5 * cfg->max_strain * sz / (std::fabs(cfg->v_load) * c->der.dt * dn);
            cap = (int64_t)est + 16;
            if (cfg->max_steps > 0 && cfg->max_steps / dn + 16 < cap) cap = cfg->max_steps / dn + 16;
        }
        p.curve_cap = cap;
        if (c->d.curve) {
            hipFree(c->d.curve);
            hipFree(c->d.bcount);
            c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)c->d.curve),
                            c->allocs.end());
            c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)c->d.bcount),
                            c->allocs.end());
            c->d.curve = nullptr;
            c->d.bcount = nullptr;
        }
        rf_status st = c->alloc(&c->d.curve, 3 * (cap + 2));
        if (st) return st;
        st = c->alloc(&c->d.bcount, cap + 2);
        if (st) return st;
        if (!(cfg->max_strain > 0.0)) h.done = 1;
    } else {
        p.curve_cap = 0;
    }
    h.halt = h.done ? 1 : 0;
    HIP_TRY(hipMemcpy(c->ctrl, &h, sizeof(rf::Ctrl), hipMemcpyHostToDevice));
    c->ld = 0.0;
    c->destruction = 0.0;
    return RF_OK;
}

rf_status rf_set_timing(rf_ctx* c, int32_t enable) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    c->timing = enable != 0;
    return RF_OK;
}

rf_status rf_set_graph(rf_ctx* c, int32_t enable) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    c->use_graph = enable != 0;
    return RF_OK;
}

rf_status rf_set_chunk(rf_ctx* c, int32_t steps) {
    if (!c || steps <= 0) return fail(RF_ERR_ARG, "bad chunk");
    c->chunk = steps;
    c->chunk_eff = steps;
    return RF_OK;
}

rf_status rf_capture_chunk(rf_ctx* c, int32_t steps) {
    if (!c || steps <= 0) return fail(RF_ERR_ARG, "rf_capture_chunk: bad arguments");
    if (!c->built) return fail(RF_ERR_STATE, "rf_capture_chunk before rf_build_bonds");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->chunk = steps;
    c->chunk_eff = steps;
    if (!c->use_graph || steps < 2) return RF_OK;
    rf_ctx::Graph* g = nullptr;
    return ensure_graph(c, c->timing ? steps - 1 : steps, &g);
}

rf_status rf_step(rf_ctx* c, int64_t nsteps, int64_t* done_steps) {
    if (done_steps) *done_steps = 0;
    if (!c) return fail(RF_ERR_ARG, "null context");
    if (!c->built) return fail(RF_ERR_STATE, "rf_step before rf_build_bonds");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    const int64_t start = c->hctrl.step;
    int64_t remaining = nsteps;
    if (!c->wall_a) HIP_TRY(hipEventCreate(&c->wall_a));
    if (!c->wall_b) HIP_TRY(hipEventCreate(&c->wall_b));
    HIP_TRY(hipEventRecord(c->wall_a, c->stream));
    double kernel_ms = 0.0;
    int64_t timed = 0;
    while (remaining > 0 && !c->hctrl.done) {
        const int64_t k = std::min<int64_t>(remaining, c->chunk_eff);
        if (c->timing && (int64_t)c->ev.size() < 2 * k) {
            while ((int64_t)c->ev.size() < 2 * k) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                c->ev.push_back(e);
            }
        }
        const int64_t before = c->hctrl.step;
        // A full chunk runs as a captured graph; with timing on, its last step
        // is launched on its own with two events that take the dispatch's own
        // timestamps (one sample per chunk).
        const bool graphed = c->use_graph && k == c->chunk_eff && k > 1;
        if (graphed) {
            st = graph_chunk(c, c->timing ? k - 1 : k);
            if (st) return st;
            if (c->timing) {
                HIP_TRY(rf::launch_step_timed(c->d, c->p, c->ctrl, c->stream, c->ev[0], c->ev[1]));
                HIP_TRY(rf::launch_controller(c->d, c->p, c->ctrl, c->stream));
            }
        } else {
            for (int64_t s = 0; s < k; ++s) {
                const bool tm = c->timing && (before + s) % c->timing == 0;
                if (tm)
                    HIP_TRY(rf::launch_step_timed(c->d, c->p, c->ctrl, c->stream, c->ev[2 * s],
                                                  c->ev[2 * s + 1]));
                else
                    HIP_TRY(rf::launch_step(c->d, c->p, c->ctrl, c->stream));
                HIP_TRY(rf::launch_controller(c->d, c->p, c->ctrl, c->stream));
            }
        }
        st = sync_ctrl(c);
        if (st) return st;
        const int64_t executed = c->hctrl.step - before;
        if (c->timing && graphed) {
            if (executed == k) {  // the sampled last step ran
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                kernel_ms += ms;
                ++timed;
            }
        } else if (c->timing) {
            for (int64_t s = 0; s < executed; ++s) {
                if ((before + s) % c->timing) continue;
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev[2 * s], c->ev[2 * s + 1]));
                kernel_ms += ms;
                ++timed;
            }
        }
        remaining -= executed;
        if (c->hctrl.err) return device_error(c);
        const bool rebuilt = c->hctrl.rebuild != 0;
        adapt_chunk(c, k, executed, rebuilt);
        if (rebuilt) {
            st = rebuild(c);
            if (st) return st;
            if (c->hctrl.err) return device_error(c);
        } else if (executed == 0 && !c->hctrl.done) {
            return fail(RF_ERR_STATE, "loading loop halted without a reason");
        }
    }
    HIP_TRY(hipEventRecord(c->wall_b, c->stream));
    HIP_TRY(hipEventSynchronize(c->wall_b));
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, c->wall_a, c->wall_b));
    c->last_total_ms = total;
    c->last_kernel_ms = kernel_ms;
    c->last_timed = timed;
    c->last_steps = c->hctrl.step - start;
    if (done_steps) *done_steps = c->last_steps;
    return RF_OK;
}

rf_status rf_last_timing(rf_ctx* c, double* kms, int64_t* timed, double* tms, int64_t* steps) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    if (kms) *kms = c->last_kernel_ms;
    if (timed) *timed = c->last_timed;
    if (tms) *tms = c->last_total_ms;
    if (steps) *steps = c->last_steps;
    return RF_OK;
}

}  // extern "C"

namespace {

// CalculateLinearityDeviation (md_results.cu:147-165) over n stored samples
// (later ones read as zero); the loop also stops at `count` samples
// (stepsave).
double linearity_deviation(const double* h, int64_t n, int64_t count, double peak_strain,
                           double peak_stress) {
    auto at = [&](int64_t i, int c) { return i < n ? h[3 * i + c] : 0.0; };
    double l_eps = 0, l_deps = 0, l_A = peak_stress / peak_strain, l_dS = 0, ld = 0;
    int64_t l_i = 0;
    do {
        ld += l_deps * l_dS * l_dS;
        ++l_i;
        l_eps = -at(l_i, 0);
        l_deps = -0.5 * (at(l_i + 1, 0) - at(l_i - 1, 0));
        l_dS = (l_A * l_eps - 0.5 * (at(l_i, 2) - at(l_i, 1)));
    } while (l_eps < peak_strain - 2 * l_deps && l_i < count);
    ld /= (peak_strain * peak_stress * peak_stress);
    ld += 1.0;
    return ld;
}

// CalculateLinearityDeviation and CalculateDestruction (md_results.cu:147-187)
// on the downloaded curve; zero-filled past the last sample.
rf_status post_process(rf_ctx* c) {
    const int64_t cap = c->p.curve_cap;
    if (cap <= 0) return RF_OK;
    std::vector<double> h(3 * (cap + 2));
    std::vector<int64_t> bc(cap + 2);
    HIP_TRY(hipMemcpy(h.data(), c->d.curve, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bc.data(), c->d.bcount, bc.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    const rf::Ctrl& t = c->hctrl;
    if (c->cfg.kind == RF_TEST_UNIAXIAL)
        c->ld = linearity_deviation(h.data(), std::min<int64_t>(t.stepsave, cap), t.stepsave,
                                    t.peak_strain, t.peak_stress);
    if (t.stepsave >= 1 && bc[0] > 0) c->destruction = (double)bc[t.stepsave - 1] / (double)bc[0];
    return RF_OK;
}

void fill_state(const rf_ctx* c, rf_test_state* o) {
    const rf::Ctrl& t = c->hctrl;
    std::memset(o, 0, sizeof(*o));
    o->top = t.top; o->bottom = t.bottom; o->v = t.vw; o->size_axis = t.size_axis;
    o->size0 = t.size0; o->strain = t.strain; o->stress = t.stress;
    o->peak_stress = t.peak_stress; o->peak_strain = t.peak_strain;
    o->sfl0 = t.sfl0; o->sfl1 = t.sfl1;
    o->linearity_deviation = c->ld; o->destruction = c->destruction;
    o->step = t.step; o->steps_stop = t.steps_stop; o->n_curve = t.stepsave;
    o->loading = t.loading; o->fracture = t.fracture; o->done = t.done;
    o->kind = c->has_test ? c->cfg.kind : 0;
}

}  // namespace

extern "C" {

rf_status rf_linearity_deviation(const double* curve, int64_t n, double peak_strain,
                                 double peak_stress, double* out) {
    if (!out || n < 0 || (n > 0 && !curve)) return fail(RF_ERR_ARG, "rf_linearity_deviation: bad arguments");
    if (!(peak_strain != 0.0) || !(peak_stress != 0.0))
        return fail(RF_ERR_ARG, "rf_linearity_deviation: zero peak strain or stress");
    *out = linearity_deviation(curve, n, n, peak_strain, peak_stress);
    return RF_OK;
}

rf_status rf_curve_moduli(const double* curve, int64_t n, double peak_strain, double peak_stress,
                          double lo, double hi, rf_moduli* out) {
    if (!out || n < 0 || (n > 0 && !curve) || !(lo <= hi))
        return fail(RF_ERR_ARG, "rf_curve_moduli: bad arguments");
    *out = rf_moduli{};
    if (!(peak_strain != 0.0)) return fail(RF_ERR_ARG, "rf_curve_moduli: zero peak strain");
    out->secant = peak_stress / peak_strain;
    // ascending branch: up to the first sample at the peak stress
    int64_t end = n;
    for (int64_t i = 0; i < n; ++i)
        if (0.5 * (curve[3 * i + 2] - curve[3 * i + 1]) >= peak_stress) {
            end = i + 1;
            break;
        }
    double sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0;
    int64_t m = 0;
    for (int64_t i = 0; i < end; ++i) {
        const double e = -curve[3 * i], s = 0.5 * (curve[3 * i + 2] - curve[3 * i + 1]);
        if (s < lo * peak_stress || s > hi * peak_stress) continue;
        sx += e; sy += s; sxx += e * e; sxy += e * s; syy += s * s;
        ++m;
    }
    out->points = m;
    if (m < 2) return RF_OK;
    const double cxx = sxx - sx * sx / m, cxy = sxy - sx * sy / m, cyy = syy - sy * sy / m;
    if (!(cxx > 0.0)) return RF_OK;
    out->lsq = cxy / cxx;
    out->lsq_intercept = (sy - out->lsq * sx) / m;
    out->r2 = cyy > 0.0 ? cxy * cxy / (cxx * cyy) : 1.0;
    return RF_OK;
}

rf_status rf_run(rf_ctx* c, rf_test_state* out) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    if (!c->has_test || c->cfg.kind == RF_TEST_NONE)
        return fail(RF_ERR_STATE, "rf_run needs a platen test (rf_set_test)");
    int64_t done = 0;
    do {
        rf_status st = rf_step(c, INT64_MAX / 4, &done);
        if (st) return st;
    } while (!c->hctrl.done && done > 0);
    rf_status st = post_process(c);
    if (st) return st;
    if (out) fill_state(c, out);
    return RF_OK;
}

static rf_status run_test(rf_ctx* c, const rf_platen_cfg* cfg, int32_t kind, const char* who,
                          rf_result* out) {
    if (!c || !cfg || !out) return fail(RF_ERR_ARG, "%s: null argument", who);
    if (cfg->kind != kind) return fail(RF_ERR_ARG, "%s: cfg->kind is %d, not %d", who, cfg->kind, kind);
    rf_status st = rf_set_test(c, cfg);
    if (st) return st;
    rf_test_state ts;
    st = rf_run(c, &ts);
    if (st) return st;
    *out = rf_result{};
    out->peak_stress = ts.peak_stress;
    out->peak_strain = ts.peak_strain;
    out->secant_modulus = ts.peak_strain != 0.0 ? ts.peak_stress / ts.peak_strain : 0.0;
    out->linearity_deviation = ts.linearity_deviation;
    out->destruction = ts.destruction;
    out->steps = ts.step;
    out->curve_points = ts.n_curve;
    out->fracture = ts.fracture;
    out->kind = ts.kind;
    return RF_OK;
}

rf_status rf_run_uniaxial(rf_ctx* c, const rf_platen_cfg* cfg, rf_result* out) {
    return run_test(c, cfg, RF_TEST_UNIAXIAL, "rf_run_uniaxial", out);
}

rf_status rf_run_brazilian(rf_ctx* c, const rf_platen_cfg* cfg, rf_result* out) {
    return run_test(c, cfg, RF_TEST_BRAZILIAN, "rf_run_brazilian", out);
}

rf_status rf_get_test_state(rf_ctx* c, rf_test_state* out) {
    if (!c || !out) return fail(RF_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    if (c->hctrl.done) {
        st = post_process(c);
        if (st) return st;
    }
    fill_state(c, out);
    return RF_OK;
}

rf_status rf_get_curve(rf_ctx* c, double* out, int64_t cap, int64_t* n) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    int64_t m = std::min<int64_t>(c->hctrl.stepsave, c->p.curve_cap);
    if (m > cap) m = cap;
    if (out && m > 0)
        HIP_TRY(hipMemcpy(out, c->d.curve, 3 * m * sizeof(double), hipMemcpyDeviceToHost));
    if (n) *n = m;
    return RF_OK;
}

rf_status rf_get_bond_counts(rf_ctx* c, int64_t* out, int64_t cap, int64_t* n) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    int64_t m = std::min<int64_t>(c->hctrl.stepsave, c->p.curve_cap);
    if (m > cap) m = cap;
    if (out && m > 0)
        HIP_TRY(hipMemcpy(out, c->d.bcount, m * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (n) *n = m;
    return RF_OK;
}

rf_status rf_get_counters(rf_ctx* c, rf_counters* out) {
    if (!c || !out) return fail(RF_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemsetAsync(c->dcount, 0, 4 * sizeof(int64_t), c->stream));
    HIP_TRY(rf::launch_counters(c->d, c->p, c->ctrl, c->dcount, c->stream));
    int64_t h[4];
    HIP_TRY(hipMemcpyAsync(h, c->dcount, sizeof h, hipMemcpyDeviceToHost, c->stream));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    std::memset(out, 0, sizeof(*out));
    out->n = c->n;
    out->occupied = h[0];
    out->bonds = h[1];
    out->contacts = h[2];
    out->candidates = h[3];
    out->steps = c->hctrl.step;
    out->rebuilds = c->rebuilds;
    return RF_OK;
}

rf_status rf_get_state(rf_ctx* c, rf_state_view* v) {
    if (!c || !v) return fail(RF_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    const int64_t n = c->n, S = c->S, B = c->B;
    const rf::DevState& d = c->d;
    const int cur = c->hctrl.cur;
    std::vector<double4> hp(n), hv(n), hw(n), hf(n), hm(n);
    HIP_TRY(hipMemcpy(hp.data(), d.pos[cur], n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hv.data(), d.vel[cur], n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hw.data(), d.angv[cur], n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hf.data(), d.F, n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hm.data(), d.M, n * sizeof(double4), hipMemcpyDeviceToHost));
    const std::vector<int32_t>& og = c->orig;  // caller id of stored particle
    auto split = [n, &og](double* dst, const std::vector<double4>& src) {
        if (!dst) return;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t o = og[i];
            dst[o] = src[i].x;
            dst[n + o] = src[i].y;
            dst[2 * n + o] = src[i].z;
        }
    };
    split(v->R, hp); split(v->V, hv); split(v->W, hw); split(v->F, hf); split(v->M, hm);
    std::vector<int32_t> nbr(n * S), nbond(n), bpart(n * B);
    std::vector<uint32_t> occ(n), bnd(n), otnz(n);
    std::vector<double4> ot(n * S);
    std::vector<int8_t> bslot(n * B);
    HIP_TRY(hipMemcpy(nbr.data(), d.nbr, n * S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(occ.data(), d.occ, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bnd.data(), d.bnd, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(otnz.data(), d.otnz, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ot.data(), d.ot, n * S * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nbond.data(), d.nbond, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bslot.data(), d.bslot, n * B, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bpart.data(), d.bpart, n * B * 4, hipMemcpyDeviceToHost));
    std::vector<rf::BondRec> brec(n * B);
    HIP_TRY(hipMemcpy(brec.data(), d.brec, n * B * sizeof(rf::BondRec), hipMemcpyDeviceToHost));
    // The device prunes a slot at the start of the next step; the reference
    // prunes it at the end of this one (d_DeleteFarLinks).  Report the state
    // after that end-of-step prune, with the kernel's predicate.
    const double acut2 = c->p.acut2;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t k = 0; k < S; ++k) {
            if (!((occ[i] >> k) & 1u)) continue;
            const int64_t j = nbr[i * S + k];
            const double dx = hp[j].x - hp[i].x, dy = hp[j].y - hp[i].y, dz = hp[j].z - hp[i].z;
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr > acut2 || rr < 1e-10) {
                occ[i] &= ~(1u << k);
                bnd[i] &= ~(1u << k);
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = og[i];
        for (int64_t k = 0; k < S; ++k) {
            const int64_t qs = i * S + k;  // stored slot
            const int64_t q = o * S + k;   // the caller's slot
            const bool oc = (occ[i] >> k) & 1u;
            const bool bd = oc && ((bnd[i] >> k) & 1u);
            if (v->nbr) v->nbr[q] = oc ? og[nbr[qs]] : -1;
            if (v->type) v->type[q] = bd ? 1 : 0;
            if (v->bidx) v->bidx[q] = -1;
            const bool onz = oc && !bd && ((otnz[i] >> k) & 1u);
            if (v->o_t) {
                v->o_t[3 * q] = onz ? ot[qs].x : 0.0;
                v->o_t[3 * q + 1] = onz ? ot[qs].y : 0.0;
                v->o_t[3 * q + 2] = onz ? ot[qs].z : 0.0;
            }
            if (v->inv_L0) v->inv_L0[q] = 0.0;
            if (v->sigma_c) v->sigma_c[q] = 0.0;
            if (v->tau_c) v->tau_c[q] = 0.0;
            for (int cc = 0; cc < 3; ++cc) {
                if (v->prev_n) v->prev_n[3 * q + cc] = 0.0;
                if (v->m_n) v->m_n[3 * q + cc] = 0.0;
                if (v->m_t) v->m_t[3 * q + cc] = 0.0;
            }
        }
        for (int b = 0; b < nbond[i]; ++b) {
            const int64_t qb = i * B + b;
            const int ks = bslot[qb];
            if (ks < 0 || !((bnd[i] >> ks) & 1u) || !((occ[i] >> ks) & 1u)) continue;
            const int64_t q = o * S + ks;
            const rf::BondRec& r = brec[qb];
            if (v->bidx) v->bidx[q] = (int8_t)b;
            if (v->inv_L0) v->inv_L0[q] = r.invL0;
            if (v->sigma_c) v->sigma_c[q] = r.sc;
            if (v->tau_c) v->tau_c[q] = r.tc;
            const double pn[3] = {r.pnx, r.pny, r.pnz}, o[3] = {r.otx, r.oty, r.otz};
            const double mn[3] = {r.mnx, r.mny, r.mnz}, mt[3] = {r.mtx, r.mty, r.mtz};
            for (int cc = 0; cc < 3; ++cc) {
                if (v->prev_n) v->prev_n[3 * q + cc] = pn[cc];
                if (v->o_t) v->o_t[3 * q + cc] = o[cc];
                if (v->m_n) v->m_n[3 * q + cc] = mn[cc];
                if (v->m_t) v->m_t[3 * q + cc] = mt[cc];
            }
        }
    }
    return RF_OK;
}

rf_status rf_set_state(rf_ctx* c, const rf_state_view* v) {
    if (!c || !v) return fail(RF_ERR_ARG, "null argument");
    if (!v->R || !v->V || !v->W || !v->nbr || !v->type || !v->bidx)
        return fail(RF_ERR_ARG, "rf_set_state needs R, V, W, nbr, type and bidx");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    const int64_t n = c->n, S = c->S, B = c->B;
    rf::DevState& d = c->d;
    const int cur = c->hctrl.cur;
    std::vector<double4> hp(n), hv(n), hw(n), hf(n), hm(n);
    HIP_TRY(hipMemcpy(hp.data(), d.pos[cur], n * sizeof(double4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hv.data(), d.vel[cur], n * sizeof(double4), hipMemcpyDeviceToHost));
    const std::vector<int32_t>& og = c->orig;  // caller id of stored particle
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = og[i];
        hp[i] = make_double4(v->R[o], v->R[n + o], v->R[2 * n + o], hp[i].w);
        hv[i] = make_double4(v->V[o], v->V[n + o], v->V[2 * n + o], hv[i].w);
        hw[i] = make_double4(v->W[o], v->W[n + o], v->W[2 * n + o], 0.0);
        hf[i] = v->F ? make_double4(v->F[o], v->F[n + o], v->F[2 * n + o], 0.0)
                     : make_double4(0.0, 0.0, 0.0, 0.0);
        hm[i] = v->M ? make_double4(v->M[o], v->M[n + o], v->M[2 * n + o], 0.0)
                     : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    HIP_TRY(hipMemcpy(d.pos[cur], hp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.vel[cur], hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.angv[cur], hw.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.F, hf.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.M, hm.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    std::vector<int32_t> nbr(n * S, 0), nbond(n, 0), bpart(n * B, 0);
    std::vector<uint32_t> occ(n, 0), bnd(n, 0), otnz(n, 0);
    std::vector<double4> ot(n * S, make_double4(0.0, 0.0, 0.0, 0.0));
    std::vector<int8_t> bslot(n * B, -1);
    std::vector<rf::BondRec> brec(n * B);
    std::memset(brec.data(), 0, brec.size() * sizeof(rf::BondRec));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = 0; k < S; ++k) {
            const int64_t qs = i * S + k;            // stored slot
            const int64_t q = (int64_t)og[i] * S + k;  // the caller's slot
            if (v->nbr[q] < 0) continue;
            if (v->nbr[q] >= n) return fail(RF_ERR_ARG, "slot partner out of range");
            occ[i] |= 1u << k;
            nbr[qs] = c->inv[v->nbr[q]];
            if (v->type[q] == 1) {
                const int b = v->bidx[q];
                if (b < 0 || b >= B) return fail(RF_ERR_ARG, "bond slot without a valid bidx");
                bnd[i] |= 1u << k;
                const int64_t qb = i * B + b;
                bslot[qb] = (int8_t)k;
                bpart[qb] = c->inv[v->nbr[q]];
                nbond[i] = std::max(nbond[i], b + 1);
                rf::BondRec& r = brec[qb];
                r.invL0 = v->inv_L0 ? v->inv_L0[q] : 0.0;
                r.L0 = 1.0 / r.invL0;
                r.sc = v->sigma_c ? v->sigma_c[q] : 0.0;
                r.tc = v->tau_c ? v->tau_c[q] : 0.0;
                double* pn[3] = {&r.pnx, &r.pny, &r.pnz};
                double* o[3] = {&r.otx, &r.oty, &r.otz};
                double* mn[3] = {&r.mnx, &r.mny, &r.mnz};
                double* mt[3] = {&r.mtx, &r.mty, &r.mtz};
                for (int cc = 0; cc < 3; ++cc) {
                    *pn[cc] = v->prev_n ? v->prev_n[3 * q + cc] : 0.0;
                    *o[cc] = v->o_t ? v->o_t[3 * q + cc] : 0.0;
                    *mn[cc] = v->m_n ? v->m_n[3 * q + cc] : 0.0;
                    *mt[cc] = v->m_t ? v->m_t[3 * q + cc] : 0.0;
                }
            } else if (v->o_t) {
                ot[qs] = make_double4(v->o_t[3 * q], v->o_t[3 * q + 1], v->o_t[3 * q + 2], 0.0);
                otnz[i] |= 1u << k;
            }
        }
    HIP_TRY(hipMemcpy(d.nbr, nbr.data(), n * S * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.occ, occ.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.bnd, bnd.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.otnz, otnz.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.ot, ot.data(), n * S * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.nbond, nbond.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.bslot, bslot.data(), n * B, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.bpart, bpart.data(), n * B * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.brec, brec.data(), n * B * sizeof(rf::BondRec), hipMemcpyHostToDevice));
    st = rebuild(c);
    if (st) return st;
    st = sync_ctrl(c);
    if (st) return st;
    if (c->hctrl.err) return device_error(c);
    c->built = true;
    return RF_OK;
}

}  // extern "C"

// ------------------------------------------------------------- batches ----
// Several samples advance with one step launch and one controller launch
// per step (k_step_batch, k_controller_batch).  Each sample keeps its own
// arrays, constants and control block; a block-to-sample table routes the
// blocks of the shared launch.  A sample stops at its own step target
// (Ctrl::stop_at), at the end of its test, or for a list rebuild, which the
// host then runs for that sample alone.
//
// Autotuning: the order of the block table is free (a block's work depends
// only on its sample and its local index).  The first steps of a batch run
// short chunks in each of five ways, timed on the host clock around the
// synchronised chunk, and the fastest is kept: one launch per step with the
// samples' blocks one sample after another, interleaved, or largest sample
// first, or no shared launch at all -- each sample stepped on its own
// stream, one after another, as rf_step does (large samples fill the GPU
// alone, and the shared launch's per-block table lookups then cost more
// than they save), or each sample on its own stream with every sample's
// chunk launched before any is waited for (the streams overlap on the
// device's queues).  Full chunks replay as captured graphs (their kernels
// read the tables and entries from device memory, so one graph per chunk
// length serves every order), and chunks shorten while list rebuilds come
// often.

constexpr int kBatchOrders = 5;       // three block orders, one after another, concurrent
constexpr int kBatchSequential = 3;
constexpr int kBatchConcurrent = 4;
// Steps of each form's timed trial chunk: enough that per-chunk host work
// (64 graph launches and waits in the concurrent form) weighs as it does in
// the run's longer chunks.
constexpr int kTuneSteps = 128;

struct rf_batch {
    std::vector<rf_ctx*> c;
    int device = 0;
    hipStream_t stream = nullptr;
    rf::BatchEntry* d_e = nullptr;
    int32_t* d_map = nullptr;
    int32_t* d_loc = nullptr;
    int32_t nblk_total = 0;
    int32_t chunk = 256;
    int32_t chunk_eff = 256;
    int32_t order = -1;      // table order in use
    int32_t chosen = -1;     // order picked by the tuner (-1: not yet)
    int32_t tuning = 0;      // next order to time
    double tune_ms[kBatchOrders] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double last_ms = 0.0;
    struct Graph {
        int64_t steps = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
};

namespace {

rf_status batch_upload(rf_batch* b) {
    std::vector<rf::BatchEntry> e(b->c.size());
    for (size_t k = 0; k < b->c.size(); ++k) {
        const rf_ctx* c = b->c[k];
        e[k].s = c->d;
        e[k].p = c->p;
        e[k].ctrl = c->ctrl;
    }
    HIP_TRY(hipMemcpyAsync(b->d_e, e.data(), e.size() * sizeof(rf::BatchEntry),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return RF_OK;
}

// Block table for an order (see rf_batch).
rf_status batch_order(rf_batch* b, int order) {
    if (order == b->order) return RF_OK;
    const size_t count = b->c.size();
    std::vector<int32_t> smp, loc;
    smp.reserve(b->nblk_total);
    loc.reserve(b->nblk_total);
    std::vector<size_t> seq(count);
    for (size_t k = 0; k < count; ++k) seq[k] = k;
    if (order == 2)
        std::stable_sort(seq.begin(), seq.end(), [&](size_t u, size_t v) {
            return b->c[u]->p.nblk > b->c[v]->p.nblk;
        });
    if (order == 1) {
        std::vector<int32_t> next(count, 0);
        for (bool more = true; more;) {
            more = false;
            for (size_t k = 0; k < count; ++k)
                if (next[k] < b->c[k]->p.nblk) {
                    smp.push_back((int32_t)k);
                    loc.push_back(next[k]++);
                    more = true;
                }
        }
    } else {
        for (size_t k : seq)
            for (int32_t q = 0; q < b->c[k]->p.nblk; ++q) {
                smp.push_back((int32_t)k);
                loc.push_back(q);
            }
    }
    HIP_TRY(hipMemcpyAsync(b->d_map, smp.data(), smp.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_loc, loc.data(), loc.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->order = order;
    return RF_OK;
}

// `k` (step, controller) launch pairs of the batch, as one captured graph
// when k is the current full chunk length.
rf_status batch_launch(rf_batch* b, int64_t k, bool graphed) {
    const int32_t count = (int32_t)b->c.size();
    if (!graphed) {
        for (int64_t s = 0; s < k; ++s) {
            HIP_TRY(rf::launch_step_batch(b->d_e, b->d_map, b->d_loc, b->nblk_total, b->stream));
            HIP_TRY(rf::launch_controller_batch(b->d_e, count, b->stream));
        }
        return RF_OK;
    }
    rf_batch::Graph* g = nullptr;
    for (auto& q : b->graphs)
        if (q.steps == k) g = &q;
    if (!g) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (int64_t s = 0; s < k; ++s) {
            const hipError_t e1 =
                rf::launch_step_batch(b->d_e, b->d_map, b->d_loc, b->nblk_total, b->stream);
            const hipError_t e2 = rf::launch_controller_batch(b->d_e, count, b->stream);
            if (le == hipSuccess) le = e1 != hipSuccess ? e1 : e2;
        }
        hipGraphExec_t ex = nullptr;
        rf_status st = end_capture(b->stream, le, &ex);
        if (st) return st;
        b->graphs.push_back(rf_batch::Graph{k, ex});
        g = &b->graphs.back();
    }
    HIP_TRY(hipGraphLaunch(g->exec, b->stream));
    return RF_OK;
}

rf_status batch_sync(rf_batch* b) {
    for (rf_ctx* c : b->c) {
        rf_status st = ensure_pinned(c);
        if (st) return st;
        HIP_TRY(hipMemcpyAsync(c->hpin, c->ctrl, sizeof(rf::Ctrl), hipMemcpyDeviceToHost,
                               b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (rf_ctx* c : b->c) c->hctrl = *c->hpin;
    return RF_OK;
}

}  // namespace

extern "C" {

rf_status rf_batch_create(rf_ctx* const* ctxs, int32_t count, rf_batch** out) {
    if (out) *out = nullptr;
    if (!ctxs || count <= 0 || !out) return fail(RF_ERR_ARG, "rf_batch_create: bad arguments");
    rf_batch* b = new rf_batch();
    b->c.assign(ctxs, ctxs + count);
    b->device = b->c[0]->device;
    int64_t total = 0;
    for (int32_t k = 0; k < count; ++k) {
        rf_ctx* c = b->c[k];
        if (!c || !c->built || c->device != b->device) {
            delete b;
            return fail(RF_ERR_ARG, "batch member %d is null, not built or on another device", k);
        }
        for (int32_t q = 0; q < k; ++q)
            if (b->c[q] == c) {
                delete b;
                return fail(RF_ERR_ARG, "context %d appears twice in the batch", k);
            }
        total += c->p.nblk;
    }
    if (total >= ((int64_t)1 << 31)) {
        delete b;
        return fail(RF_ERR_ARG, "batch too large");
    }
    b->nblk_total = (int32_t)total;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipMalloc(&b->d_e, count * sizeof(rf::BatchEntry));
    if (e == hipSuccess) e = hipMalloc(&b->d_map, std::max<int64_t>(total, 1) * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_loc, std::max<int64_t>(total, 1) * sizeof(int32_t));
    if (e != hipSuccess) {
        rf_batch_destroy(b);
        return fail(RF_ERR_HIP, "batch allocation failed: %s", hipGetErrorString(e));
    }
    rf_status st = batch_order(b, 0);
    if (st) {
        rf_batch_destroy(b);
        return st;
    }
    *out = b;
    return RF_OK;
}

rf_status rf_batch_destroy(rf_batch* b) {
    if (!b) return RF_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->d_e) hipFree(b->d_e);
    if (b->d_map) hipFree(b->d_map);
    if (b->d_loc) hipFree(b->d_loc);
    for (auto& g : b->graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return RF_OK;
}

rf_status rf_batch_step(rf_batch* b, int64_t nsteps, int64_t* done_steps) {
    if (!b || nsteps < 0) return fail(RF_ERR_ARG, "rf_batch_step: bad arguments");
    HIP_TRY(hipSetDevice(b->device));
    const size_t count = b->c.size();
    for (rf_ctx* c : b->c) HIP_TRY(hipStreamSynchronize(c->stream));
    rf_status st = batch_sync(b);
    if (st) return st;
    std::vector<int64_t> start(count);
    for (size_t k = 0; k < count; ++k) {
        start[k] = b->c[k]->hctrl.step;
        const int64_t stop = nsteps > INT64_MAX / 4 ? INT64_MAX / 4 : start[k] + nsteps;
        HIP_TRY(rf::launch_set_stop(b->c[k]->ctrl, stop, b->stream));
    }
    st = batch_upload(b);
    if (st) return st;
    st = batch_sync(b);
    if (st) return st;
    std::vector<int64_t> stop(count);
    for (size_t q = 0; q < count; ++q)
        stop[q] = nsteps > INT64_MAX / 4 ? INT64_MAX / 4 : start[q] + nsteps;
    using clock = std::chrono::steady_clock;
    const auto t_begin = clock::now();
    rf_status result = RF_OK;
    for (;;) {
        bool active = false;
        for (rf_ctx* c : b->c) active = active || !c->hctrl.halt;
        if (!active) break;
        const bool tune = b->chosen < 0;
        const int mode = tune ? b->tuning : b->chosen;
        const bool seq = mode == kBatchSequential;
        const bool conc = mode == kBatchConcurrent;
        if (tune && !seq && !conc) {
            st = batch_order(b, b->tuning);
            if (st) { result = st; break; }
        }
        const int64_t k = tune ? kTuneSteps : b->chunk_eff;
        std::vector<int64_t> before(count), reb(count);
        std::vector<char> idle(count);  // halted for good before the chunk
        for (size_t q = 0; q < count; ++q) {
            before[q] = b->c[q]->hctrl.step;
            reb[q] = b->c[q]->rebuilds;
            idle[q] = b->c[q]->hctrl.halt != 0;
        }
        if (conc && k > 1) {  // graphs captured before the chunk (and its timing) starts
            for (size_t q = 0; q < count && !result; ++q) {
                rf_ctx* c = b->c[q];
                const int64_t left = stop[q] - c->hctrl.step;
                if (idle[q] || c->hctrl.done || left <= 0 || !c->use_graph) continue;
                rf_ctx::Graph* g = nullptr;
                st = ensure_graph(c, std::min<int64_t>(k, left), &g);
                if (st) result = st;
            }
            if (result) break;
        }
        const auto ta = clock::now();
        bool progress = false, rebuilt = false, full = true;
        if (seq) {
            // every sample on its own stream, one after another (rf_step
            // handles its rebuilds and keeps its control block in sync)
            for (size_t q = 0; q < count && !result; ++q) {
                rf_ctx* c = b->c[q];
                if (idle[q] || c->hctrl.done) continue;
                const int64_t left = stop[q] - c->hctrl.step;
                if (left <= 0) continue;
                st = rf_step(c, std::min<int64_t>(k, left), nullptr);
                if (st) result = st;
            }
            if (result) break;
            for (size_t q = 0; q < count; ++q) {
                rf_ctx* c = b->c[q];
                progress = progress || c->hctrl.step != before[q];
                full = full && (idle[q] || c->hctrl.step - before[q] == k);
                rebuilt = rebuilt || c->rebuilds != reb[q];
            }
        } else if (conc) {
            // every active sample's chunk launched on its own stream before
            // any is waited for, so the samples' launches overlap on the device
            std::vector<char> launched(count, 0);
            for (size_t q = 0; q < count; ++q) {
                rf_ctx* c = b->c[q];
                if (idle[q] || c->hctrl.done) continue;
                const int64_t left = stop[q] - c->hctrl.step;
                if (left <= 0) continue;
                st = launch_chunk(c, std::min<int64_t>(k, left));
                if (st) { result = st; break; }
                launched[q] = 1;
            }
            for (size_t q = 0; q < count; ++q) {  // after an error too: drain what was launched
                if (!launched[q]) continue;
                st = sync_ctrl(b->c[q]);
                if (st && !result) result = st;
            }
            if (result) break;
            for (size_t q = 0; q < count && !result; ++q) {
                rf_ctx* c = b->c[q];
                progress = progress || c->hctrl.step != before[q];
                full = full && (idle[q] || c->hctrl.step - before[q] == k);
                if (!launched[q]) continue;
                if (c->hctrl.err) { result = device_error(c); break; }
                if (c->hctrl.rebuild) {
                    st = rebuild(c);
                    if (st) { result = st; break; }
                    if (c->hctrl.err) { result = device_error(c); break; }
                    rebuilt = true;
                }
            }
            if (result) break;
        } else {
            st = batch_launch(b, k, !tune && k > 1);
            if (st) { result = st; break; }
            st = batch_sync(b);
            if (st) { result = st; break; }
            for (size_t q = 0; q < count && !result; ++q) {
                rf_ctx* c = b->c[q];
                progress = progress || c->hctrl.step != before[q];
                full = full && (idle[q] || c->hctrl.step - before[q] == k);
                if (c->hctrl.err) { result = device_error(c); break; }
                if (c->hctrl.rebuild) {
                    st = rebuild(c);
                    if (st) { result = st; break; }
                    if (c->hctrl.err) { result = device_error(c); break; }
                    rebuilt = true;
                }
            }
            if (result) break;
        }
        const double chunk_ms = std::chrono::duration<double, std::milli>(clock::now() - ta).count();
        if (tune && full && !rebuilt) {  // a clean timing: every active sample ran every step
            b->tune_ms[b->tuning] = chunk_ms / (double)k;
            if (++b->tuning == kBatchOrders) {
                int best = 0;
                for (int o = 1; o < kBatchOrders; ++o)
                    if (b->tune_ms[o] < b->tune_ms[best]) best = o;
                b->chosen = best;
                if (best != kBatchSequential && best != kBatchConcurrent) {
                    st = batch_order(b, best);
                    if (st) { result = st; break; }
                }
            }
        } else if (!tune) {
            if (rebuilt && !full) {
                int64_t m = 0;
                for (size_t q = 0; q < count; ++q) m = std::max(m, b->c[q]->hctrl.step - before[q]);
                int64_t len = 8;
                while (len < m) len <<= 1;
                b->chunk_eff = (int32_t)std::min<int64_t>(len, b->chunk);
            } else if (full) {
                b->chunk_eff = (int32_t)std::min<int64_t>(2 * (int64_t)b->chunk_eff, b->chunk);
            }
        }
        if (rebuilt) {
            st = batch_upload(b);  // a rebuild may have narrowed the skin
            if (st) { result = st; break; }
        } else if (!progress) {
            result = fail(RF_ERR_STATE, "batch halted without a reason");
            break;
        }
    }
    b->last_ms = std::chrono::duration<double, std::milli>(clock::now() - t_begin).count();
    for (size_t k = 0; k < count; ++k) HIP_TRY(rf::launch_set_stop(b->c[k]->ctrl, 0, b->stream));
    st = batch_sync(b);
    if (st && !result) result = st;
    if (done_steps)
        for (size_t k = 0; k < count; ++k) done_steps[k] = b->c[k]->hctrl.step - start[k];
    return result;
}

rf_status rf_batch_run(rf_batch* b, rf_test_state* out) {
    if (!b) return fail(RF_ERR_ARG, "null batch");
    for (rf_ctx* c : b->c)
        if (!c->has_test || c->cfg.kind == RF_TEST_NONE)
            return fail(RF_ERR_STATE, "rf_batch_run needs a platen test on every sample");
    rf_status st = rf_batch_step(b, INT64_MAX, nullptr);
    if (st) return st;
    for (size_t k = 0; k < b->c.size(); ++k) {
        rf_ctx* c = b->c[k];
        if (!c->hctrl.done) return fail(RF_ERR_STATE, "batch sample %zu did not finish", k);
        st = post_process(c);
        if (st) return st;
        if (out) fill_state(c, out + k);
    }
    return RF_OK;
}

rf_status rf_batch_tuning(rf_batch* b, int32_t* order, double* ms_per_step) {
    if (!b) return fail(RF_ERR_ARG, "null batch");
    if (order) *order = b->chosen;
    if (ms_per_step)
        for (int o = 0; o < kBatchOrders; ++o) ms_per_step[o] = b->tune_ms[o];
    return RF_OK;
}

rf_status rf_batch_set_form(rf_batch* b, int32_t form) {
    if (!b || form < -1 || form >= kBatchOrders)
        return fail(RF_ERR_ARG, "rf_batch_set_form: form %d not in -1..%d", form, kBatchOrders - 1);
    HIP_TRY(hipSetDevice(b->device));
    if (form < 0) {  // back to the autotuner
        b->chosen = -1;
        b->tuning = 0;
        for (int o = 0; o < kBatchOrders; ++o) b->tune_ms[o] = 0.0;
        return RF_OK;
    }
    b->chosen = form;
    if (form != kBatchSequential && form != kBatchConcurrent) {
        rf_status st = batch_order(b, form);
        if (st) return st;
        st = batch_sync(b);
        if (st) return st;
    }
    return RF_OK;
}

rf_status rf_batch_last_ms(rf_batch* b, double* total_ms) {
    if (!b || !total_ms) return fail(RF_ERR_ARG, "null argument");
    *total_ms = b->last_ms;
    return RF_OK;
}

}  // extern "C"

// -------------------------------------------------------- sensitivities ---
// Forward-mode derivatives of the loading response with respect to material
// parameters (rf_sens.hip): one run carries up to kSensK parameter
// directions next to the values.

namespace {

// d(E, G, hm_E, hm_G, mu, mu_r, p_A, E_wall, G_wall, ratio, b_abs)/d(param),
// differentiating rf_derive's formulas; the wall factor scales hm_E, hm_G.
rf_status material_tangent(const rf_material* m, double wall_factor, int32_t param, double out[11]) {
    for (int u = 0; u < 11; ++u) out[u] = 0.0;
    const double nu = m->nu, Eb = m->E_bond, Ec = m->E_contact;
    const double Gc = 0.5 * Ec / (1.0 + nu);
    switch (param) {
        case RF_PARAM_E_BOND:
            out[0] = 1.0;
            out[1] = 0.5 / (1.0 + nu);
            break;
        case RF_PARAM_NU: {
            out[1] = -0.5 * Eb / ((1.0 + nu) * (1.0 + nu));
            const double dGc = -0.5 * Ec / ((1.0 + nu) * (1.0 + nu));
            out[2] = Ec * nu / ((1.0 - nu * nu) * (1.0 - nu * nu));
            out[3] = 0.5 * (dGc * (2.0 - nu) + Gc) / ((2.0 - nu) * (2.0 - nu));
            break;
        }
        case RF_PARAM_E_CONTACT:
            out[2] = 0.5 / (1.0 - nu * nu);
            out[3] = 0.5 * (0.5 / (1.0 + nu)) / (2.0 - nu);
            break;
        case RF_PARAM_FRICTION: out[4] = 1.0; break;
        case RF_PARAM_ROLLING_FRICTION: out[5] = 1.0; break;
        case RF_PARAM_BOND_RADIUS_RATIO:
            if (!(m->bond_radius_ratio > 0.0))
                return fail(RF_ERR_ARG, "bond radius ratio is not in use (absolute bond radius)");
            out[9] = 1.0;
            break;
        default: return fail(RF_ERR_ARG, "unknown parameter %d", param);
    }
    out[7] = wall_factor * out[2];
    out[8] = wall_factor * out[3];
    return RF_OK;
}

// The particles of each wave of the sensitivity step.  A wave walks its
// particles' bonds over 64 / nl lanes per parameter direction, so its time
// follows ceil(bond entries / lanes).  Runs of consecutive particles (at most
// one per lane) are cut where the entries would exceed `target` per lane,
// with the smallest target whose runs the device holds at once; bond-dense
// regions then get shorter runs, and their lanes without a particle take
// bond positions handed out by the others.  The runs only balance the waves:
// every particle's sums keep their order whatever run it falls in.
rf_status sens_wave_table(rf_ctx* c) {
    rf::SensState& t = c->sens;
    const int64_t n = c->n;
    const int ppw = 64 / t.nl, B = c->B;
    std::vector<int32_t> nb((size_t)n);
    HIP_TRY(hipMemcpy(nb.data(), c->d.nbond, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const int64_t resident = (int64_t)std::max(cus, 1) * rf::kSensWavesPerCU;
    int64_t total = 0;
    for (int64_t q = 0; q < n; ++q) total += std::min(nb[q], B);
    std::vector<int32_t> wb;
    int64_t target = std::max<int64_t>(1, (total + resident * ppw - 1) / (resident * ppw));
    for (;; ++target) {
        wb.assign(1, 0);
        int cnt = 0;
        int64_t ent = 0;
        for (int64_t q = 0; q < n; ++q) {
            const int x = std::min(nb[q], B);
            if (cnt == ppw || (cnt > 0 && ent + x > ppw * target)) {
                wb.push_back((int32_t)q);
                cnt = 0;
                ent = 0;
            }
            ++cnt;
            ent += x;
        }
        wb.push_back((int32_t)n);
        // at target B every run is ppw particles long: no shorter runs exist
        if ((int64_t)wb.size() - 1 <= resident || target >= B) break;
    }
    t.nwaves = (int32_t)(wb.size() - 1);
    HIP_TRY(hipMemcpy(t.wbase, wb.data(), wb.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return RF_OK;
}

}  // namespace

extern "C" {

rf_status rf_sens_material_tangent(const rf_material* m, double wall_factor, int32_t param,
                                   double* out11) {
    if (!m || !out11) return fail(RF_ERR_ARG, "null argument");
    return material_tangent(m, wall_factor, param, out11);
}

rf_status rf_sens_begin(rf_ctx* c, int32_t nparam, const int32_t* params) {
    if (!c || !params || nparam < 1 || nparam > rf::kSensK)
        return fail(RF_ERR_ARG, "rf_sens_begin: 1..%d parameters", rf::kSensK);
    if (!c->built) return fail(RF_ERR_STATE, "rf_sens_begin before rf_build_bonds");
    // A derived time step (rf_material.dt <= 0) follows the parameters of
    // the time scale that sets it (rf_derive): the Rayleigh time moves with
    // E_contact and nu, the bond time with E_bond and the bond radius.  The
    // tangents hold dt constant, so such a direction is refused.
    if (!(c->mat.dt > 0.0)) {
        const bool rayleigh = c->der.T_rayleigh < c->der.T_bond;
        for (int k = 0; k < nparam; ++k) {
            const int32_t q = params[k];
            const bool moves = rayleigh ? (q == RF_PARAM_E_CONTACT || q == RF_PARAM_NU)
                                        : (q == RF_PARAM_E_BOND || q == RF_PARAM_BOND_RADIUS_RATIO);
            if (moves)
                return fail(RF_ERR_STATE,
                            "rf_sens_begin: dt is derived from the material (rf_material.dt <= 0) "
                            "and moves with parameter %d; set rf_material.dt to hold it fixed",
                            (int)q);
        }
    }
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    rf::SensState& t = c->sens;
    const int64_t n = c->n, K = rf::kSensK;
    if (!c->sens_on) {
        for (int b = 0; b < 2; ++b) {
            rf_status st = c->alloc(&t.tpos[b], K * n);
            if (!st) st = c->alloc(&t.tvel[b], K * n);
            if (!st) st = c->alloc(&t.tang[b], K * n);
            if (st) return st;
        }
        rf_status st = c->alloc(&t.tbrec, K * n * c->B * 12);
        if (!st) st = c->alloc(&t.tot, K * n * c->S);
        if (!st) st = c->alloc(&t.pp, 2 * n);
        if (!st) st = c->alloc(&t.pnb, n);
        if (!st) st = c->alloc(&t.tpp, 2 * K * n);
        if (!st) st = c->alloc(&t.tpart_b, K * c->p.nblk);
        if (!st) st = c->alloc(&t.tpart_t, K * c->p.nblk);
        if (!st) st = c->alloc(&t.wbase, n + 2);
        if (st) return st;
        c->sens_on = true;
    }
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMemsetAsync(t.tpos[b], 0, K * n * sizeof(double4), c->stream));
        HIP_TRY(hipMemsetAsync(t.tvel[b], 0, K * n * sizeof(double4), c->stream));
        HIP_TRY(hipMemsetAsync(t.tang[b], 0, K * n * sizeof(double4), c->stream));
    }
    HIP_TRY(hipMemsetAsync(t.tbrec, 0, K * n * c->B * 12 * sizeof(double), c->stream));
    HIP_TRY(hipMemsetAsync(t.tot, 0, K * n * c->S * sizeof(double4), c->stream));
    const int64_t cap = std::max<int64_t>(c->p.curve_cap, 1);
    if (t.tcurve) {
        hipFree(t.tcurve);
        c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)t.tcurve),
                        c->allocs.end());
        t.tcurve = nullptr;
    }
    rf_status st = c->alloc(&t.tcurve, cap * K * 2);
    if (st) return st;
    const double wf = c->has_test ? c->cfg.wall_modulus_factor : 0.0;
    for (int k = 0; k < K; ++k) {
        if (k < nparam) {
            st = material_tangent(&c->mat, wf, params[k], t.dmat[k]);
            if (st) return st;
        } else {
            for (int u = 0; u < 11; ++u) t.dmat[k][u] = 0.0;
        }
    }
    t.nl = nparam;
    st = sens_wave_table(c);
    if (st) return st;
    t.br_step0 = c->hctrl.step;
    c->sens_np = nparam;
    c->sens_curve0 = c->hctrl.stepsave;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return RF_OK;
}

rf_status rf_sens_branches(rf_ctx* c, int32_t mode, int64_t steps, const rf_ctx* recorded) {
    if (!c) return fail(RF_ERR_ARG, "null context");
    rf::SensState& t = c->sens;
    if (mode == RF_BRANCH_DECIDE) {
        t.br_mode = 0;
        return RF_OK;
    }
    HIP_TRY(hipSetDevice(c->device));
    if (mode == RF_BRANCH_RECORD) {
        if (steps <= 0) return fail(RF_ERR_ARG, "rf_sens_branches: steps must be > 0");
        if (!c->br_own || c->br_cap < steps) {
            for (void* q : {(void*)c->br_bits, (void*)c->br_w, (void*)c->br_vw}) {
                if (!q) continue;
                hipFree(q);
                c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), q), c->allocs.end());
            }
            c->br_bits = nullptr;
            c->br_w = nullptr;
            c->br_vw = nullptr;
            rf_status st = c->alloc(&c->br_bits, (size_t)steps * c->n * rf::kBrWords);
            if (!st) st = c->alloc(&c->br_w, (size_t)steps * c->n);
            if (!st) st = c->alloc(&c->br_vw, (size_t)steps);
            if (st) return st;
            c->br_cap = steps;
            c->br_own = true;
        }
        t.br_mode = 1;
        t.br_steps = steps;
        t.br_bits = c->br_bits;
        t.br_w = c->br_w;
        t.br_vw = c->br_vw;
        c->br_recorded = 0;
        return RF_OK;
    }
    if (mode == RF_BRANCH_REPLAY) {
        if (!recorded || recorded == c)
            return fail(RF_ERR_ARG, "rf_sens_branches: replay needs another, recorded context");
        if (recorded->n != c->n || recorded->S != c->S || recorded->device != c->device ||
            recorded->orig != c->orig)
            return fail(RF_ERR_ARG, "rf_sens_branches: the recorded context holds another sample");
        if (!recorded->br_own || recorded->br_recorded <= 0)
            return fail(RF_ERR_STATE, "rf_sens_branches: the context recorded no window");
        t.br_mode = 2;
        t.br_steps = recorded->br_recorded;
        t.br_bits = recorded->br_bits;
        t.br_w = recorded->br_w;
        t.br_vw = recorded->br_vw;
        return RF_OK;
    }
    return fail(RF_ERR_ARG, "rf_sens_branches: unknown mode %d", mode);
}

rf_status rf_sens_step(rf_ctx* c, int64_t nsteps, int64_t* done_steps) {
    if (done_steps) *done_steps = 0;
    if (!c || !c->sens_on || c->sens_np < 1) return fail(RF_ERR_STATE, "rf_sens_step before rf_sens_begin");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    const int64_t start = c->hctrl.step;
    const rf::SensState& tb = c->sens;
    if (tb.br_mode && start + nsteps > tb.br_step0 + tb.br_steps)
        return fail(RF_ERR_ARG, "rf_sens_step: the window would run past the %lld steps of its "
                                "frozen-branch record", (long long)tb.br_steps);
    int64_t remaining = nsteps;
    while (remaining > 0 && !c->hctrl.done) {
        const int64_t k = std::min<int64_t>(remaining, c->chunk_eff);
        const int64_t before = c->hctrl.step;
        for (int64_t s = 0; s < k; ++s) {
            HIP_TRY(rf::launch_sens_step(c->d, c->p, c->sens, c->ctrl, c->stream));
            HIP_TRY(rf::launch_controller_sens(c->d, c->p, c->sens, c->ctrl, c->stream));
        }
        st = sync_ctrl(c);
        if (st) return st;
        const int64_t executed = c->hctrl.step - before;
        remaining -= executed;
        if (c->hctrl.err & rf::kErrSensList)
            return fail(RF_ERR_STATE, "sensitivity window: a touching pair has no slot, two "
                                      "particles coincide, or a particle left the box");
        if (c->hctrl.err) return device_error(c);
        adapt_chunk(c, k, executed, c->hctrl.rebuild != 0);
        if (c->hctrl.rebuild) {
            st = rebuild(c);
            if (st) return st;
            if (c->hctrl.err) return device_error(c);
        } else if (executed == 0 && !c->hctrl.done) {
            return fail(RF_ERR_STATE, "sensitivity loop halted without a reason");
        }
    }
    if (done_steps) *done_steps = c->hctrl.step - start;
    if (c->sens.br_mode == 1)
        c->br_recorded = std::max<int64_t>(c->br_recorded, c->hctrl.step - c->sens.br_step0);
    return RF_OK;
}

rf_status rf_sens_get_curve(rf_ctx* c, double* out, int64_t cap, int64_t* n) {
    if (n) *n = 0;
    if (!c || !c->sens_on) return fail(RF_ERR_STATE, "no sensitivity run");
    HIP_TRY(hipSetDevice(c->device));
    rf_status st = sync_ctrl(c);
    if (st) return st;
    const int64_t m = std::min<int64_t>(c->hctrl.stepsave, c->p.curve_cap) - c->sens_curve0;
    const int64_t cnt = std::max<int64_t>(0, std::min<int64_t>(m, cap));
    std::vector<double> h((size_t)std::max<int64_t>(c->p.curve_cap, 1) * rf::kSensK * 2);
    HIP_TRY(hipMemcpy(h.data(), c->sens.tcurve, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int np = c->sens_np;
    for (int64_t q = 0; q < cnt; ++q)
        for (int k = 0; k < np; ++k)
            for (int e = 0; e < 2; ++e)
                out[(q * np + k) * 2 + e] = h[((c->sens_curve0 + q) * rf::kSensK + k) * 2 + e];
    if (n) *n = cnt;
    return RF_OK;
}

rf_status rf_adjoint_run(rf_ctx* c, const rf_adjoint_cfg* a, double* values, double* grad) {
    if (!c || !a || !values || !grad) return fail(RF_ERR_ARG, "rf_adjoint_run: null argument");
    if (a->steps < 1) return fail(RF_ERR_ARG, "rf_adjoint_run: the window needs at least one step");
    if (!c->has_test || c->cfg.kind == RF_TEST_NONE)
        return fail(RF_ERR_STATE, "rf_adjoint_run needs a platen test (rf_set_test)");
    rf_status st = rf_sens_begin(c, a->nparam, a->params);
    if (st) return st;
    int64_t done = 0;
    st = rf_sens_step(c, a->steps, &done);
    if (st) return st;
    if (done != a->steps)
        return fail(RF_ERR_STATE, "rf_adjoint_run: the test ended after %lld of %lld window steps",
                    (long long)done, (long long)a->steps);
    const int np = a->nparam;
    const int64_t p0 = c->sens_curve0;
    int64_t m = 0;
    std::vector<double> curve(3 * (size_t)std::max<int64_t>(c->p.curve_cap, 1));
    st = rf_get_curve(c, curve.data(), c->p.curve_cap, &m);
    if (st) return st;
    const int64_t cnt = m - p0;
    if (cnt < 1) return fail(RF_ERR_STATE, "rf_adjoint_run: the window recorded no curve point");
    std::vector<double> tan((size_t)cnt * np * 2);
    int64_t got = 0;
    st = rf_sens_get_curve(c, tan.data(), cnt, &got);
    if (st) return st;
    if (got != cnt) return fail(RF_ERR_STATE, "rf_adjoint_run: curve and tangents differ in length");
    auto sgn = [](double x) { return (double)((x > 0.0) - (x < 0.0)); };
    double sum = 0.0, best = -1.0;
    int64_t q = 0;
    for (int64_t i = 0; i < cnt; ++i) {
        const double* pt = &curve[3 * (p0 + i)];
        const double sa = 0.5 * (std::fabs(pt[1]) + std::fabs(pt[2]));
        sum += sa;
        if (sa > best) {  // first point of the highest stress
            best = sa;
            q = i;
        }
    }
    const double eps = -curve[3 * (p0 + q)];
    if (!(eps > 0.0))
        return fail(RF_ERR_STATE, "rf_adjoint_run: the window's highest stress comes before any "
                                  "platen travel");
    values[0] = sum / (double)cnt;
    values[1] = best / eps;
    for (int k = 0; k < np; ++k) {
        double dsum = 0.0, dpeak = 0.0;
        for (int64_t i = 0; i < cnt; ++i) {
            const double* pt = &curve[3 * (p0 + i)];
            const double* t = &tan[((size_t)i * np + k) * 2];
            const double d = 0.5 * (sgn(pt[1]) * t[0] + sgn(pt[2]) * t[1]);
            dsum += d;
            if (i == q) dpeak = d;
        }
        grad[k] = dsum / (double)cnt;
        grad[np + k] = dpeak / eps;
    }
    return RF_OK;
}

}  // extern "C"
