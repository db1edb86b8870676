// RockFit-DEM -- command-line drop-in of the reference binary on the
// MI355X library (include/rockfit_dem.h).
//
// Argument grammar of kernel.cu:23-124 (the same flags and the same quirks:
// the "CS" shift carries over to later arguments), stdout lines the Python
// farm reads (NN_MP.py:85-86, 769-771):
//   There are <n> GPUs
//   Correct GPU request! Requested: <d>
//   INDEX <i>
//   ResultCalculateOneSample[E,sigma_c,tau_c,b_r,gap,porosity,
//                            epsU,sigmaU,LinDevU,DestrU,epsB,sigmaB,0,DestrB]
//   Time <start> | <end>
// and the three drivers of the reference binary:
//   M(...)  CalculateOneSample (mp_calculateonesample.cu:17-250)
//   DCG     DetermineConjGradient (mp_determineconjgradient.cu:146-368): the
//           finite-difference runs of each step go to the GPU as one batch
//   TSS     TestSeveralSamples (mp_testseveralsamples.cu:16-161)
// Material from SetInitialParameters (mp_determineconjgradient.cu:15-47) and
// SetVariationParameters (:49-71); samples ./result/Sample_<k><part>.txt
// (md_createcylindersample.cu:144-182, 441-466); curves
// ./result/Uniaxial_<i>.txt and ./result/Brazil_<i>.txt (md_results.cu:168-180).
// Sample creation grows a seeded random packing with FIRE on the GPU
// (rf_fire_relax); the reference's packing is time-seeded.
//
// Environment:
//   RF_SEED               seed of sample creation (default 20261203)
//   RF_CLI_DRY_RUN=1      print the parsed flags to stderr and stop
//   RF_CG_STEPS           DCG steps (default 50, the reference's NSteps)
//   RF_CG_REUSE_SAMPLES=1 DCG loads ./result/Sample_<k>.txt when it exists
//                         instead of creating it
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "rf_cg.h"
#include "rockfit_dem.h"

namespace {

constexpr double kStress = 1e9;    // Pa per reduced stress unit
constexpr double kLength = 1e-3;   // m per mm
constexpr double kVelocity = 1e3;  // m/s per reduced velocity unit
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kPr = 0.5;  // particle radius, mm (SetInitialParameters)

struct Geometry {
    double L[3];      // LSi - 2 p_r (centres), mm
    double space[3];  // SpaceSize, mm
};
// mp_calculateonesample.cu:91: uniaxial cylinder and Brazilian disc
const Geometry kGeom[2] = {{{25.0 - 2 * kPr, 25.0 - 2 * kPr, 50.0 - 2 * kPr}, {40.0, 40.0, 60.0}},
                           {{50.0 - 2 * kPr, 50.0 - 2 * kPr, 20.0 - 2 * kPr}, {65.0, 65.0, 30.0}}};

// Loading of one test: velocity (m/s), MinStress (Pa), MaxStrain,
// ConditionFracture, dNsave.
struct Loading {
    int kind;
    double v_mps, min_stress_pa, max_strain, cond;
    int dnsave;
};
// CalculateOneSample, mp_calculateonesample.cu:120-124, dNSaveU/B :80-81
const Loading kOneU{RF_TEST_UNIAXIAL, -5.0 * 0.02, 10e6, 0.01, 0.1, 100};
const Loading kOneB{RF_TEST_BRAZILIAN, -5.0 * 0.02, 2.1e6, 0.03, 0.1, 1000};
// DetermineConjGradient, mp_determineconjgradient.cu:321-328
const Loading kCgU{RF_TEST_UNIAXIAL, -5.0 * 0.02, 10e6, 0.025, 0.2, 100};
const Loading kCgB{RF_TEST_BRAZILIAN, -5.0 * 0.02, 2.1e6, 0.04, 0.8, 1000};
// TestSeveralSamples, mp_testseveralsamples.cu:91-95, 51-52
const Loading kTssU{RF_TEST_UNIAXIAL, -2.0 * 0.02, 10e6, 0.03, 0.1, 100};
const Loading kTssB{RF_TEST_BRAZILIAN, -10.0 * 0.02, 2.1e6, 0.04, 0.9, 1000};

bool ok(rf_status s, const char* what) {
    if (s == RF_OK) return true;
    std::cerr << what << " failed: " << rf_last_error() << "\n";
    return false;
}

struct Sample {
    std::vector<double> x, y, z;
};

std::string sample_path(int k, const char* part) {
    char name[256];
    std::snprintf(name, sizeof name, "./result/Sample_%d%s.txt", k, part);
    return name;
}

// LoadParticleDATAtxt (md_createcylindersample.cu:165-182): N, then x y z in
// metres, scaled by 1/length.
bool read_sample(const std::string& path, Sample& s) {
    std::ifstream f(path);
    size_t n = 0;
    if (!(f >> n)) {
        std::cerr << "cannot read " << path << "\n";
        return false;
    }
    const double lc = 1.0 / kLength;
    s.x.resize(n);
    s.y.resize(n);
    s.z.resize(n);
    std::string line;
    std::getline(f, line);
    for (size_t i = 0; i < n; ++i) {
        double a, b, c;
        if (!std::getline(f, line) || std::sscanf(line.c_str(), "%lf %lf %lf", &a, &b, &c) != 3) {
            std::cerr << path << ": short file\n";
            return false;
        }
        s.x[i] = a * lc;
        s.y[i] = b * lc;
        s.z[i] = c * lc;
    }
    return true;
}

// SaveParticleDATAtxt (md_createcylindersample.cu:144-163): particles at the
// hidden point are left out.
bool write_sample(const std::string& path, const Sample& s, const double hidden[3]) {
    std::vector<size_t> keep;
    for (size_t i = 0; i < s.x.size(); ++i) {
        const double dx = s.x[i] - hidden[0], dy = s.y[i] - hidden[1], dz = s.z[i] - hidden[2];
        if (dx * dx + dy * dy + dz * dz >= 1e-10) keep.push_back(i);
    }
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::cerr << "cannot write " << path << "\n";
        return false;
    }
    std::fprintf(f, "%zu\n", keep.size());
    for (size_t i : keep)
        std::fprintf(f, "%.17g %.17g %.17g\n", s.x[i] * kLength, s.y[i] * kLength, s.z[i] * kLength);
    std::fclose(f);
    return true;
}

// Material point P (reduced units: E, sigma_c, tau_c, b_r, gap) with the
// fixed constants of SetInitialParameters: nu 0.247, rho 2610, e 0.6,
// mu 0.45, mu_r 0.05, one modulus for bonds and contacts, b_d = 2 p_r + gap,
// acut = acut_factor p_r (3.2 there, 2.9 in SetVariationParameters).
rf_material material_of(const double P[5], double acut_factor) {
    rf_material m{};
    m.E_bond = P[0];
    m.E_contact = P[0];
    m.nu = 0.247;
    m.density = 2.610;
    m.restitution = 0.6;
    m.friction = 0.45;
    m.rolling_friction = 0.05;
    m.bond_radius = P[3];
    m.bond_radius_ratio = 0.0;
    m.bond_gap = P[4];
    m.sigma_c = P[1];
    m.tau_c = P[2];
    m.acut = acut_factor * kPr;
    m.quirks = RF_QUIRKS_REFERENCE;
    return m;
}

bool box_of(int k, double acut, rf_box* box) {
    return ok(rf_box_reference(kGeom[k].space, 2.0 * acut, box), "rf_box_reference");
}

// The box of a run whose cell edge differs from the one the space box was
// set up with: DetermineConjGradient sets the sample boxes up with 2 x 3.2 p_r
// (SetInitialParameters before sample_data::CalculateParameters) and its
// runs re-grid them with 2 x 2.9 p_r (SetVariationParameters, then
// cell_data::CalculateParameters over the same B, md_data_types.h:42-55).
rf_box regrid(const rf_box& base, double cell) {
    rf_box b = base;
    b.cell = cell;
    const double inv = 1.0 / cell;
    for (int d = 0; d < 3; ++d) {
        const double B = base.center[d] + base.half[d];
        b.dims[d] = (int32_t)std::floor(B * inv) + (d == 2 ? 4 : 1);
    }
    return b;
}

// Sample creation: PN = (1 - porosity) Vext / p_V spheres
// (mp_calculateonesample.cu:99, md_data_types.h:114) with uniform random
// centres in the centre cylinder (radius L.x/2, half height L.z/2, the
// reference's d_CylinderRestrictionZ bounds, md_createcylindersample.cu:304,
// 387), then grown to full size: FIRE relaxation (rf_fire_relax with the
// harmonic overlap force and absorbing container walls, under which FIRE
// converges on dense packings) at radii 0.6, 0.75, 0.85, 0.92, 0.97 and
// 1.0 p_r.  The last stage runs to MaximumParticleIntersection (gap_tol)
// and, like d_ReduceParticleNumber in the reference loop
// (md_createcylindersample.cu:404-415), removes the worst overlapping
// particle at checks that still fail late in the run.  Returns the porosity
// of the kept spheres (md_createcylindersample.cu:423-424).
bool create_sample(int k, double porosity, double gap_tol, const rf_box& box, int device,
                   const std::string& path, double* out_porosity) {
    const Geometry& g = kGeom[k];
    const double R0 = 0.5 * g.L[0], H0 = 0.5 * g.L[2];
    const double Vext = kPi * (R0 + kPr) * (R0 + kPr) * (g.L[2] + 2.0 * kPr);
    const double pV = 4.0 * (1.0 / 3.0) * kPi * kPr * kPr * kPr;
    const size_t n = (size_t)((1.0 - porosity) * Vext / pV);
    uint64_t seed = 20261203;
    if (const char* s = std::getenv("RF_SEED")) seed = std::strtoull(s, nullptr, 10);
    std::mt19937_64 rng(seed + 1000 * k);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    Sample s;
    s.x.resize(n);
    s.y.resize(n);
    s.z.resize(n);
    std::vector<double> r(n);
    for (size_t i = 0; i < n; ++i) {
        double X, Y;
        do {
            X = R0 * u(rng);
            Y = R0 * u(rng);
        } while (X * X + Y * Y > R0 * R0);
        s.x[i] = box.center[0] + X;
        s.y[i] = box.center[1] + Y;
        s.z[i] = box.center[2] + H0 * u(rng);
    }
    rf_fire_cfg c;
    if (!ok(rf_fire_defaults(&c), "rf_fire_defaults")) return false;
    c.D = 1.0;
    c.dt0 = 0.02;
    c.density = 2.610;
    c.quirks = 0;
    c.law = RF_FIRE_LAW_HARMONIC;
    c.container = RF_CONTAINER_CYLINDER_Z_ABSORB;
    for (int d = 0; d < 3; ++d) {
        c.center[d] = box.center[d];
        c.hidden[d] = box.hidden[d];
    }
    c.radius = R0;
    c.half_height = H0;
    c.hide_z = box.hidden[2] - 0.5;
    c.device = device;
    const double stages[] = {0.6, 0.75, 0.85, 0.92, 0.97, 1.0};
    rf_fire_result res{};
    for (double sc : stages) {
        const bool last = sc == 1.0;
        std::fill(r.begin(), r.end(), sc * kPr);
        c.max_steps = 50;
        c.hard_cap = last ? 60000 : 20000;
        c.gap_tol = last ? gap_tol : std::min(gap_tol, -1e-3);
        c.check_every = last ? 50 : 20;
        c.delete_after = last ? 10000 : -1;
        if (!ok(rf_fire_relax((int64_t)n, s.x.data(), s.y.data(), s.z.data(), r.data(), &c, nullptr,
                              &res),
                "rf_fire_relax"))
            return false;
        std::cerr << "FIRE radius " << sc * kPr << " steps " << res.steps << " min gap " << res.min_gap
                  << " removed " << res.removed << " stop " << res.stop_reason << "\n";
    }
    const size_t kept = n - (size_t)res.removed;
    std::cerr << "FIN FIRE! " << res.steps << " " << res.min_gap << "\n";
    *out_porosity = 1.0 - (double)kept * pV / Vext;
    std::cerr << "Result density:porosity " << 1.0 - *out_porosity << " porosity: " << *out_porosity
              << "\n";
    return write_sample(path, s, box.hidden);
}

// CalculateUParameters / CalculateBParameters inputs: size_real and
// center_real of the loaded sample (md_createcylindersample.cu:448-465).
rf_platen_cfg platen_of(const Sample& s, const Loading& L) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < s.x.size(); ++i) {
        const double p[3] = {s.x[i], s.y[i], s.z[i]};
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], p[d]);
            hi[d] = std::max(hi[d], p[d]);
        }
    }
    rf_platen_cfg cfg{};
    cfg.kind = L.kind;
    cfg.dnsave = L.dnsave;
    cfg.v_load = L.v_mps / kVelocity;
    cfg.min_stress = L.min_stress_pa / kStress;
    cfg.max_strain = L.max_strain;
    cfg.cond_fracture = L.cond;
    for (int d = 0; d < 3; ++d) {
        cfg.size[d] = hi[d] - lo[d] + 2.0 * kPr;
        cfg.center[d] = 0.5 * (hi[d] + lo[d]);
    }
    cfg.r_ref = kPr;
    // wall moduli: 2 hm_E uniaxial (md_calculationuniaxialcompression.cu:278),
    // hm_E Brazilian (md_calculationbraziltest.cu:198)
    cfg.wall_modulus_factor = L.kind == RF_TEST_UNIAXIAL ? 2.0 : 1.0;
    return cfg;
}

bool write_curve(rf_ctx* ctx, int64_t n_curve, const std::string& path) {
    int64_t m = n_curve;
    std::vector<double> curve(3 * (size_t)std::max<int64_t>(m, 1));
    if (!ok(rf_get_curve(ctx, curve.data(), m, &m), "rf_get_curve")) return false;
    std::ofstream f(path);
    f << "EpsZ StressZ_bottom StressZ_top\n";
    for (int64_t i = 0; i < m; ++i)
        f << curve[3 * i] << " " << curve[3 * i + 1] * kStress << " " << curve[3 * i + 2] * kStress << "\n";
    return true;
}

// h_CalculationUniaxialCompression / h_CalculationBrazilTest of one sample
// for each material.  One material runs alone (its chunks replayed as
// captured graphs); several run as one batch, one step launch for all.
// The curve of run `curve_of` is written to curve_path.
bool run_tests(const Sample& s, const std::vector<rf_material>& mats, const rf_box& box,
               const Loading& L, int device, std::vector<rf_test_state>& out,
               const std::string& curve_path, size_t curve_of) {
    const size_t n = s.x.size();
    std::vector<double> r(n, kPr);
    rf_sample smp{(int64_t)n, s.x.data(), s.y.data(), s.z.data(), r.data()};
    rf_options opt{};
    opt.slots = 32;  // IonP (mp_determineconjgradient.cu:17)
    opt.cand_cap = 32;
    opt.skin = -1.0;  // library default
    opt.device = device;
    opt.block = 256;
    const rf_platen_cfg cfg = platen_of(s, L);
    std::vector<rf_ctx*> ctxs;
    bool good = true;
    for (const rf_material& m : mats) {
        rf_ctx* c = nullptr;
        good = ok(rf_create(&smp, &m, &box, &opt, &c), "rf_create");
        if (!good) break;
        ctxs.push_back(c);
        good = ok(rf_build_bonds(c), "rf_build_bonds") && ok(rf_set_test(c, &cfg), "rf_set_test");
        if (!good) break;
    }
    out.assign(mats.size(), rf_test_state{});
    if (good && ctxs.size() == 1) {
        rf_ctx* ctx = ctxs[0];
        rf_test_state ts{};
        while (good) {  // progress every 20000 steps, then rf_run finishes
            int64_t done = 0;
            good = ok(rf_step(ctx, 20000, &done), "rf_step") &&
                   ok(rf_get_test_state(ctx, &ts), "rf_get_test_state");
            if (!good || ts.done || done == 0) break;
            std::cerr << "Steps " << ts.step << " strain " << ts.strain << " stress "
                      << ts.stress * kStress << "\n";
        }
        good = good && ok(rf_run(ctx, &out[0]), "rf_run");
    } else if (good && !ctxs.empty()) {
        rf_batch* b = nullptr;
        good = ok(rf_batch_create(ctxs.data(), (int32_t)ctxs.size(), &b), "rf_batch_create");
        std::vector<int64_t> done(ctxs.size(), 0), total(ctxs.size(), 0);
        while (good) {
            good = ok(rf_batch_step(b, 20000, done.data()), "rf_batch_step");
            int64_t moved = 0;
            for (size_t k = 0; k < done.size(); ++k) {
                total[k] += done[k];
                moved += done[k];
            }
            if (!good || moved == 0) break;
            std::cerr << "Steps";
            for (int64_t t : total) std::cerr << " " << t;
            std::cerr << "\n";
        }
        good = good && ok(rf_batch_run(b, out.data()), "rf_batch_run");
        if (b) rf_batch_destroy(b);
    }
    if (good && !curve_path.empty() && curve_of < ctxs.size())
        good = write_curve(ctxs[curve_of], out[curve_of].n_curve, curve_path);
    for (rf_ctx* c : ctxs) rf_destroy(c);
    return good;
}

int calculate_one_sample(const std::string& ms, bool create, bool U, bool B, bool fusc, bool fbsc,
                         int device) {
    std::cerr << "Start CalculateOneSample\n";
    double mp[6];
    size_t start = ms.find('(');
    for (int i = 0; i < 6; ++i) {
        if (start == std::string::npos) {
            std::cerr << "bad material string " << ms << "\n";
            return 1;
        }
        const size_t fin = ms.find_first_of(",)", start + 1);
        mp[i] = std::stod(ms.substr(start + 1, fin - start - 1));
        start = fin;
    }
    const auto t0 = std::chrono::system_clock::now();
    // mp_calculateonesample.cu:68-73: SI -> reduced units
    const double P[5] = {mp[0] / kStress, mp[1] / kStress, mp[2] / kStress, mp[3] / kLength,
                         mp[4] / kLength};
    const rf_material mat = material_of(P, 3.2);
    rf_box box[2];
    for (int k = 0; k < 2; ++k)
        if (!box_of(k, mat.acut, &box[k])) return 1;
    double mps5 = mp[5];
    double RF[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double UPorosity = mps5;
    const double gap_tol = -0.0001;  // MaximumParticleIntersection, mp_calculateonesample.cu:106
    if ((U && create) || fusc) {
        if (!create_sample(0, mps5, gap_tol, box[0], device, sample_path(0, "_0"), &UPorosity)) return 1;
    }
    std::vector<rf_test_state> ts;
    Sample s;
    if (U) {
        if (!read_sample(sample_path(0, "_0"), s) ||
            !run_tests(s, {mat}, box[0], kOneU, device, ts, "./result/Uniaxial_0.txt", 0))
            return 1;
        RF[0] = ts[0].peak_strain;
        RF[1] = ts[0].peak_stress;
        RF[2] = ts[0].linearity_deviation;
        RF[3] = ts[0].destruction;
        std::cerr << "RU " << RF[0] << " " << RF[1] * kStress << "\n";
    } else {
        RF[0] = 1e-2;
        RF[1] = 1e6 / kStress;
        RF[2] = 1.0;
        RF[3] = 1.0;
    }
    mps5 = UPorosity > mps5 ? UPorosity : mps5;
    double BPorosity = mps5;
    if ((B && create) || fbsc) {
        if (!create_sample(1, mps5, gap_tol, box[1], device, sample_path(1, "_0"), &BPorosity)) return 1;
    }
    mps5 = UPorosity > BPorosity ? UPorosity : BPorosity;
    if (B) {
        if (!read_sample(sample_path(1, "_0"), s) ||
            !run_tests(s, {mat}, box[1], kOneB, device, ts, "./result/Brazil_0.txt", 0))
            return 1;
        RF[4] = ts[0].peak_strain;
        RF[5] = ts[0].peak_stress;
        RF[7] = ts[0].destruction;
        std::cerr << "RB " << RF[5] * kStress << "\n";
    } else {
        RF[4] = 1e-2;
        RF[5] = 1e6 / kStress;
        RF[7] = 1.0;
    }
    RF[6] = 0;
    std::cout << "ResultCalculateOneSample[" << P[0] * kStress << "," << P[1] * kStress << ","
              << P[2] * kStress << "," << P[3] * kLength << "," << P[4] * kLength << "," << mps5 << ","
              << RF[0] << "," << RF[1] * kStress << "," << RF[2] << "," << RF[3] << "," << RF[4] << ","
              << RF[5] * kStress << "," << RF[6] << "," << RF[7] << "]\n";
    const auto t1 = std::chrono::system_clock::now();
    const std::time_t s0 = std::chrono::system_clock::to_time_t(t0);
    const std::time_t s1 = std::chrono::system_clock::to_time_t(t1);
    std::cout << "Time " << std::ctime(&s0) << " | " << std::ctime(&s1) << "\n";
    std::cerr << "Fin\n";
    return 0;
}

// DetermineConjGradient (mp_determineconjgradient.cu:146-368): 50 steps of
// the reference's conjugate-gradient calibration of (E, sigma_c, tau_c, b_r,
// gap) against the granite goals, variation type 7 (strengths and bond
// radius), uniaxial test only.  Each step needs the base run and one run per
// varied parameter; they run on the GPU as one batch.  Records every step
// in ./result/CalculationResults.dat (mp_task.h:313-321).
int determine_conj_gradient(int device) {
    std::cerr << "Start DetermineConjGradient\n";
    rfcg::Task t;
    t.delta_factor = 0.1;
    const double goal[4] = {0.00480002, 104.0497482e6 / kStress, 10.5e6 / kStress, 1.0};
    const double mult[4] = {50, 100, 1, 10000};
    for (int j = 0; j < 4; ++j) {
        t.goal[j] = goal[j];
        t.weight_mult[j] = mult[j];
    }
    // the last of the start points assigned in the reference (:281-285)
    const double P0[5] = {0.45 * 1.71066e10 / kStress, 1.0 * 6.18389e8 / kStress,
                          1.0 * 5.03332e8 / kStress, 4 * 0.000409345 / kLength, 0.00046797 / kLength};
    for (int i = 0; i < 5; ++i) t.params[i] = P0[i];
    t.u_test = true;
    t.b_test = false;
    t.variation_type = 7;
    int nsteps = 50;
    if (const char* e = std::getenv("RF_CG_STEPS")) nsteps = std::atoi(e);
    const char* reuse_env = std::getenv("RF_CG_REUSE_SAMPLES");
    const bool reuse = reuse_env && reuse_env[0] == '1';
    const double porosity[2] = {0.375, 0.37};
    const double gap_tol = -0.001;  // MaximumParticleIntersection of this driver
    const Loading* L[2] = {&kCgU, &kCgB};
    const bool test_on[2] = {t.u_test, t.b_test};
    // samples are created with SetInitialParameters' acut (3.2 p_r); the
    // runs use SetVariationParameters' 2.9 p_r in the same space box
    rf_box cbox[2], rbox[2];
    Sample smp[2];
    for (int k = 0; k < 2; ++k) {
        if (!box_of(k, 3.2 * kPr, &cbox[k])) return 1;
        rbox[k] = regrid(cbox[k], 2.0 * 2.9 * kPr);
        if (!test_on[k]) continue;
        const std::string path = sample_path(k, "");
        double got = 0.0;
        std::ifstream probe(path);
        if (!(reuse && probe.good()) &&
            !create_sample(k, porosity[k], gap_tol, cbox[k], device, path, &got))
            return 1;
        if (!read_sample(path, smp[k])) return 1;
    }
    t.step = 0;
    t.start();
    for (int step = 0; step < nsteps; ++step) {
        t.step = step;
        for (int k = 0; k < 2; ++k) {
            std::vector<rf_material> mats;
            std::vector<int> which;
            for (int run = 0; run <= rfcg::kParams; ++run) {
                if (!t.runs(run) || !test_on[k]) {  // CalculateStepCG's skipped runs
                    if (k == 0) {
                        t.R[run][0] = 0;
                        t.R[run][1] = 0;
                    } else {
                        t.R[run][2] = 1.0;
                    }
                    continue;
                }
                double P[5];
                t.point(run, P);
                mats.push_back(material_of(P, 2.9));
                which.push_back(run);
            }
            if (mats.empty()) continue;
            std::vector<rf_test_state> out;
            const std::string curve = k == 0 ? "./result/Uniaxial_0.txt" : "./result/Brazil_0.txt";
            if (!run_tests(smp[k], mats, rbox[k], *L[k], device, out, curve, mats.size() - 1)) return 1;
            for (size_t q = 0; q < which.size(); ++q) {
                const int run = which[q];
                if (k == 0) {
                    t.R[run][0] = out[q].peak_strain;
                    t.R[run][1] = out[q].peak_stress;
                    t.R[run][3] = out[q].linearity_deviation;
                    std::cerr << "RU " << out[q].peak_strain << " " << out[q].peak_stress * kStress << "\n";
                } else {
                    t.R[run][2] = out[q].peak_stress;
                    std::cerr << "RB " << out[q].peak_stress * kStress << "\n";
                }
            }
        }
        double base[5];
        for (int i = 0; i < 5; ++i) base[i] = t.params[i];
        t.update();
        if (!t.save("./result/CalculationResults.dat", base)) {
            std::cerr << "cannot write ./result/CalculationResults.dat\n";
            return 1;
        }
        std::cerr << "StepDG " << step << "\n";
    }
    std::cerr << "Fin\n";
    return 0;
}

// TestSeveralSamples (mp_testseveralsamples.cu:16-161): one uniaxial sample
// created (porosity 0.375) and loaded at 0.04 m/s; appends to
// ./result/CalculationSamplesResults.dat (mp_task.h:324-333).  The
// Brazilian columns hold 0: the reference's test is off and it prints
// uninitialised memory there.
int test_several_samples(int device) {
    std::cerr << "Start TestSeveralSamples\n";
    const double P[5] = {1.71066e10 / kStress, 6.18389e8 / kStress, 5.03332e8 / kStress,
                         0.000409345 / kLength, 0.00046797 / kLength};
    const rf_material mat = material_of(P, 3.2);
    const int nsamples = 1;
    const bool u_test = true, b_test = false;
    rf_box box[2];
    for (int k = 0; k < 2; ++k)
        if (!box_of(k, mat.acut, &box[k])) return 1;
    for (int i = 0; i < nsamples; ++i) {
        char part[32];
        std::snprintf(part, sizeof part, "_%d", i);
        double got = 0.0;
        if (u_test && !create_sample(0, 0.375, -0.0001, box[0], device, sample_path(0, part), &got))
            return 1;
        if (b_test && !create_sample(1, 0.37, -0.0001, box[1], device, sample_path(1, part), &got))
            return 1;
    }
    std::cerr << "TestSeveralSamples(): Samples Created\n";
    for (int i = 0; i < nsamples; ++i) {
        char part[32], curve[64];
        std::snprintf(part, sizeof part, "_%d", i);
        double R[4] = {0, 0, 0, 0};
        Sample s;
        std::vector<rf_test_state> out;
        if (u_test) {
            std::snprintf(curve, sizeof curve, "./result/Uniaxial_%d.txt", i);
            if (!read_sample(sample_path(0, part), s) ||
                !run_tests(s, {mat}, box[0], kTssU, device, out, curve, 0))
                return 1;
            R[0] = out[0].peak_strain;
            R[1] = out[0].peak_stress;
            std::cerr << "RU " << R[0] << " " << R[1] * kStress << "\n";
        }
        if (b_test) {
            std::snprintf(curve, sizeof curve, "./result/Brazil_%d.txt", i);
            if (!read_sample(sample_path(1, part), s) ||
                !run_tests(s, {mat}, box[1], kTssB, device, out, curve, 0))
                return 1;
            R[2] = out[0].peak_strain;
            R[3] = out[0].peak_stress;
            std::cerr << "RB " << R[3] * kStress << "\n";
        }
        std::ofstream f("./result/CalculationSamplesResults.dat", std::ios::app);
        if (i == 0)
            f << "Sample UniaxialFractureStrain UniaxialFractureStress BrasilFractureStrain "
                 "BrasilFractureStress\n";
        f << i << " " << R[0] << " " << R[1] * kStress << " " << R[2] << " " << R[3] * kStress << "\n";
    }
    std::cerr << "Fin\n";
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    bool flags[9] = {false, false, false, false, false, false, false, false, false};
    std::string material, param, index;
    int count = 0, device = 0;
    size_t shift = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    std::cout << "There are " << count << " GPUs" << std::endl;
    for (int i = 1; i < argc; ++i) {
        param = argv[i];
        if (param.compare(0, 1, "G") == 0) {
            const int d = argv[i][1] - '0';
            if (d >= 0 && d < count) {
                std::cout << "Correct GPU request! Requested: " << d << "\n";
                device = d;
                flags[0] = true;
            } else {
                std::cerr << "Incorrect GPU request! Requested:" << d << "\n";
                return 1;
            }
        }
        if (param.compare(0, 1, "M") == 0) {
            material = param;
            flags[1] = true;
            std::cerr << "Set Calculate One Sample\n";
        }
        if (param.size() >= 2 && param.compare(0, 2, "CS") == 0) {
            flags[2] = true;
            shift = 2;
        }
        if (param.size() >= shift + 1 && param.compare(shift, 1, "U") == 0) {
            flags[5] = true;
            ++shift;
        }
        if (param.size() >= shift + 1 && param.compare(shift, 1, "B") == 0) {
            flags[6] = true;
            ++shift;
        }
        if (param.compare(0, 3, "DCG") == 0) {
            flags[3] = true;
            std::cerr << "Set DetermineConjGradient\n";
        }
        if (param.compare(0, 3, "TSS") == 0) {
            flags[4] = true;
            std::cerr << "Set TestSeveralSamples\n";
        }
        if (param.compare(0, 4, "FUSC") == 0) {
            flags[7] = true;
            std::cerr << "Set Force uniaxial sample creation\n";
        }
        if (param.compare(0, 4, "FBSC") == 0) {
            flags[8] = true;
            std::cerr << "Set Force brazil sample creation\n";
        }
        if (param.compare(0, 1, "I") == 0) {
            index = param.substr(1);
            std::cout << "INDEX " << index << "\n";
        }
    }
    if (!flags[0]) std::cerr << "Selected by default GPU 0\n";
    if (const char* dry = std::getenv("RF_CLI_DRY_RUN"); dry && dry[0] == '1') {
        std::cerr << "FLAGS";
        for (bool f : flags) std::cerr << " " << (f ? 1 : 0);
        std::cerr << " MATERIAL " << material << " DEVICE " << device << "\n";
        return 0;
    }
    mkdir("./result", 0755);
    int rc = 0;
    if (flags[1] && !rc)
        rc = calculate_one_sample(material, flags[2], flags[5], flags[6], flags[7], flags[8], device);
    if (flags[3] && !rc) rc = determine_conj_gradient(device);
    if (flags[4] && !rc) rc = test_several_samples(device);
    return rc;
}
