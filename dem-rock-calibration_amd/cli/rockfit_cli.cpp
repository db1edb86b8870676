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
// and CalculateOneSample (mp_calculateonesample.cu:17-250): material from
// M(...) restated from SetInitialParameters (mp_determineconjgradient.cu:15-47),
// samples ./result/Sample_<k>_0.txt (md_createcylindersample.cu:144-182,
// 441-466), curves ./result/Uniaxial_0.txt and ./result/Brazil_0.txt
// (md_results.cu:168-180).  Sample creation (CS, FUSC, FBSC) places the
// spheres on a jittered lattice of the sample cylinder and relaxes them with
// FIRE on the GPU (rf_fire_relax); the reference's packing is time-seeded,
// this one is seeded by RF_SEED (default 20261203).
//
// RF_CLI_DRY_RUN=1 prints the parsed flags to stderr and stops before any
// computation (argument checks without a GPU).
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

#include "rockfit_dem.h"

namespace {

constexpr double kStress = 1e9;   // Pa per reduced stress unit
constexpr double kLength = 1e-3;  // m per mm
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kPr = 0.5;       // particle radius, mm (SetInitialParameters)

struct Geometry {
    double L[3];     // LSi - 2 p_r (centres), mm
    double space[3];  // SpaceSize, mm
};
// mp_calculateonesample.cu:91: uniaxial cylinder and Brazilian disc
const Geometry kGeom[2] = {{{25.0 - 2 * kPr, 25.0 - 2 * kPr, 50.0 - 2 * kPr}, {40.0, 40.0, 60.0}},
                           {{50.0 - 2 * kPr, 50.0 - 2 * kPr, 20.0 - 2 * kPr}, {65.0, 65.0, 30.0}}};

bool ok(rf_status s, const char* what) {
    if (s == RF_OK) return true;
    std::cerr << what << " failed: " << rf_last_error() << "\n";
    return false;
}

rf_material reference_material(const double mp[6]) {
    rf_material m{};
    m.E_bond = mp[0] / kStress;
    m.E_contact = m.E_bond;
    m.nu = 0.247;
    m.density = 2.610;
    m.restitution = 0.6;
    m.friction = 0.45;
    m.rolling_friction = 0.05;
    m.bond_radius = mp[3] / kLength;
    m.bond_radius_ratio = 0.0;
    m.bond_gap = mp[4] / kLength;  // b_d = 2 p_r + gap
    m.sigma_c = mp[1] / kStress;
    m.tau_c = mp[2] / kStress;
    m.acut = 3.2 * kPr;
    m.quirks = RF_QUIRKS_REFERENCE;
    return m;
}

bool load_sample(int k, std::vector<double>& x, std::vector<double>& y, std::vector<double>& z) {
    char name[256];
    std::snprintf(name, sizeof name, "./result/Sample_%d_0.txt", k);
    std::ifstream f(name);
    size_t n = 0;
    if (!(f >> n)) {
        std::cerr << "cannot read " << name << "\n";
        return false;
    }
    // LoadParticleDATAtxt (md_createcylindersample.cu:165-182): x * (1/length)
    const double lc = 1.0 / kLength;
    x.resize(n);
    y.resize(n);
    z.resize(n);
    std::string line;
    std::getline(f, line);
    for (size_t i = 0; i < n; ++i) {
        double a, b, c;
        if (!std::getline(f, line) || std::sscanf(line.c_str(), "%lf %lf %lf", &a, &b, &c) != 3) {
            std::cerr << name << ": short file\n";
            return false;
        }
        x[i] = a * lc;
        y[i] = b * lc;
        z[i] = c * lc;
    }
    return true;
}

// SaveParticleDATAtxt: count, then x y z in metres; removed particles
// (at the hidden point) are left out.
bool save_sample(int k, const std::vector<double>& x, const std::vector<double>& y,
                 const std::vector<double>& z, const double hidden[3]) {
    char name[256];
    std::snprintf(name, sizeof name, "./result/Sample_%d_0.txt", k);
    std::vector<size_t> keep;
    for (size_t i = 0; i < x.size(); ++i) {
        const double dx = x[i] - hidden[0], dy = y[i] - hidden[1], dz = z[i] - hidden[2];
        if (dx * dx + dy * dy + dz * dz >= 1e-10) keep.push_back(i);
    }
    FILE* f = std::fopen(name, "w");
    if (!f) {
        std::cerr << "cannot write " << name << "\n";
        return false;
    }
    std::fprintf(f, "%zu\n", keep.size());
    for (size_t i : keep)
        std::fprintf(f, "%.17g %.17g %.17g\n", x[i] * kLength, y[i] * kLength, z[i] * kLength);
    std::fclose(f);
    return true;
}

// Sample creation: PN = (1 - porosity) Vext / p_V spheres
// (mp_calculateonesample.cu:99, md_data_types.h:114) with uniform random
// centres in the centre cylinder (radius L.x/2, half height L.z/2, the
// reference's d_CylinderRestrictionZ bounds, md_createcylindersample.cu:304,
// 387), then grown to full size: FIRE relaxation (rf_fire_relax with the
// harmonic overlap force and absorbing container walls, under which FIRE
// converges on dense packings) at radii 0.6, 0.75, 0.85, 0.92, 0.97 and
// 1.0 p_r.  The last stage runs to MaximumParticleIntersection = -1e-4
// (mp_calculateonesample.cu:106) and, like d_ReduceParticleNumber in the
// reference loop (md_createcylindersample.cu:404-415), removes the worst
// overlapping particle at checks that still fail late in the run.  Returns
// the porosity of the kept spheres (md_createcylindersample.cu:423-424).
bool create_sample(int k, double porosity, const rf_box& box, int device, double* out_porosity) {
    const Geometry& g = kGeom[k];
    const double R0 = 0.5 * g.L[0], H0 = 0.5 * g.L[2];
    const double Vext = kPi * (R0 + kPr) * (R0 + kPr) * (g.L[2] + 2.0 * kPr);
    const double pV = 4.0 * (1.0 / 3.0) * kPi * kPr * kPr * kPr;
    const size_t n = (size_t)((1.0 - porosity) * Vext / pV);
    uint64_t seed = 20261203;
    if (const char* s = std::getenv("RF_SEED")) seed = std::strtoull(s, nullptr, 10);
    std::mt19937_64 rng(seed + 1000 * k);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> x(n), y(n), z(n), r(n);
    for (size_t i = 0; i < n; ++i) {
        double X, Y;
        do {
            X = R0 * u(rng);
            Y = R0 * u(rng);
        } while (X * X + Y * Y > R0 * R0);
        x[i] = box.center[0] + X;
        y[i] = box.center[1] + Y;
        z[i] = box.center[2] + H0 * u(rng);
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
        c.gap_tol = last ? -1e-4 : -1e-3;
        c.check_every = last ? 50 : 20;
        c.delete_after = last ? 10000 : -1;
        if (!ok(rf_fire_relax((int64_t)n, x.data(), y.data(), z.data(), r.data(), &c, nullptr, &res),
                "rf_fire_relax"))
            return false;
        std::cerr << "FIRE radius " << sc * kPr << " steps " << res.steps << " min gap " << res.min_gap
                  << " removed " << res.removed << " stop " << res.stop_reason << "\n";
    }
    const size_t kept = n - (size_t)res.removed;
    std::cerr << "FIN FIRE! " << res.steps << " " << res.min_gap << "\n";
    *out_porosity = 1.0 - (double)kept * pV / Vext;
    std::cerr << "Result density:porosity " << 1.0 - *out_porosity << " porosity: " << *out_porosity << "\n";
    return save_sample(k, x, y, z, box.hidden);
}

struct TestOut {
    double strain = 0.0, stress = 0.0, ld = 1.0, destruction = 1.0;
};

// h_CalculationUniaxialCompression / h_CalculationBrazilTest on a loaded
// sample, with CalculateUParameters / CalculateBParameters inputs of
// mp_calculateonesample.cu:120-124, 156-159, 202-205.
bool run_test(int k, const rf_material& mat, const rf_box& box, int device, TestOut* t) {
    std::vector<double> x, y, z;
    if (!load_sample(k, x, y, z)) return false;
    const size_t n = x.size();
    std::vector<double> r(n, kPr);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < n; ++i) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], p[d]);
            hi[d] = std::max(hi[d], p[d]);
        }
    }
    rf_platen_cfg cfg{};
    cfg.kind = k == 0 ? RF_TEST_UNIAXIAL : RF_TEST_BRAZILIAN;
    cfg.dnsave = k == 0 ? 100 : 1000;
    cfg.v_load = -5.0 * 0.02 / 1e3;  // m/s -> mm/us
    cfg.min_stress = (k == 0 ? 10e6 : 2.1e6) / kStress;
    cfg.max_strain = k == 0 ? 0.01 : 0.03;
    cfg.cond_fracture = 0.1;
    for (int d = 0; d < 3; ++d) {
        cfg.size[d] = hi[d] - lo[d] + 2.0 * kPr;
        cfg.center[d] = 0.5 * (hi[d] + lo[d]);
    }
    cfg.r_ref = kPr;
    cfg.wall_modulus_factor = k == 0 ? 2.0 : 1.0;
    rf_sample s{(int64_t)n, x.data(), y.data(), z.data(), r.data()};
    rf_options opt{};
    opt.slots = 32;  // IonP (mp_determineconjgradient.cu:17)
    opt.cand_cap = 32;
    opt.skin = -1.0;  // library default
    opt.device = device;
    opt.block = 256;
    rf_ctx* ctx = nullptr;
    if (!ok(rf_create(&s, &mat, &box, &opt, &ctx), "rf_create")) return false;
    bool good = ok(rf_build_bonds(ctx), "rf_build_bonds") && ok(rf_set_test(ctx, &cfg), "rf_set_test");
    rf_test_state ts{};
    // progress on stderr every 20000 steps, then rf_run ends the test and
    // post-processes the curve
    while (good) {
        int64_t done = 0;
        good = ok(rf_step(ctx, 20000, &done), "rf_step") && ok(rf_get_test_state(ctx, &ts), "rf_get_test_state");
        if (!good || ts.done || done == 0) break;
        std::cerr << "Steps " << ts.step << " strain " << ts.strain << " stress " << ts.stress * kStress
                  << "\n";
    }
    good = good && ok(rf_run(ctx, &ts), "rf_run");
    if (good) {
        t->strain = ts.peak_strain;
        t->stress = ts.peak_stress;
        t->ld = ts.linearity_deviation;
        t->destruction = ts.destruction;
        int64_t m = ts.n_curve;
        std::vector<double> curve(3 * (size_t)std::max<int64_t>(m, 1));
        good = ok(rf_get_curve(ctx, curve.data(), m, &m), "rf_get_curve");
        char name[256];
        std::snprintf(name, sizeof name, k == 0 ? "./result/Uniaxial_0.txt" : "./result/Brazil_0.txt");
        std::ofstream f(name);
        f << "EpsZ StressZ_bottom StressZ_top\n";
        for (int64_t i = 0; i < m; ++i)
            f << curve[3 * i] << " " << curve[3 * i + 1] * kStress << " " << curve[3 * i + 2] * kStress
              << "\n";
    }
    rf_destroy(ctx);
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
    const rf_material mat = reference_material(mp);
    rf_box box[2];
    for (int k = 0; k < 2; ++k)
        if (!ok(rf_box_reference(kGeom[k].space, 2.0 * mat.acut, &box[k]), "rf_box_reference"))
            return 1;
    double mps5 = mp[5];
    double RF[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double UPorosity = mps5;
    if ((U && create) || fusc) {
        if (!create_sample(0, mps5, box[0], device, &UPorosity)) return 1;
    }
    if (U) {
        TestOut t;
        if (!run_test(0, mat, box[0], device, &t)) return 1;
        RF[0] = t.strain; RF[1] = t.stress; RF[2] = t.ld; RF[3] = t.destruction;
        std::cerr << "RU " << t.strain << " " << t.stress * kStress << "\n";
    } else {
        RF[0] = 1e-2; RF[1] = 1e6 / kStress; RF[2] = 1.0; RF[3] = 1.0;
    }
    mps5 = UPorosity > mps5 ? UPorosity : mps5;
    double BPorosity = mps5;
    if ((B && create) || fbsc) {
        if (!create_sample(1, mps5, box[1], device, &BPorosity)) return 1;
    }
    mps5 = UPorosity > BPorosity ? UPorosity : BPorosity;
    if (B) {
        TestOut t;
        if (!run_test(1, mat, box[1], device, &t)) return 1;
        RF[4] = t.strain; RF[5] = t.stress; RF[7] = t.destruction;
        std::cerr << "RB " << t.stress * kStress << "\n";
    } else {
        RF[4] = 1e-2; RF[5] = 1e6 / kStress; RF[7] = 1.0;
    }
    RF[6] = 0;
    std::cout << "ResultCalculateOneSample[" << mat.E_bond * kStress << "," << mat.sigma_c * kStress
              << "," << mat.tau_c * kStress << "," << mat.bond_radius * kLength << ","
              << mat.bond_gap * kLength << "," << mps5 << "," << RF[0] << "," << RF[1] * kStress
              << "," << RF[2] << "," << RF[3] << "," << RF[4] << "," << RF[5] * kStress << ","
              << RF[6] << "," << RF[7] << "]\n";
    const auto t1 = std::chrono::system_clock::now();
    const std::time_t s0 = std::chrono::system_clock::to_time_t(t0);
    const std::time_t s1 = std::chrono::system_clock::to_time_t(t1);
    std::cout << "Time " << std::ctime(&s0) << " | " << std::ctime(&s1) << "\n";
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
    if (flags[1]) rc = calculate_one_sample(material, flags[2], flags[5], flags[6], flags[7], flags[8], device);
    if (flags[3] || flags[4])
        std::cerr << "DCG / TSS: run the calibration driver (rockfit.calibrate) or a batch "
                     "(rf_batch_*) instead\n";
    return rc;
}
