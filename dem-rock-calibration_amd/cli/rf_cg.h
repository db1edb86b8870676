// Conjugate-gradient calibration step of the reference's mp_task_data
// (mp_task.h): weights and first variations (CalculateStartParameters
// :134-151), the forward-difference Jacobian, Fletcher-Reeves direction and
// Gauss-Newton step length (CalculateNewParameters2 :233-311), the lower
// bounds (SetMaterialContractions :110-133) and the per-step record
// (SaveResults :313-321).  Host code only; the runs the differences need are
// the caller's (the command line batches them on the GPU).
//
// Parameters, in the reference's reduced units: 0 bond modulus E, 1 normal
// strength sigma_c, 2 shear strength tau_c, 3 bond radius b_r, 4 bond gap.
// Results per run: 0 uniaxial peak strain, 1 uniaxial peak stress,
// 2 Brazilian peak stress, 3 uniaxial linearity deviation.
#pragma once

#include <cstdio>

namespace rfcg {

constexpr int kParams = 5;
constexpr int kResults = 4;
constexpr double kStress = 1e9;   // Pa per reduced stress unit
constexpr double kLength = 1e-3;  // m per mm

// Which parameters a variation type moves (ParameterVariationType,
// mp_task.h:146, 303): 0 strengths and E, 1 bond geometry, 2 all, 3 all but
// E, 4 shear strength and geometry, 5 E and geometry, 6 E, 7 strengths and
// bond radius.
constexpr int kVaried[8][kParams] = {{1, 1, 1, 0, 0}, {0, 0, 0, 1, 1}, {1, 1, 1, 1, 1},
                                     {0, 1, 1, 1, 1}, {0, 0, 1, 1, 1}, {1, 0, 0, 1, 1},
                                     {1, 0, 0, 0, 0}, {0, 1, 1, 1, 0}};

struct Task {
    int step = 0;
    int variation_type = 2;
    double delta_factor = 0.1;  // MaterialDeltaFactor
    bool u_test = true, b_test = false;
    double goal[kResults] = {0, 0, 0, 0};
    double weight_mult[kResults] = {0, 0, 0, 0};
    double weight[kResults] = {0, 0, 0, 0};
    double params[kParams] = {0, 0, 0, 0, 0};     // this step's base point
    double variation[kParams] = {0, 0, 0, 0, 0};  // run k+1 moves parameter k by this
    double R[kParams + 1][kResults] = {};         // results of run 0 (base) and runs 1..5
    double dR[kParams][kResults] = {};            // forward differences
    double dF[kParams] = {0, 0, 0, 0, 0};         // gradient of the functional
    double dd[kParams] = {0, 0, 0, 0, 0};         // search direction
    double D[kResults] = {0, 0, 0, 0};            // goal - result
    double F[kResults] = {0, 0, 0, 0};            // weighted squares
    double t = 0.0;                               // last step length

    // CalculateStartParameters: weights from the goals, first variations.
    void start() {
        for (int j = 0; j < kResults; ++j) weight[j] = weight_mult[j] / (goal[j] * goal[j]);
        if (!u_test) {  // as written in the reference: weight 1 is left alone
            weight[0] = 1;
            weight[3] = 1;
            goal[0] = 1;
            goal[1] = 1;
            goal[3] = 1;
        }
        if (!b_test) {
            weight[2] = 100;
            goal[2] = 1;
        }
        set_variation();
    }

    void set_variation() {
        for (int i = 0; i < kParams; ++i)
            variation[i] = kVaried[variation_type][i] * delta_factor * params[i];
    }

    // SetVariationParameters' skip rule: run k > 0 is not computed when its
    // parameter does not move.
    bool runs(int k) const { return k == 0 || !(variation[k - 1] < 1e-12); }

    // The material point of run k.
    void point(int k, double out[kParams]) const {
        for (int i = 0; i < kParams; ++i) out[i] = params[i] + (k > 0 && i == k - 1 ? variation[i] : 0.0);
    }

    // CalculateNewParameters2 (normalisation coefficients 1) followed by
    // SetMaterialContractions and the next step's variations.
    void update() {
        for (int i = 0; i < kParams; ++i)
            for (int j = 0; j < kResults; ++j)
                dR[i][j] = variation[i] > 1e-12 ? (R[i + 1][j] - R[0][j]) / variation[i] : 0.0;
        for (int j = 0; j < kResults; ++j) {
            D[j] = goal[j] - R[0][j];
            F[j] = weight[j] * D[j] * D[j];
        }
        // Fletcher-Reeves ratio of the new and old gradients; as in the
        // reference the direction adds beta times the old gradient (its
        // direction array is overwritten with the old gradient first)
        double g_old = 0.0, g_new = 0.0;
        for (int i = 0; i < kParams; ++i) {
            dd[i] = dF[i];
            g_old += dF[i] * dF[i];
            double g = 0.0;
            for (int j = 0; j < kResults; ++j) g += -2 * weight[j] * D[j] * dR[i][j];
            dF[i] = g;
            g_new += g * g;
        }
        const double beta = g_old > 1e-12 ? g_new / g_old : 0.0;
        for (int i = 0; i < kParams; ++i) dd[i] = dF[i] + beta * dd[i];
        double a0 = 0.0, a1 = 0.0;
        for (int j = 0; j < kResults; ++j)
            for (int i = 0; i < kParams; ++i) {
                a1 += weight[j] * D[j] * dR[i][j] * dd[i];
                for (int k = 0; k < kParams; ++k)
                    a0 += weight[j] * dR[i][j] * dd[i] * dR[k][j] * dd[k];
            }
        t = -a1 / a0;
        for (int i = 0; i < kParams; ++i) params[i] = params[i] - t * dd[i];
        contract();
        set_variation();
    }

    // SetMaterialContractions: lower bounds of every parameter.
    void contract() {
        if (params[0] < 1e5 / kStress) params[0] = 1e5 / kStress;
        if (params[1] < 1e5 / kStress) params[1] = 1e5 / kStress;
        if (params[2] < 1e8 / kStress) params[2] = 1e8 / kStress;
        if (params[3] < 1e-6 / kLength) params[3] = 1e-6 / kLength;
        if (params[4] < 1e-5 / kLength) params[4] = 1e-5 / kLength;
    }

    // SaveResults: appended after update(), with this step's base point.
    bool save(const char* path, const double base[kParams]) const {
        FILE* f = std::fopen(path, "a");
        if (!f) return false;
        std::fprintf(f, "Step: %d\n", step);
        std::fprintf(f, "Parameters (E, Ecrit, Gcrit, boundR, boundDistance): %g %g %g %g %g \n",
                     base[0] * kStress, base[1] * kStress, base[2] * kStress, base[3] * kLength,
                     base[4] * kLength);
        std::fprintf(f, "Results (UniaxialFractureStrain UniaxialFractureStress BrasilFractureStress): %g %g %g\n",
                     R[0][0], R[0][1] * kStress, R[0][2] * kStress);
        double sum = 0.0;
        std::fprintf(f, "DFunctional");
        for (int j = 0; j < kResults; ++j) {
            std::fprintf(f, " %g", F[j]);
            sum += F[j];
        }
        std::fprintf(f, " | %g\n", sum);
        std::fclose(f);
        return true;
    }
};

}  // namespace rfcg
