// Drives rfcg::Task (the command line's conjugate-gradient step) on a linear
// response model R = A P + b, for tests/test_cg.py.
// stdin: type delta steps u_test b_test, goal[4], mult[4], P0[5], A[4][5], b[4]
// argv[1]: the SaveResults file to append to.
#include <cstdio>

#include "rf_cg.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    rfcg::Task t;
    int steps = 0, u = 1, b = 0;
    double A[4][5], bb[4];
    if (std::scanf("%d %lf %d %d %d", &t.variation_type, &t.delta_factor, &steps, &u, &b) != 5) return 2;
    t.u_test = u != 0;
    t.b_test = b != 0;
    int got = 0;
    for (double& g : t.goal) got += std::scanf("%lf", &g);
    for (double& m : t.weight_mult) got += std::scanf("%lf", &m);
    for (double& p : t.params) got += std::scanf("%lf", &p);
    for (auto& row : A)
        for (double& a : row) got += std::scanf("%lf", &a);
    for (double& v : bb) got += std::scanf("%lf", &v);
    if (got != 4 + 4 + 5 + 20 + 4) return 2;
    t.step = 0;
    t.start();
    std::printf("weights %.17g %.17g %.17g %.17g\n", t.weight[0], t.weight[1], t.weight[2], t.weight[3]);
    for (int s = 0; s < steps; ++s) {
        t.step = s;
        for (int run = 0; run <= rfcg::kParams; ++run) {
            std::printf("run %d %d\n", run, t.runs(run) ? 1 : 0);
            double P[5];
            t.point(run, P);
            double R[4];
            for (int j = 0; j < 4; ++j) {
                R[j] = bb[j];
                for (int i = 0; i < 5; ++i) R[j] += A[j][i] * P[i];
            }
            // as the command line fills them: uniaxial results 0, 1, 3 and
            // Brazilian result 2, zeros / ones for runs not made
            if (t.runs(run) && t.u_test) {
                t.R[run][0] = R[0];
                t.R[run][1] = R[1];
                t.R[run][3] = R[3];
            } else {
                t.R[run][0] = 0;
                t.R[run][1] = 0;
            }
            t.R[run][2] = t.runs(run) && t.b_test ? R[2] : 1.0;
        }
        double base[5];
        for (int i = 0; i < 5; ++i) base[i] = t.params[i];
        t.update();
        if (!t.save(argv[1], base)) return 3;
        std::printf("step %d t %.17g\n", s, t.t);
        std::printf("params");
        for (double p : t.params) std::printf(" %.17g", p);
        std::printf("\ndF");
        for (double g : t.dF) std::printf(" %.17g", g);
        std::printf("\ndd");
        for (double d : t.dd) std::printf(" %.17g", d);
        std::printf("\nF");
        for (double f : t.F) std::printf(" %.17g", f);
        std::printf("\n");
    }
    return 0;
}
