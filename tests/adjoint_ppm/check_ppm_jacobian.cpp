// Host check of the adjoint's closed-form PPM Jacobian (para_jacobian, the
// limiter's active branches in closed form) against its forward-mode form over
// all five inputs (para_jacobian_dual): equal values and tangents on random,
// near-linear, spiked and tied stencils.  Exit status 0 when they agree.
#include <cstdio>
#include <cmath>
#include <random>
#include "hamr_adjoint.h"
using namespace hamr;
int main()
{
    std::mt19937_64 g(7);
    std::normal_distribution<double> nd(0., 1.);
    std::uniform_int_distribution<int> kind(0, 3);
    double worst = 0.;
    long nz = 0, branches[4] = {0, 0, 0, 0};
    for (long it = 0; it < 2000000; it++) {
        double x[5];
        const int kk = kind(g);
        for (int k = 0; k < 5; k++) {
            if (kk == 0) x[k] = nd(g);
            else if (kk == 1) x[k] = k + 0.3 * nd(g);           // near-linear
            else if (kk == 2) x[k] = (k == 2 ? 3. : 0.) + 0.1 * nd(g);  // spike
            else x[k] = std::floor(2 * nd(g));                   // ties
        }
        double l1, r1, l2, r2, Jl1[5], Jr1[5], Jl2[5], Jr2[5];
        para_jacobian_dual(x, &l1, &r1, Jl1, Jr1);
        para_jacobian(x, &l2, &r2, Jl2, Jr2);
        if (l1 != l2 || r1 != r2) { printf("value mismatch at %ld\n", it); return 1; }
        for (int k = 0; k < 5; k++) {
            const double d = std::fmax(std::fabs(Jl1[k] - Jl2[k]), std::fabs(Jr1[k] - Jr2[k]));
            worst = std::fmax(worst, d);
            nz += (Jl1[k] != 0.) + (Jr1[k] != 0.);
        }
        branches[kk]++;
    }
    printf("max |J_closed - J_dual| = %.3e over 2e6 inputs (%ld nonzero entries)\n", worst, nz);
    return worst < 1e-14 ? 0 : 2;
}
