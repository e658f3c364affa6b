// pp_growth.h -- the linear growth tables on the device: Dlinear_ab's
// 4-point Lagrange look-up (src/Cosmology/Cosmology.jl:273-293), shared by
// the collapse-table generation and the per-peak bookkeeping.
#pragma once

#include <hip/hip_runtime.h>

namespace pp {

struct GrowthDev {
    int ntab;
    double amin, amax, D_inf, D_0, HD_inf, HD_0;
};

// the reference's HUNT bisection over an ascending table, 1-based result
__device__ __forceinline__ int hunt_asc(const double *xx, int n, double x)
{
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo > 1) {
        const int jm = (jhi + jlo) >> 1;
        if (x > xx[jm - 1]) jlo = jm;
        else jhi = jm;
    }
    return jlo;
}

// Dlinear_ab: D = D_a(a) a and dlnD/dlna at scale factor a; D_a is the
// interpolated table value (its third output).  Inside the table the
// 4-point Lagrange rule on rows kk-2 .. kk+1 (kk = HUNT index), the 2-point
// rule next to its ends (where the 4 rows would leave the table).
__device__ inline void dlinear(const double *at, const double *Dt, const double *Ht, const GrowthDev &g, double a,
                               double &D, double &HD, double &Da)
{
    if (a > g.amin && a < g.amax) {
        const int kk = hunt_asc(at, g.ntab, a);
        int nl, kl;
        if (kk < 2 || kk >= g.ntab) { nl = 2; kl = kk - 1 > 1 ? kk - 1 : 1; }
        else { nl = 4; kl = kk - 2; }
        if (kl < 1) { nl = 2; kl = 1; }
        const double *xa = at + kl - 1, *v1 = Dt + kl - 1, *v2 = Ht + kl - 1;
        if (nl == 4) {
            const double a1 = (a - xa[1]) / (xa[0] - xa[1]) * (a - xa[2]) / (xa[0] - xa[2]) * (a - xa[3]) / (xa[0] - xa[3]);
            const double a2 = (a - xa[0]) / (xa[1] - xa[0]) * (a - xa[2]) / (xa[1] - xa[2]) * (a - xa[3]) / (xa[1] - xa[3]);
            const double a3 = (a - xa[0]) / (xa[2] - xa[0]) * (a - xa[1]) / (xa[2] - xa[1]) * (a - xa[3]) / (xa[2] - xa[3]);
            const double a4 = (a - xa[0]) / (xa[3] - xa[0]) * (a - xa[1]) / (xa[3] - xa[1]) * (a - xa[2]) / (xa[3] - xa[2]);
            Da = a1 * v1[0] + a2 * v1[1] + a3 * v1[2] + a4 * v1[3];
            HD = a1 * v2[0] + a2 * v2[1] + a3 * v2[2] + a4 * v2[3];
        } else {
            const double a1 = (a - xa[1]) / (xa[0] - xa[1]);
            const double a2 = 1.0 - a1;
            Da = a1 * v1[0] + a2 * v1[1];
            HD = a1 * v2[0] + a2 * v2[1];
        }
    } else if (a >= g.amax) {
        Da = g.D_0;
        HD = g.HD_0;
    } else {
        Da = g.D_inf;
        HD = g.HD_inf;
    }
    D = Da * a;
}

}  // namespace pp
