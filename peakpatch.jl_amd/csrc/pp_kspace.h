// pp_kspace.h -- half-spectrum helpers shared by the field kernels: the
// wave vector of an rfft index, a Float64 factor times a ComplexF32 mode
// rounded once, and the tabulated P(k) of load_pk
// (src/InitialConditions/PowerSpectrum.jl:11-23).
#pragma once

#include <hip/hip_runtime.h>

namespace pp {

// k of half-spectrum index (ix, iy, iz): FFTW.rfftfreq / fftfreq(n, n dk),
// i.e. index times (n dk) / n
struct KGrid {
    int n, nk;
    double mult;
};

__device__ __forceinline__ void kvec(const KGrid &g, long long e, int &ix, int &iy, int &iz, double &kx, double &ky,
                                     double &kz)
{
    ix = (int)(e % g.nk);
    const long long r = e / g.nk;
    iy = (int)(r % g.n);
    iz = (int)(r / g.n);
    const int fy = iy < (g.n + 1) / 2 ? iy : iy - g.n, fz = iz < (g.n + 1) / 2 ? iz : iz - g.n;
    kx = (double)ix * g.mult;
    ky = (double)fy * g.mult;
    kz = (double)fz * g.mult;
}

__device__ __forceinline__ float2 cmul(float2 v, double re_f, double im_f)
{
    // (re_f + i im_f) (v.x + i v.y) in Float64, rounded to Float32
    const double a = (double)v.x, b = (double)v.y;
    return make_float2((float)(re_f * a - im_f * b), (float)(re_f * b + im_f * a));
}

// load_pk: log P linear in log k between table rows, extrapolated linearly.
// The bracket is the last row lo in [0, np-2] with lk[lo] <= log k.  When the
// rows are evenly spaced in log k (inv_dlk > 0: 1 / spacing, lk0 = lk[0]),
// the row is first guessed from the spacing and then moved to that bracket;
// otherwise it is found by bisection.  Either way the same row results.
__device__ inline double pk_eval(const double *__restrict__ lk, const double *__restrict__ lp, int np, double k,
                                 double lk0 = 0.0, double inv_dlk = 0.0)
{
    const double x = log(k);
    int lo;
    if (inv_dlk > 0.0) {
        const double g = floor((x - lk0) * inv_dlk);
        lo = g < 0.0 ? 0 : (g > (double)(np - 2) ? np - 2 : (int)g);
        while (lo > 0 && x < lk[lo]) lo--;
        while (lo < np - 2 && x >= lk[lo + 1]) lo++;
    } else {
        lo = 0;
        int hi = np - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (x >= lk[mid]) lo = mid;
            else hi = mid;
        }
    }
    const double t = (x - lk[lo]) / (lk[lo + 1] - lk[lo]);
    return exp(lp[lo] + t * (lp[lo + 1] - lp[lo]));
}

// 1 / the row spacing of a log k table evenly spaced to 1e-9 of it, else 0
inline double pk_uniform_inv_spacing(const double *lk, int np)
{
    if (np < 3) return 0.0;
    const double d = (lk[np - 1] - lk[0]) / (np - 1);
    if (!(d > 0.0)) return 0.0;
    for (int q = 1; q < np; q++)
        if (fabs((lk[q] - lk[q - 1]) - d) > 1e-9 * d) return 0.0;
    return 1.0 / d;
}

}  // namespace pp
