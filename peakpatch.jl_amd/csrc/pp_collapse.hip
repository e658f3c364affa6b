// pp_collapse.hip -- collapse-table generation: the homogeneous-ellipsoid
// collapse ODE integrated for every (log10 F, e, p/e) point of the table.
//
// Reference: evolve_ellipse_full with the Fortran-compatible RK4 solver and
// the default EllipsoidParams (src/EllipsoidalCollapse/EllipsoidalCollapse.jl:
// 60-361, 527-624: iforce_strat 4, ivir_strat 2, fcoll (0.01, 0.171, 0.171),
// tfac 0.01, nstepmax 10000, zinit_fac 20, Carlson RD shape integrals), on
// the table grid of make_table (src/EllipsoidalCollapse/CollapseTable.jl:
// 33-55).  The table value is 1/a at the first axis's virialisation
// (1 + z_collapse), or -1.
//
// Design: one lane per table point, everything in Float64; the linear growth
// tables (DlinearTables columns, 2 000 rows) sit in LDS, so the four
// Dlinear_ab look-ups of each RK4 step are a binary search over LDS.  Points
// need from a few hundred to 10 000 steps; lanes of one wave take
// neighbouring points (similar step counts), and a wave ends when its last
// lane has virialised.
#include <cmath>
#include <vector>

#include "pp_common.h"
#include "pp_growth.h"

namespace pp {

constexpr int kGrowthMax = 2048;

struct EllState {
    double Rvac, Rcurv;
    int vir[3];
    double a_eq[3], a_3eq2, a_3eq1, Ha;
    double lam[3], F3;
};

// Carlson's RD by duplication, stopping at relative spread 0.05, with the
// fifth-order series (Numerical Recipes' rd, as the reference's
// _elliptic_rd, EllipsoidalCollapse.jl:63-107)
__device__ double carlson_rd(double x, double y, double z)
{
    double xt = x, yt = y, zt = z, sum = 0.0, fac = 1.0, dx, dy, dz, ave;
    for (;;) {
        const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        const double lam = sx * (sy + sz) + sy * sz;
        sum += fac / (sz * (zt + lam));
        fac *= 0.25;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        ave = 0.2 * (xt + yt + 3.0 * zt);
        dx = (ave - xt) / ave;
        dy = (ave - yt) / ave;
        dz = (ave - zt) / ave;
        if (fmax(fabs(dx), fmax(fabs(dy), fabs(dz))) <= 0.05) break;
    }
    const double C1 = 3.0 / 14.0, C2 = 1.0 / 6.0, C3 = 9.0 / 22.0, C4 = 3.0 / 26.0;
    const double C5 = 0.25 * C3, C6 = 1.5 * C4;
    const double ea = dx * dy, eb = dz * dz, ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
    return 3.0 * sum + fac * (1.0 + ed * (-C1 + C5 * ed - C6 * dz * ee) +
                              dz * (C2 * ee + dz * (-C3 * ec + dz * C4 * ea))) / (ave * sqrt(ave));
}

// shape integrals b_i of the ellipsoid with axes a1, a2, a3 (get_b_2,
// EllipsoidalCollapse.jl:113-146): spheroid closed forms when a3/a1 >= a2/a1
// (exactly 1/3 each within 0.1% of a sphere), Carlson RD otherwise
__device__ void shape_b(double a1, double a2, double a3, double b[3])
{
    const double third = 1.0 / 3.0;
    const double r2 = a2 / a1, r3 = a3 / a1;
    if (r3 >= r2) {
        if (r3 >= 0.999) {
            b[0] = b[1] = b[2] = third;
            return;
        }
        double b3;
        if (r3 <= 0.001) {
            const double e2 = 1.0 - r3 * r3, o2 = r3 * r3, o1 = 0.5 * r3 * r3, e = sqrt(e2);
            b3 = 0.5 * (o2 / e2 * (1.0 / o2 - log((1.0 + e) / o1) * 0.5 / e));
        } else {
            const double e2 = 1.0 - r3 * r3, e = sqrt(e2);
            b3 = 0.5 * ((1.0 - e2) / e2 * (1.0 / (1.0 - e2) - log((1.0 + e) / (1.0 - e)) * 0.5 / e));
        }
        b[0] = 1.0 - 2.0 * b3;
        b[1] = b3;
        b[2] = b3;
        return;
    }
    const double q = third * r2 * r3;
    const double b3 = q * carlson_rd(r2 * r2, 1.0, r3 * r3);
    const double b2 = q * carlson_rd(r3 * r3, 1.0, r2 * r2);
    b[0] = 1.0 - (b2 + b3);
    b[1] = b2;
    b[2] = b3;
}

__device__ __forceinline__ double hubble(double a, const EllState &s)
{
    return sqrt((1.0 + s.Rvac * a * a * a + s.Rcurv * a) / a);
}

// right-hand side at scale factor t of y = (a1, a2, a3, a1', a2', a3'):
// the axes virialise in the order 3, 2, 1 at fcoll_i a_b and freeze (their
// y entries are reset, as get_derivs! does to its argument)
__device__ void derivs(double t, double y[6], double dy[6], EllState &s, const double *at, const double *Dt,
                       const double *Ht, const GrowthDev &g)
{
    const double fc[3] = {0.01, 0.171, 0.171};
    const double a3b = t * t * t;
    const double Ha = hubble(t, s);
    s.Ha = Ha;
    const double Hi = 1.0 / Ha;
    if (s.vir[2] == 0 && y[2] <= fc[2] * t) { s.vir[2] = 1; s.a_eq[2] = fc[2] * t; }
    if (s.vir[1] == 0) s.a_3eq2 = fc[1] * t;
    if (s.vir[0] == 0) s.a_3eq1 = fc[0] * t;
    if (s.vir[2] == 1 && y[1] <= s.a_3eq2 && s.vir[1] == 0) { s.vir[1] = 1; s.a_eq[1] = s.a_3eq2; }
    if (s.vir[2] == 1 && y[0] <= s.a_3eq1 && s.vir[0] == 0) { s.vir[0] = 1; s.a_eq[0] = s.a_3eq1; }
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        dy[ax] = 0.0;
        dy[ax + 3] = 0.0;
        if (s.vir[ax] == 1) { y[ax] = s.a_eq[ax]; y[ax + 3] = 0.0; }
    }
    const double d1e = a3b / (y[0] * y[1] * y[2]);
    const double d0 = (0.5 / a3b - s.Rvac) * Hi;
    const double d1_int = 1.5 * (d1e - 1.0) / a3b * Hi;
    double D, HD, Da;
    dlinear(at, Dt, Ht, g, t, D, HD, Da);
    const double d1_ext = 1.5 / a3b * Hi * D;
    double b[3];
    shape_b(y[0], y[1], y[2], b);
#pragma unroll
    for (int ax = 0; ax < 3; ax++)
        if (s.vir[ax] == 0) {
            dy[ax] = y[ax + 3] * Hi;
            dy[ax + 3] = -y[ax] * (d1_int * b[ax] + d0 + d1_ext * (s.lam[ax] - s.F3));
        }
}

__global__ void __launch_bounds__(64) k_collapse_table(const double *__restrict__ g_at, const double *__restrict__ g_Dt,
                                                       const double *__restrict__ g_Ht, GrowthDev g, double Om,
                                                       double OL, int nx, int ny, int nz, double X1, double dX,
                                                       double Y1, double dY, double Z1, double dZ,
                                                       float *__restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double at[kGrowthMax], Dt[kGrowthMax], Ht[kGrowthMax];
    for (int q = threadIdx.x; q < g.ntab; q += blockDim.x) {
        at[q] = g_at[q];
        Dt[q] = g_Dt[q];
        Ht[q] = g_Ht[q];
    }
    __syncthreads();
    const long long npts = (long long)nx * ny * nz;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long long)nx * ny));
    const double x = X1 + i * dX, e = Y1 + j * dY, poe = Z1 + k * dZ;
    const double F = pow(10.0, x), pv = poe * e;
    EllState s;
    s.Rvac = OL / Om;
    s.Rcurv = (1.0 - Om - OL) / Om;
    s.vir[0] = s.vir[1] = s.vir[2] = 0;
    s.a_eq[0] = s.a_eq[1] = s.a_eq[2] = 0.0;
    s.a_3eq2 = s.a_3eq1 = 0.0;
    s.lam[2] = F / 3.0 * (1.0 + 3.0 * e + pv);
    s.lam[1] = F / 3.0 * (1.0 - 2.0 * pv);
    s.lam[0] = F / 3.0 * (1.0 - 3.0 * e + pv);
    s.F3 = F / 3.0;
    const double ab = 1.0 / (20.0 * F);
    double t = ab, D, HD, Da;
    dlinear(at, Dt, Ht, g, ab, D, HD, Da);
    s.Ha = hubble(ab, s);
    double y[6], dy[6], yt[6], dyt[6], dym[6];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        y[ax] = ab * (1.0 - D * s.lam[ax]);
        y[ax + 3] = s.Ha * (1.0 - (1.0 + HD) * D * s.lam[ax]);
    }
    double zvir = -1.0;
    for (int step = 1;; step++) {
        const double h = 0.01 * y[2] * sqrt(y[2]) * s.Ha;
        derivs(t, y, dy, s, at, Dt, Ht, g);
        // the reference's RK4 (rk4_homel!): derivatives at t, t + h/2 twice, t + h
        const double hh = h * 0.5, h6 = h / 6.0, th = t + hh;
#pragma unroll
        for (int q = 0; q < 6; q++) yt[q] = y[q] + hh * dy[q];
        derivs(th, yt, dyt, s, at, Dt, Ht, g);
#pragma unroll
        for (int q = 0; q < 6; q++) yt[q] = y[q] + hh * dyt[q];
        derivs(th, yt, dym, s, at, Dt, Ht, g);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            yt[q] = y[q] + h * dym[q];
            dym[q] = dyt[q] + dym[q];
        }
        derivs(t + h, yt, dyt, s, at, Dt, Ht, g);
#pragma unroll
        for (int q = 0; q < 6; q++) y[q] = y[q] + h6 * (dy[q] + dyt[q] + 2.0 * dym[q]);
        t += h;
        if (s.vir[0] == 1) {
            zvir = 1.0 / t;
            break;
        }
        if (step >= 10000 || t >= 2.0) break;
    }
    out[p] = (float)zvir;
}

}  // namespace pp

using namespace pp;

extern "C" int pp_make_collapse_table(pp_ctx *ctx, double Om, double OL, const pp_growth *growth, int nx, int ny,
                                      int nz, double X1, double X2, double Y1, double Y2, double Z1, double Z2,
                                      float *table_out, int install)
{
    PP_ARG(ctx && growth && growth->atab && growth->D_a && growth->dlnD_dlna, "NULL argument");
    PP_ARG(growth->ntab >= 2 && growth->ntab <= kGrowthMax, "growth table rows %d outside [2, %d]", growth->ntab,
           kGrowthMax);
    PP_ARG(nx >= 2 && ny >= 2 && nz >= 2, "collapse table needs >= 2 points per axis (%d,%d,%d)", nx, ny, nz);
    PP_ARG(X2 > X1 && Y2 > Y1 && Z2 > Z1, "collapse table bounds must be increasing");
    PP_ARG(Om > 0, "Omega_m must be positive");
    PP_ARG(table_out || install, "nothing to do: no output array and install = 0");
    PP_CHECK(ensure_device(ctx));
    const int nt = growth->ntab;
    for (int q = 1; q < nt; q++)
        PP_ARG(growth->atab[q] > growth->atab[q - 1], "growth table scale factors must increase");
    const size_t npts = (size_t)nx * ny * nz, tb = (size_t)nt * sizeof(double);
    PP_CHECK(ctx->d_misc.grow(3 * tb + npts * sizeof(float)));
    double *d_at = ctx->d_misc.as<double>(), *d_Dt = d_at + nt, *d_Ht = d_Dt + nt;
    float *d_out = (float *)(d_Ht + nt);
    PP_HIP(hipMemcpyAsync(d_at, growth->atab, tb, hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(d_Dt, growth->D_a, tb, hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(d_Ht, growth->dlnD_dlna, tb, hipMemcpyHostToDevice, ctx->stream));
    GrowthDev g;
    g.ntab = nt;
    g.amin = growth->atab[0];
    g.amax = growth->atab[nt - 1];
    g.D_inf = growth->D_a[0];
    g.D_0 = growth->D_a[nt - 1];
    g.HD_inf = growth->dlnD_dlna[0];
    g.HD_0 = growth->dlnD_dlna[nt - 1];
    const double dX = (X2 - X1) / (nx - 1), dY = (Y2 - Y1) / (ny - 1), dZ = (Z2 - Z1) / (nz - 1);
    const unsigned blocks = (unsigned)((npts + 63) / 64);
    hipLaunchKernelGGL(k_collapse_table, dim3(blocks), dim3(64), 0, ctx->stream, (const double *)d_at,
                       (const double *)d_Dt, (const double *)d_Ht, g, Om, OL, nx, ny, nz, X1, dX, Y1, dY, Z1, dZ,
                       d_out);
    PP_HIP(hipGetLastError());
    if (install) {
        PP_CHECK(ctx->d_table.grow(npts * sizeof(float)));
        PP_HIP(hipMemcpyAsync(ctx->d_table.ptr, d_out, npts * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (table_out)
        PP_HIP(hipMemcpyAsync(table_out, d_out, npts * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    if (install) {
        ctx->tnx = nx; ctx->tny = ny; ctx->tnz = nz;
        ctx->tX1 = X1; ctx->tX2 = X2; ctx->tY1 = Y1; ctx->tY2 = Y2; ctx->tZ1 = Z1; ctx->tZ2 = Z2;
        ctx->have_table = true;
    }
    return PP_OK;
}
