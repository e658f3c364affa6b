/*
 * pp_oracle.c -- FP64 CPU restatement of PeakPatch.jl's halo-finding hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is built into
 * oracle/liboracle.so and may be loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and there only
 * as the checker (or as the timed CPU baseline).  The product library
 * (peakpatch.jl_amd/) never links, loads or calls it.
 *
 * Every function restates the reference literally, quirks included, and cites
 * the file:line it follows.  Indices inside analyse_peak are 1-based exactly
 * as in the Julia source so that the restatement can be read side by side.
 *
 * Parity pins (see tests/test_oracle_pins.py):
 *   hRinteg, atab4, precompute_shells, normalize_strain, get_evals,
 *   tophat/gaussian windows, find_peaks, collapse-table I/O + trilinear,
 *   analyse_peak (Fortran get_homel golden vector + synthetic collapse /
 *   no-collapse cases), Dlinear tables (D(1)=1, fsc_of_z(0)=1.686,
 *   Einstein-de Sitter closed form), chi_to_z round trip.
 *   Unpinned at the last bit (no reference-held vectors exist): the FP64
 *   last bits of analyse_peak on arbitrary fields, Interpolations.jl's exact
 *   summation order, QuadGK's chi(z) last digits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* hRinteg  (src/HaloFinder/RadialShell.jl:100-127)                          */
/* ------------------------------------------------------------------------ */
double orc_hRinteg(double u0, double u12)
{
    double x0 = fabs(u0);
    if (x0 >= 2.0) return 0.0;
    double one7 = 1.0 / 7.0;
    double result = 1.6 * u12 - 6.4 * one7;
    double x02 = x0 * x0;
    double x04 = x02 * x02;
    if (x0 >= 1.0) {
        double h1 = x04 * (2.0 - x0 * (2.4 - x0 * (1.0 - x0 * one7)));
        double h2 = x02 * (4.0 - x0 * (4.0 - x0 * (1.5 - x0 * 0.2)));
        result = result - h2 * u12 + h1;
    } else {
        result = result - 0.2 * (u12 - one7);
        if (x0 > 0.0) {
            double h1 = x04 * (1.0 - x02 * (1.0 - 3.0 * x0 * one7));
            double h2 = x02 * (2.0 - x02 * (1.5 - 0.6 * x0));
            result = result - h2 * u12 + h1;
        }
    }
    return result;
}

/* ------------------------------------------------------------------------ */
/* atab4  (src/HaloFinder/RadialShell.jl:133-155); akk[1..1101] -> [0..1100] */
/* ------------------------------------------------------------------------ */
void orc_atab4(double *akk)
{
    double ca = 3.0 / (2.0 * ORC_PI);
    double cb = 1.0 / (4.0 * ORC_PI);
    for (int i = 1; i <= 1100; i++) {
        double u = (i - 1) * 0.01;
        double ru = sqrt(u);
        double v;
        if (ru < 1.0) v = ca * (2.0 / 3.0 - u + 0.5 * u * ru);
        else if (ru <= 2.0) { double t = 2.0 - ru; v = cb * (t * t * t); }
        else v = 0.0;
        akk[i - 1] = v;
    }
    akk[1100] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* precompute_shells  (src/HaloFinder/RadialShell.jl:160-175)                */
/* Generation order dk, dj, di; stable sort by r2 (counting sort = stable). */
/* Returns the number of cells; writes up to cap entries.                   */
/* ------------------------------------------------------------------------ */
int64_t orc_precompute_shells(int rmax, int32_t *di, int32_t *dj, int32_t *dk,
                              int32_t *r2, int64_t cap)
{
    int r2max = rmax * rmax;
    int64_t *count = (int64_t *)calloc((size_t)r2max + 2, sizeof(int64_t));
    int64_t total = 0;
    for (int k = -rmax; k <= rmax; k++)
        for (int j = -rmax; j <= rmax; j++)
            for (int i = -rmax; i <= rmax; i++) {
                int q = i * i + j * j + k * k;
                if (q <= r2max) { count[q + 1]++; total++; }
            }
    if (di == NULL || total > cap) { free(count); return total; }
    for (int q = 1; q <= r2max + 1; q++) count[q] += count[q - 1];
    for (int k = -rmax; k <= rmax; k++)
        for (int j = -rmax; j <= rmax; j++)
            for (int i = -rmax; i <= rmax; i++) {
                int q = i * i + j * j + k * k;
                if (q <= r2max) {
                    int64_t p = count[q]++;
                    di[p] = i; dj[p] = j; dk[p] = k; r2[p] = q;
                }
            }
    free(count);
    return total;
}

/* ------------------------------------------------------------------------ */
/* normalize_strain  (src/HaloFinder/RadialShell.jl:201-212)                 */
/* ------------------------------------------------------------------------ */
void orc_normalize_strain(double *m /* 3x3 column-major */, double F)
{
    double tr = m[0] + m[4] + m[8];
    if (tr != 0.0) {
        double s = F / tr;
        for (int i = 0; i < 9; i++) m[i] *= s;
    }
}

/* ------------------------------------------------------------------------ */
/* get_evals  (src/HaloFinder/RadialShell.jl:217-222)                        */
/* LAPACK eigvals(Symmetric(mat)) restated as cyclic Jacobi on the upper    */
/* triangle, iterated to convergence; eigenvalues ascending.               */
/* ------------------------------------------------------------------------ */
void orc_get_evals(const double *m /* column-major, upper triangle used */, double *lam)
{
    double a[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            int rr = r < c ? r : c, cc = r < c ? c : r;   /* Symmetric(:U) */
            a[r][c] = m[rr + 3 * cc];
        }
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off == 0.0) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0.0 ? 1.0 : -1.0) /
                           (fabs(theta) + sqrt(theta * theta + 1.0));
                if (!isfinite(theta)) t = 0.0;
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                if (t == 0.0) { a[p][q] = a[q][p] = 0.0; continue; }
                for (int k = 0; k < 3; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                a[p][q] = a[q][p] = 0.0;
            }
    }
    double l0 = a[0][0], l1 = a[1][1], l2 = a[2][2], t;
    if (l0 > l1) { t = l0; l0 = l1; l1 = t; }
    if (l1 > l2) { t = l1; l1 = l2; l2 = t; }
    if (l0 > l1) { t = l0; l0 = l1; l1 = t; }
    lam[0] = l0; lam[1] = l1; lam[2] = l2;
}

/* ------------------------------------------------------------------------ */
/* Collapse table: read_homeltab layout (CollapseTable.jl:102-122) is read  */
/* by the Python side; the trilinear lookup with -1 fill                    */
/* (CollapseTable.jl:128-145, Interpolations.jl 0.15.1 BSpline(Linear())    */
/* scaled on range(X1,X2,length=Nx) + extrapolate(fill=-1)) is here.        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz;
    double X1, X2, Y1, Y2, Z1, Z2;
    const float *data; /* x fastest */
} orc_table;

static void lin_pos(double x, double lo, double hi, int n, int *i0, double *f)
{
    double step = (hi - lo) / (double)(n - 1);
    double p = (x - lo) / step + 1.0;          /* 1-based position */
    double fl = floor(p);
    if (p == (double)n) fl = fl - 1.0;         /* upper edge uses (n-1, n) */
    if (fl < 1.0) fl = 1.0;
    if (fl > (double)(n - 1)) fl = (double)(n - 1);
    *i0 = (int)fl;
    *f = p - fl;
}

double orc_interp(const orc_table *t, double x, double y, double z)
{
    if (!(x >= t->X1 && x <= t->X2 && y >= t->Y1 && y <= t->Y2 &&
          z >= t->Z1 && z <= t->Z2))
        return -1.0;
    int ix, iy, iz;
    double fx, fy, fz;
    lin_pos(x, t->X1, t->X2, t->nx, &ix, &fx);
    lin_pos(y, t->Y1, t->Y2, t->ny, &iy, &fy);
    lin_pos(z, t->Z1, t->Z2, t->nz, &iz, &fz);
    ix--; iy--; iz--;
    const float *d = t->data;
    long nx = t->nx, nxy = (long)t->nx * t->ny;
#define T(a, b, c) ((double)d[(ix + (a)) + (iy + (b)) * nx + (iz + (c)) * nxy])
    double c00 = T(0, 0, 0) * (1.0 - fx) + T(1, 0, 0) * fx;
    double c10 = T(0, 1, 0) * (1.0 - fx) + T(1, 1, 0) * fx;
    double c01 = T(0, 0, 1) * (1.0 - fx) + T(1, 0, 1) * fx;
    double c11 = T(0, 1, 1) * (1.0 - fx) + T(1, 1, 1) * fx;
#undef T
    double c0 = c00 * (1.0 - fy) + c10 * fy;
    double c1 = c01 * (1.0 - fy) + c11 * fy;
    return c0 * (1.0 - fz) + c1 * fz;
}

double orc_interp_c(const float *data, int nx, int ny, int nz, double X1, double X2,
                    double Y1, double Y2, double Z1, double Z2,
                    double x, double y, double z)
{
    orc_table t = {nx, ny, nz, X1, X2, Y1, Y2, Z1, Z2, data};
    return orc_interp(&t, x, y, z);
}

/* ------------------------------------------------------------------------ */
/* Windows  (src/HaloFinder/Filters.jl:18-36)                                */
/* ------------------------------------------------------------------------ */
double orc_tophat_window(double kR)
{
    if (kR == 0.0) return 1.0;
    return 3.0 * (sin(kR) - kR * cos(kR)) / (kR * kR * kR);
}
double orc_gaussian_window(double kR) { return exp(-(kR * kR) / 2.0); }
double orc_gaussian_window_fortran(double kR) { return orc_gaussian_window(kR / 2.0); }

/* k-space multiply of smooth_field (Filters.jl:79-106) without the FFT:
 * out[ix,iy,iz] = in * Float32(W(k*Rf)), k computed in Float64, k=0 left
 * unscaled, Nyquist kept.  in/out are interleaved complex64, shape
 * (n/2+1, n, n), x fastest.  in == out is allowed. */
void orc_window_multiply(const float *in, float *out, int n, double boxsize,
                         double Rf, int wsmooth)
{
    int nk = n / 2 + 1;
    double dk = 2.0 * ORC_PI / boxsize;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < n; iz++) {
        int fz = iz < (n + 1) / 2 ? iz : iz - n;          /* fftfreq(n, n) */
        double kz = dk * (double)fz;
        for (int iy = 0; iy < n; iy++) {
            int fy = iy < (n + 1) / 2 ? iy : iy - n;
            double ky = dk * (double)fy;
            for (int ix = 0; ix < nk; ix++) {
                double kx = dk * (double)ix;              /* rfftfreq(n, n) */
                size_t e = 2 * ((size_t)ix + (size_t)nk * ((size_t)iy + (size_t)n * iz));
                float re = in[e], im = in[e + 1];
                double k = sqrt(kx * kx + ky * ky + kz * kz);
                double fkR = k * Rf;
                if (fkR == 0.0) { out[e] = re; out[e + 1] = im; continue; }
                double w;
                if (wsmooth == 0) w = orc_gaussian_window_fortran(fkR);
                else if (wsmooth == 3) w = orc_gaussian_window_fortran(fkR) * (k * k);
                else w = orc_tophat_window(fkR);
                float wf = (float)w;
                out[e] = re * wf;
                out[e + 1] = im * wf;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* find_peaks  (src/HaloFinder/PeakFind.jl:15-78)                            */
/* Outputs 1-based i, j, k as the reference; returns the peak count.        */
/* ------------------------------------------------------------------------ */
int64_t orc_find_peaks(const float *delta, int8_t *mask, int n1, int n2, int n3,
                       int nbuff, float fcrit, int64_t max_peaks,
                       int32_t *oi, int32_t *oj, int32_t *ok, float *odelta)
{
    int64_t np = 0;
    size_t s1 = (size_t)n1, s12 = (size_t)n1 * n2;
#define D(i, j, k) delta[((size_t)(i) - 1) + ((size_t)(j) - 1) * s1 + ((size_t)(k) - 1) * s12]
#define M(i, j, k) mask[((size_t)(i) - 1) + ((size_t)(j) - 1) * s1 + ((size_t)(k) - 1) * s12]
    for (int k = nbuff + 1; k <= n3 - nbuff; k++)
        for (int j = nbuff + 1; j <= n2 - nbuff; j++)
            for (int i = nbuff + 1; i <= n1 - nbuff; i++) {
                if (M(i, j, k) == 1) continue;
                float ff = D(i, j, k);
                if (ff < fcrit) continue;
                int is_max = 1;
                for (int kk = -1; kk <= 1 && is_max; kk++)
                    for (int jj = -1; jj <= 1 && is_max; jj++)
                        for (int ii = -1; ii <= 1; ii++)
                            if (ff < D(i + ii, j + jj, k + kk)) { is_max = 0; break; }
                if (!is_max) continue;
                oi[np] = i; oj[np] = j; ok[np] = k; odelta[np] = ff;
                np++;
                M(i, j, k) = 1;
                if (np >= max_peaks) return np;
            }
#undef D
#undef M
    return np;
}

/* ------------------------------------------------------------------------ */
/* Cosmology  (src/Cosmology/Cosmology.jl:68-372)                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    int ntab;
    double atab[2000], D_atab[2000], dlnD[2000];
    double amintab, amaxtab, D_ainfty, D_a0, HD_Hainfty, HD_Ha0;
} orc_dlin;

static void deriv_dlin(const double *y, double *d, double a, double h, double Omnr,
                       double Omvac, double Omcurv, double OmX, double OmB)
{
    /* Cosmology.jl:123-146 with Omer = Omhdm = fhdmclus = Omnrd = 0, ldecay = 0 */
    double Omer = 0.0, Omerd = 0.0, Omnrd = 0.0, fhdmclus = 0.0, Omhdm = 0.0, taud = 1.0;
    double hub = Omer / a + Omnr + Omvac * (a * a * a) + Omcurv * a + Omerd / a + Omnrd;
    d[0] = 1.0 / (h * sqrt(hub / (a * a * a)));
    double fb = y[2];
    d[1] = fb;
    double q = Omer / a + Omnr + Omvac * (a * a * a) + Omcurv * a + Omerd / a + Omnrd;
    double qs = 0.5 / q;
    double q_val = qs * (2.0 * Omer / a + Omnr - 2.0 * Omvac * (a * a * a) + 2.0 * Omerd / a + Omnrd);
    double fnr = (OmX + OmB + fhdmclus * Omhdm) / Omnr;
    if (OmB != 0.0 && a <= 1.0e-3) fnr = (OmX + fhdmclus * Omhdm) / Omnr;
    double fer = 0.0;
    double qf = qs * (2.0 * Omer / a * fer + Omnr * fnr + Omnrd * fnr);
    d[2] = 3.0 * qf - fb * fb - (1.0 - q_val) * fb;
    d[3] = Omnrd * a * d[0] / taud;
    d[4] = d[0] / a;
}

/* Dlinear_tables (Cosmology.jl:148-270).  Om = total matter, OB, OL, h. */
void orc_dlinear_tables(orc_dlin *T, double Om, double OB, double OL, double h)
{
    double Omnr = Om, Omvac = OL, Omcurv = 1.0 - Om - OL, OmX = Om - OB;
    double amaxtab = 1.0 / 0.5, amintab_target = 1.0 / 1001.0, ainit = 1.0e-7;
    int nmax = 2000;
    double a = ainit, alogn = log(ainit), amlog = log(amaxtab);
    double hh = (amlog - alogn) / nmax;
    double t0 = 1.0 / h / sqrt(Omnr);
    double aeq = 0.0, yaeq = sqrt(a + aeq), yaeq0 = sqrt(aeq);
    double y[5];
    y[0] = t0 * (2.0 / 3.0 * (yaeq * yaeq * yaeq - yaeq0 * yaeq0 * yaeq0) - 2.0 * aeq * (yaeq - yaeq0));
    y[1] = log(a);
    double fbclus = 0.0, fhdmclus = 0.0, Omhdm = 0.0;
    y[2] = 0.25 * (sqrt(24.0 * (OmX + fbclus * OB + fhdmclus * Omhdm) / Omnr + 1.0) - 1.0);
    y[3] = 0.0;
    y[4] = t0 * 2.0 * (yaeq - yaeq0);
    int ip = 0, lnorm = 1, ip0 = 0;
    double anorm = 0.001, bnorm = 1.0;
    for (int i = 1; i <= nmax; i++) {
        double w1[5], w2[5], w3[5], w4[5], yt[5];
        deriv_dlin(y, w1, a, h, Omnr, Omvac, Omcurv, OmX, OB);
        double alogn_half = alogn + hh * 0.5, a_half = exp(alogn_half);
        for (int c = 0; c < 5; c++) yt[c] = y[c] + hh * w1[c] * 0.5;
        deriv_dlin(yt, w2, a_half, h, Omnr, Omvac, Omcurv, OmX, OB);
        for (int c = 0; c < 5; c++) yt[c] = y[c] + hh * w2[c] * 0.5;
        deriv_dlin(yt, w3, a_half, h, Omnr, Omvac, Omcurv, OmX, OB);
        double alogn_new = alogn + hh, a_new = exp(alogn_new);
        for (int c = 0; c < 5; c++) yt[c] = y[c] + hh * w3[c];
        deriv_dlin(yt, w4, a_new, h, Omnr, Omvac, Omcurv, OmX, OB);
        for (int c = 0; c < 5; c++) y[c] = y[c] + hh / 6.0 * (w1[c] + w4[c] + 2.0 * (w2[c] + w3[c]));
        alogn = alogn_new;
        a = a_new;
        if (lnorm == 1 && a >= anorm) { lnorm = 0; bnorm = exp(y[1]) / a; }
        if (a >= amintab_target && a <= amaxtab) {
            T->atab[ip] = a;
            T->D_atab[ip] = exp(y[1]) / a;
            T->dlnD[ip] = y[2];
            ip++;
            if (a >= 1.0 && ip0 == 0) ip0 = ip;               /* 1-based */
        }
    }
    int ntab = ip;
    double fac;
    if (ip0 > 0 && ip0 > 1) {
        double ainterp = (T->atab[ip0 - 1] - 1.0) / (T->atab[ip0 - 1] - T->atab[ip0 - 2]);
        double Dg = (T->D_atab[ip0 - 1] * (1.0 - ainterp) + T->D_atab[ip0 - 2] * ainterp) / bnorm;
        fac = 1.0 / (Dg * bnorm);
    } else fac = 1.0 / bnorm;
    for (int i = 0; i < ntab; i++) T->D_atab[i] *= fac;
    T->ntab = ntab;
    T->amintab = T->atab[0];
    T->amaxtab = T->atab[ntab - 1];
    T->D_ainfty = T->D_atab[0];
    T->D_a0 = T->D_atab[ntab - 1];
    T->HD_Hainfty = T->dlnD[0];
    T->HD_Ha0 = T->dlnD[ntab - 1];
}

static int hunt_asc(const double *xx, int n, double x)
{   /* _hunt (Cosmology.jl:105-119), 1-based result */
    int ascnd = xx[n - 1] > xx[0];
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo > 1) {
        int jm = (jhi + jlo) / 2;
        if ((x > xx[jm - 1]) == ascnd) jlo = jm; else jhi = jm;
    }
    return jlo;
}

/* Dlinear_ab (Cosmology.jl:273-293): returns Dlinear; *HD_Ha and *D_a optional */
double orc_dlinear_ab(const orc_dlin *T, double ai, double *HD_Ha, double *D_a)
{
    int ntab = T->ntab;
    double Da, HD;
    if (ai > T->amintab && ai < T->amaxtab) {
        int kk = hunt_asc(T->atab, ntab, ai);
        int nlint, kkl;
        if (kk < 2 || kk >= ntab) { nlint = 2; kkl = kk - 1 > 1 ? kk - 1 : 1; }
        else { nlint = 4; kkl = kk - 2; }
        /* kk == 2 makes the reference index atab[0] (a BoundsError at
         * z ~ 1000); use the two-point rule there instead of reading out of
         * bounds. */
        if (kkl < 1) { nlint = 2; kkl = 1; }
        const double *xa = &T->atab[kkl - 1], *v1 = &T->D_atab[kkl - 1], *v2 = &T->dlnD[kkl - 1];
        double x = ai;
        if (nlint == 4) {
            double a1 = (x - xa[1]) / (xa[0] - xa[1]) * (x - xa[2]) / (xa[0] - xa[2]) * (x - xa[3]) / (xa[0] - xa[3]);
            double a2 = (x - xa[0]) / (xa[1] - xa[0]) * (x - xa[2]) / (xa[1] - xa[2]) * (x - xa[3]) / (xa[1] - xa[3]);
            double a3 = (x - xa[0]) / (xa[2] - xa[0]) * (x - xa[1]) / (xa[2] - xa[1]) * (x - xa[3]) / (xa[2] - xa[3]);
            double a4 = (x - xa[0]) / (xa[3] - xa[0]) * (x - xa[1]) / (xa[3] - xa[1]) * (x - xa[2]) / (xa[3] - xa[2]);
            Da = a1 * v1[0] + a2 * v1[1] + a3 * v1[2] + a4 * v1[3];
            HD = a1 * v2[0] + a2 * v2[1] + a3 * v2[2] + a4 * v2[3];
        } else {
            double a1 = (x - xa[1]) / (xa[0] - xa[1]);
            double a2 = 1.0 - a1;
            Da = a1 * v1[0] + a2 * v1[1];
            HD = a1 * v2[0] + a2 * v2[1];
        }
    } else if (ai >= T->amaxtab) { Da = T->D_a0; HD = T->HD_Ha0; }
    else { Da = T->D_ainfty; HD = T->HD_Hainfty; }
    if (HD_Ha) *HD_Ha = HD;
    if (D_a) *D_a = Da;
    return Da * ai;
}

/* fsc_of_z (RadialShell.jl:191-196) */
double orc_fsc_of_z(const orc_dlin *T, double z)
{
    double a = 1.0 / (1.0 + z);
    return 1.686 * orc_dlinear_ab(T, a, NULL, NULL);
}

size_t orc_dlin_size(void) { return sizeof(orc_dlin); }

/* chi(z) (Cosmology.jl:21-24): (2.998e5/100) * h * int_0^z dz/E(z).
 * QuadGK is restated as adaptive Gauss-Kronrod (7-15) to 1e-13 relative. */
static double inv_E(double z, double Om, double OL)
{
    double zp = 1.0 + z;
    return 1.0 / sqrt(Om * zp * zp * zp + OL);
}
static const double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.0};
static const double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
static double gk15(double a, double b, double Om, double OL, double *err)
{
    double c = 0.5 * (a + b), hl = 0.5 * (b - a);
    double fc = inv_E(c, Om, OL);
    double rk = fc * gk_wk[7], rg = fc * gk_wg[3];
    for (int j = 0; j < 7; j++) {
        double dx = hl * gk_x[j];
        double f = inv_E(c - dx, Om, OL) + inv_E(c + dx, Om, OL);
        rk += gk_wk[j] * f;
        if (j & 1) rg += gk_wg[j / 2] * f;
    }
    *err = fabs((rk - rg) * hl);
    return rk * hl;
}
static double gk_adapt(double a, double b, double Om, double OL, int depth)
{
    double err, r = gk15(a, b, Om, OL, &err);
    if (depth > 30 || err <= 1e-14 * fabs(r)) return r;
    double m = 0.5 * (a + b);
    return gk_adapt(a, m, Om, OL, depth + 1) + gk_adapt(m, b, Om, OL, depth + 1);
}
double orc_chi(double z, double Om, double OL, double h)
{
    if (z == 0.0) return 0.0;
    return (2.998e5 / 100.0) * h * gk_adapt(0.0, z, Om, OL, 0);
}

/* chi_to_z (Cosmology.jl:344-358) on a table built as build_chi_to_z
 * (Cosmology.jl:333-337): zs = range(0, z_max, length=npts). */
double orc_chi_to_z(const double *zs, const double *chis, int npts, double z_max, double chi_val)
{
    if (chi_val <= chis[0]) return 0.0;
    if (chi_val >= chis[npts - 1]) return z_max;
    int lo = 1, hi = npts;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (chis[mid - 1] <= chi_val) lo = mid; else hi = mid;
    }
    double frac = (chi_val - chis[lo - 1]) / (chis[hi - 1] - chis[lo - 1]);
    return zs[lo - 1] + frac * (zs[hi - 1] - zs[lo - 1]);
}

void orc_build_chi_to_z(double Om, double OL, double h, double z_max, int npts,
                        double *zs, double *chis)
{
    for (int i = 0; i < npts; i++) {
        zs[i] = z_max * (double)i / (double)(npts - 1);
        chis[i] = orc_chi(zs[i], Om, OL, h);
    }
}

/* ------------------------------------------------------------------------ */
/* analyse_peak  (src/HaloFinder/RadialShell.jl:228-891)                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const float *delta, *etax, *etay, *etaz;
    const float *eta2x, *eta2y, *eta2z;   /* NULL if ilpt < 2 */
    int8_t *mask;                         /* NULL if none */
    const float *lapd;                    /* NULL if none */
    int n1, n2, n3;
} orc_grid;

typedef struct {
    double RTHL, Srb, Sbar[3], Sbar2[3], strain[9], gradpk[3], gradpkrf[3];
    double Fbarx, e_v, p_v, zvir_half, gradpkf[3], d2F;
} orc_result;

enum { DUMP_NONE = 0, DUMP_PREMASK = 1, DUMP_NO_FCRIT = 2, DUMP_M0_ONE = 3,
       DUMP_NO_COLLAPSE = 4, DUMP_RTHL_NEG = 5, DUMP_NONFINITE = 6 };

/* get_evals goes through LAPACK.syevr!, which raises ArgumentError on a
 * matrix holding Inf/NaN (chkuplofinite).  kernel_strain at mp == 1 divides
 * by rad[1] == 0 (RadialShell.jl:263), so a walk that reaches mp == 1 makes
 * the reference raise.  The restatement reports such a peak as
 * DUMP_NONFINITE with the no_collapse() record instead of aborting. */
static int nonfinite9(const double *E)
{
    for (int i = 0; i < 9; i++) if (!isfinite(E[i])) return 1;
    return 0;
}

typedef struct {
    const int32_t *di, *dj, *dk, *r2;
    int64_t ncells;
} orc_shells;

static void no_collapse(orc_result *r)
{
    memset(r, 0, sizeof(*r));
    r->RTHL = -1.0;
    r->zvir_half = -1.0;
}

/* Per-shell profile arrays, 1-based: X[1..maxm] */
typedef struct {
    double *rad, *Fbar;
    int64_t *nshell;
    double *S[3], *S2[3], *G[3], *Gf[3], *SR[3][3];
} orc_prof;

static void kernel_strain(const orc_prof *P, int mp, int mlow, int mupp, const double *akk,
                          double wRnor, double aRnor, double hl1, double hl2,
                          double Ebar[9], double grad[3], double gradf[3])
{   /* RadialShell.jl:228-277; Ebar column-major (L,K) -> L + 3K */
    for (int i = 0; i < 9; i++) Ebar[i] = 0.0;
    for (int i = 0; i < 3; i++) grad[i] = gradf[i] = 0.0;
    const double *rad = P->rad;
    if (mp > 1) {
        for (int m1 = mlow; m1 <= mupp; m1++) {
            double u0 = hl1 * (rad[mp] - rad[m1]);
            double u02 = u0 * u0;
            if (u02 < 4.0) {
                double u12 = hl2 * (rad[mp] * rad[mp] + rad[m1] * rad[m1]);
                double den = u12 - u02;
                double wt = orc_hRinteg(u0, u12) / (den * den);
                for (int L = 0; L < 3; L++) {
                    grad[L] += wt * P->G[L][m1];
                    gradf[L] += wt * P->Gf[L][m1];
                    for (int K = 0; K < 3; K++)
                        Ebar[L + 3 * K] -= 0.5 * wt * (P->SR[L][K][m1] + P->SR[K][L][m1]);
                }
            }
        }
        double scale = wRnor / rad[mp];
        for (int i = 0; i < 9; i++) Ebar[i] *= scale;
        for (int i = 0; i < 3; i++) { grad[i] *= scale; gradf[i] *= scale; }
    } else {
        for (int m1 = mlow; m1 <= mupp; m1++) {
            double u0 = hl1 * rad[m1];
            double con = 100.0 * u0 * u0;
            int icon = (int)floor(con);
            double diff = con - icon;
            int i2c = icon + 1;                               /* 1-based */
            double aww = akk[i2c - 1] + diff * (akk[i2c] - akk[i2c - 1]);
            double wt = 0.5 * aww / rad[m1];
            for (int L = 0; L < 3; L++) {
                grad[L] += wt * P->G[L][m1];
                gradf[L] += wt * P->Gf[L][m1];
                for (int K = 0; K < 3; K++)
                    Ebar[L + 3 * K] -= 0.5 * wt * (P->SR[L][K][m1] + P->SR[K][L][m1]);
            }
        }
        for (int i = 0; i < 9; i++) Ebar[i] *= aRnor;
        for (int i = 0; i < 3; i++) { grad[i] *= aRnor; gradf[i] *= aRnor; }
    }
}

static void evals_ep(const double *E, double *lam, double *Frho, double *e_v, double *p_v)
{
    orc_get_evals(E, lam);
    *Frho = lam[0] + lam[1] + lam[2];
    *e_v = 0.0; *p_v = 0.0;
    if (*Frho > 0.0) {
        *e_v = 0.5 * (lam[2] - lam[0]) / *Frho;
        *p_v = 0.5 * (lam[2] + lam[0] - 2.0 * lam[1]) / *Frho;
    }
}

/* Optional per-peak profile export (for shell-count / profile parity).     */
typedef struct {
    int m;               /* last finalised shell index (1-based count)    */
    int cap;             /* capacity of the arrays below                  */
    int64_t *nshell;     /* [cap], entry s-1 = nshell[s]                  */
    double *Fbar, *Fsum; /* [cap]                                         */
    int m0_final;        /* m0 after the walk, 0 if no collapse           */
} orc_profile_out;

static int analyse_peak_impl(const orc_grid *pg, int64_t ipp, double alatt, int64_t ir2min,
                             double ZZon, double Rfclvi, const orc_table *ct,
                             const orc_shells *sh, int nbuff, double rmax2rs,
                             double fcrit, const double *akk, orc_result *res,
                             orc_profile_out *pout)
{
    const int32_t *S_di = sh->di, *S_dj = sh->dj, *S_dk = sh->dk, *S_r2 = sh->r2;
    /* shells[jp] (1-based) -> index jp-1 */
#define SH_R2(jp) S_r2[(jp) - 1]
    int64_t npartmax = sh->ncells, npart;
    if (rmax2rs > 0.0) {
        double f3p = 4.0 / 3.0 * ORC_PI;
        double t = rmax2rs * Rfclvi / alatt;
        double v = floor(f3p * (t * t * t));
        npart = (int64_t)v < npartmax ? (int64_t)v : npartmax;
    } else npart = npartmax;

    double hlatt = 1.0, anor = 1.0 / (hlatt * hlatt * hlatt);
    double aRnor = 3.0 * anor / alatt;
    double wnor = anor / (4.0 * ORC_PI);
    double wRnor = 3.0 * wnor / alatt;

    int n1 = pg->n1, n2 = pg->n2, n3 = pg->n3;
    int64_t n1xn2 = (int64_t)n1 * n2;
    int ik = (int)((ipp - 1) / n1xn2 + 1);
    int ij = (int)(((ipp - 1) % n1xn2) / n1 + 1);
    int ii = (int)((ipp - 1) % n1 + 1);
#define IDX(a, b, c) (((size_t)(a) - 1) + ((size_t)(b) - 1) * (size_t)n1 + ((size_t)(c) - 1) * (size_t)n1xn2)

    if (rmax2rs > 0.0 && pg->mask != NULL) {
        if ((double)pg->mask[IDX(ii, ij, ik)] > floor(Rfclvi / alatt * rmax2rs)) {
            no_collapse(res);
            return DUMP_PREMASK;
        }
    }

    int64_t maxm = npart + 2;
    /* one block of doubles for all profile arrays, 1-based (index 0 unused) */
    size_t L1 = (size_t)maxm + 2;
    double *buf = (double *)calloc(L1 * 29, sizeof(double));
    int64_t *nsh_arr = (int64_t *)calloc(L1, sizeof(int64_t));
    orc_prof P;
    P.rad = buf; P.Fbar = buf + L1;
    P.nshell = nsh_arr;
    for (int L = 0; L < 3; L++) {
        P.S[L] = buf + L1 * (2 + L);
        P.S2[L] = buf + L1 * (5 + L);
        P.G[L] = buf + L1 * (8 + L);
        P.Gf[L] = buf + L1 * (11 + L);
        for (int K = 0; K < 3; K++) P.SR[L][K] = buf + L1 * (14 + 3 * L + K);
    }
    double *Fsum_dbg = buf + L1 * 23;
    double *rad = P.rad, *Fbar = P.Fbar;
    int64_t *nshell = P.nshell;

    /* ---- Phase 1+2 (RadialShell.jl:333-453) ---- */
    int m = 1, m0 = 1;
    rad[1] = 0.0;
    Fbar[1] = pg->delta[IDX(ii, ij, ik)];
    nshell[1] = 1;
    int ilpt2 = pg->eta2x != NULL;
    double ev0[3] = {pg->etax[IDX(ii, ij, ik)], pg->etay[IDX(ii, ij, ik)], pg->etaz[IDX(ii, ij, ik)]};
    double e2v0[3] = {0, 0, 0};
    if (ilpt2) {
        e2v0[0] = pg->eta2x[IDX(ii, ij, ik)];
        e2v0[1] = pg->eta2y[IDX(ii, ij, ik)];
        e2v0[2] = pg->eta2z[IDX(ii, ij, ik)];
    }
    for (int L = 0; L < 3; L++) {
        P.S[L][1] = ev0[L]; P.S2[L][1] = e2v0[L];
        P.S[L][2] = 0.0; P.S2[L][2] = 0.0;
        P.G[L][1] = P.G[L][2] = 0.0;
        P.Gf[L][1] = P.Gf[L][2] = 0.0;
        for (int K = 0; K < 3; K++) P.SR[L][K][1] = P.SR[L][K][2] = 0.0;
    }
    double Fshell = 0.0;
    int64_t nsh = 0;
    int64_t ir2p = SH_R2(2);
    double dFbarp = Fbar[1];
    double rupp_init = sqrt((double)ir2min) + 2.0 * hlatt;
    int ifcrit = 1, mupp = 1;
    (void)rupp_init;

    for (int64_t jp = 2; jp <= npart; jp++) {
        int64_t ir2 = SH_R2(jp);
        if (ir2 != ir2p) {
            m += 1;
            rad[m] = sqrt((double)ir2p);
            nshell[m] = nsh;
            Fsum_dbg[m] = Fshell;
            double dFbar = nsh > 0 ? Fshell / (double)nsh : 0.0;
            double rad3p = rad[m - 1] * rad[m - 1] * rad[m - 1];
            double rad3 = rad[m] * rad[m] * rad[m];
            Fbar[m] = (rad3p * Fbar[m - 1] + 0.5 * (dFbarp + dFbar) * (rad3 - rad3p)) / rad3;
            dFbarp = dFbar;
            if (ifcrit == 1) {
                m0 = m;
                if (ir2p > ir2min) {
                    rupp_init = rad[m0] + 2.0 * hlatt;
                    ifcrit = 0;
                    mupp = m;
                }
            } else {
                if (rad[m] < rupp_init) mupp = m;
            }
            ir2p = ir2;
            Fshell = 0.0;
            nsh = 0;
            for (int L = 0; L < 3; L++) {
                P.S[L][m + 1] = 0.0; P.S2[L][m + 1] = 0.0;
                P.G[L][m] /= rad[m]; P.Gf[L][m] /= rad[m];
                P.G[L][m + 1] = 0.0; P.Gf[L][m + 1] = 0.0;
                for (int K = 0; K < 3; K++) {
                    P.SR[L][K][m] /= rad[m];
                    P.SR[L][K][m + 1] = 0.0;
                }
            }
        }
        int v1 = ii + S_di[jp - 1], v2 = ij + S_dj[jp - 1], v3 = ik + S_dk[jp - 1];
        if (v1 < 1 || v1 > n1 || v2 < 1 || v2 > n2 || v3 < 1 || v3 > n3) continue;
        nsh += 1;
        double dv = pg->delta[IDX(v1, v2, v3)];
        Fshell += dv;
        double ev[3] = {pg->etax[IDX(v1, v2, v3)], pg->etay[IDX(v1, v2, v3)], pg->etaz[IDX(v1, v2, v3)]};
        double e2v[3] = {0, 0, 0};
        if (ilpt2) {
            e2v[0] = pg->eta2x[IDX(v1, v2, v3)];
            e2v[1] = pg->eta2y[IDX(v1, v2, v3)];
            e2v[2] = pg->eta2z[IDX(v1, v2, v3)];
        }
        double dT = pg->delta[IDX(v3, v1, v2)];   /* quirk Q6: permuted read */
        double ixs[3] = {S_di[jp - 1], S_dj[jp - 1], S_dk[jp - 1]};
        for (int L = 0; L < 3; L++) {
            P.S[L][m + 1] += ev[L];
            P.S2[L][m + 1] += e2v[L];
            P.G[L][m + 1] += dv * ixs[L];
            P.Gf[L][m + 1] += dT * ixs[L];
            for (int K = 0; K < 3; K++) P.SR[L][K][m + 1] += ev[L] * ixs[K];
        }
    }

    if (pout) {
        pout->m = m;
        int lim = m < pout->cap ? m : pout->cap;
        for (int s = 1; s <= lim; s++) {
            pout->nshell[s - 1] = nshell[s];
            pout->Fbar[s - 1] = Fbar[s];
            pout->Fsum[s - 1] = s == 1 ? Fbar[1] : Fsum_dbg[s];
        }
        pout->m0_final = 0;
    }

    /* ---- Phase 3 (RadialShell.jl:456-490) ---- */
    double rmrf = 10.0;
    int mrf = 1, mrfi = 1;
    int64_t ir2p_rf = SH_R2(2);
    double tgt = (Rfclvi / alatt) * (Rfclvi / alatt);
    for (int64_t jp = 2; jp <= npart; jp++) {
        int64_t ir2 = SH_R2(jp);
        double dd = fabs((double)ir2 - tgt);
        if (dd < rmrf) { mrf = mrfi; rmrf = rmrf < dd ? rmrf : dd; }
        if (ir2 != ir2p_rf) { mrfi += 1; ir2p_rf = ir2; }
    }
    double rlow_rf = rad[mrf] - 2.0 > 0.0 ? rad[mrf] - 2.0 : 0.0;
    int mlow_rf = mrf;
    if (mrf > 1)
        for (int m1 = mrf - 1; m1 >= 1; m1--) if (rad[m1] > rlow_rf) mlow_rf = m1;
    double rupp_rf = rad[mrf] + 2.0;
    int mupp_rf = mrf;
    for (int m1 = mrf + 1; m1 <= m; m1++) if (rad[m1] < rupp_rf) mupp_rf = m1;
    double Etmp[9], gradpkrf[3], gtmp[3];
    kernel_strain(&P, mrf, mlow_rf, mupp_rf, akk, wRnor, aRnor, 1.0, 1.0, Etmp, gradpkrf, gtmp);

    int status = DUMP_NONE;
    /* ---- Phase 4 (RadialShell.jl:492-538) ---- */
    if (Fbar[m0] >= fcrit) {
        int found = 0;
        for (int mm = m0; mm <= m; mm++)
            if (Fbar[mm] < fcrit) { m0 = mm; found = 1; break; }
        if (!found) mupp = m;
        if (found) {
            double rupp_m0 = rad[m0] + 2.0;
            mupp = m0;
            for (int mm = m0 + 1; mm <= m; mm++) if (rad[mm] < rupp_m0) mupp = mm;
        }
    } else {
        int mstart = m0 - 1 > 1 ? m0 - 1 : 1;
        int found = 0;
        for (int mp = mstart; mp >= 1; mp--) {
            if (Fbar[mp] >= fcrit) { found = 1; break; }
            m0 = mp;
        }
        if (!found) { status = DUMP_NO_FCRIT; goto nocol; }
        double rupp_m0 = rad[m0] + 2.0;
        int mupp_new = m0;
        for (int mm = m0 + 1; mm <= m; mm++) if (rad[mm] < rupp_m0) mupp_new = mm;
        mupp = mupp_new;
    }

    /* ---- Phase 5 (RadialShell.jl:540-704) ---- */
    {
        double rlow = rad[m0] - 2.0 > 0.0 ? rad[m0] - 2.0 : 0.0;
        int mlow = m0;
        if (m0 > 1)
            for (int m1 = m0 - 1; m1 >= 1; m1--) if (rad[m1] > rlow) mlow = m1;
        double zvir1 = -1.0, zvir1p = -1.0;
        double Ebar_last[9] = {0}, gradpk_last[3] = {0}, Frhoh = 0.0, Frho = 0.0;
        double e_v = 0.0, p_v = 0.0, strain_mat[9] = {0}, gradpk_prev[3] = {0};
        double gradpkf_last[3] = {0}, gradpkf_prev[3] = {0};
        double Frhpk = 0.0, Fnupk = 0.0, Fevpk = 0.0, Fpvpk = 0.0;
        int collapsed = 0;
        double Ebar_m0[9], gradpk_m0[3], gradpkf_m0[3], lam[3], Frho_m0, e_v_m0, p_v_m0;
        kernel_strain(&P, m0, mlow, mupp, akk, wRnor, aRnor, 1.0, 1.0, Ebar_m0, gradpk_m0, gradpkf_m0);
        if (nonfinite9(Ebar_m0)) { status = DUMP_NONFINITE; goto nocol; }
        evals_ep(Ebar_m0, lam, &Frho_m0, &e_v_m0, &p_v_m0);
        double Frhoh_m0 = Frho_m0;
        Frho_m0 = Fbar[m0];
        double zvir1p_m0 = -1.0;   /* quirk Q3 */
        if (zvir1p_m0 >= ZZon) {
            zvir1 = -1.0; zvir1p = zvir1p_m0;
            memcpy(Ebar_last, Ebar_m0, sizeof Ebar_last);
            memcpy(gradpk_last, gradpk_m0, sizeof gradpk_last);
            memcpy(gradpkf_last, gradpkf_m0, sizeof gradpkf_last);
            memcpy(strain_mat, Ebar_m0, sizeof strain_mat);
            memcpy(gradpk_prev, gradpk_m0, sizeof gradpk_prev);
            memcpy(gradpkf_prev, gradpkf_m0, sizeof gradpkf_prev);
            Frhoh = Frhoh_m0; Frho = Frho_m0; e_v = e_v_m0; p_v = p_v_m0;
            Frhpk = Frhoh_m0; Fnupk = Frho_m0; Fevpk = e_v_m0; Fpvpk = p_v_m0;
            collapsed = 1;
        } else {
            zvir1 = zvir1p_m0;
            Frhpk = Frhoh_m0; Fnupk = Frho_m0; Fevpk = e_v_m0; Fpvpk = p_v_m0;
            memcpy(gradpk_prev, gradpk_m0, sizeof gradpk_prev);
            memcpy(gradpkf_prev, gradpkf_m0, sizeof gradpkf_prev);
            memcpy(strain_mat, Ebar_m0, sizeof strain_mat);
            Frhoh = Frhoh_m0; Frho = Frho_m0; e_v = e_v_m0; p_v = p_v_m0;
            memcpy(Ebar_last, Ebar_m0, sizeof Ebar_last);
            memcpy(gradpk_last, gradpk_m0, sizeof gradpk_last);
            memcpy(gradpkf_last, gradpkf_m0, sizeof gradpkf_last);
            if (m0 == 1) { status = DUMP_M0_ONE; goto nocol; }
            for (int mp = m0 - 1; mp >= 1; mp--) {
                double rupp_mp = rad[mp] + 2.0;
                double rlow_mp = rad[mp] - 2.0 > 0.0 ? rad[mp] - 2.0 : 0.0;
                if (rad[mupp] > rupp_mp) {
                    int mupp_new = mp;
                    for (int m1 = mp + 1; m1 <= mupp; m1++) if (rad[m1] < rupp_mp) mupp_new = m1;
                    mupp = mupp_new;
                }
                if (mlow > 1 && rad[mlow - 1] > rlow_mp) {
                    int mlownew = mlow;
                    for (int m1 = mlow - 1; m1 >= 1; m1--) if (rad[m1] > rlow_mp) mlownew = m1;
                    mlow = mlownew;
                }
                double Ebar_mp[9], gradpk_mp[3], gradpkf_mp[3], Frho_mp, e_v_mp, p_v_mp;
                kernel_strain(&P, mp, mlow, mupp, akk, wRnor, aRnor, 1.0, 1.0, Ebar_mp, gradpk_mp, gradpkf_mp);
                if (nonfinite9(Ebar_mp)) { status = DUMP_NONFINITE; goto nocol; }
                evals_ep(Ebar_mp, lam, &Frho_mp, &e_v_mp, &p_v_mp);
                double Frhoh_mp = Frho_mp;
                Frho_mp = Fbar[mp];
                double zvir1p_mp = -1.0;
                if (Frho_mp > 0.0) {
                    double poe = e_v_mp < 1e-5 ? 0.0 : p_v_mp / e_v_mp;
                    zvir1p_mp = orc_interp(ct, log10(Frho_mp), e_v_mp, poe);
                }
                int m0_next = mp + 1;
                if (zvir1p_mp >= ZZon) {
                    m0 = m0_next;
                    zvir1p = zvir1p_mp;
                    memcpy(Ebar_last, Ebar_mp, sizeof Ebar_last);
                    memcpy(gradpk_last, gradpk_mp, sizeof gradpk_last);
                    memcpy(gradpkf_last, gradpkf_mp, sizeof gradpkf_last);
                    Frhoh = Frhoh_mp; Frho = Frho_mp; e_v = e_v_mp; p_v = p_v_mp;
                    collapsed = 1;
                    break;
                }
                zvir1 = zvir1p_mp;
                Frhpk = Frhoh_mp; Fnupk = Frho_mp; Fevpk = e_v_mp; Fpvpk = p_v_mp;
                memcpy(gradpk_prev, gradpk_mp, sizeof gradpk_prev);
                memcpy(gradpkf_prev, gradpkf_mp, sizeof gradpkf_prev);
                memcpy(strain_mat, Ebar_mp, sizeof strain_mat);
                Frhoh = Frhoh_mp; Frho = Frho_mp; e_v = e_v_mp; p_v = p_v_mp;
                memcpy(Ebar_last, Ebar_mp, sizeof Ebar_last);
                memcpy(gradpk_last, gradpk_mp, sizeof gradpk_last);
                memcpy(gradpkf_last, gradpkf_mp, sizeof gradpkf_last);
            }
        }
        if (!collapsed) { status = DUMP_NO_COLLAPSE; goto nocol; }

        /* ---- Phase 6 (RadialShell.jl:706-754) ---- */
        double dZvir = zvir1p - zvir1, RTHL;
        if (zvir1 > 0.0 && dZvir != 0.0) {
            double a3 = rad[m0 - 1] * rad[m0 - 1] * rad[m0 - 1];
            double b3 = rad[m0] * rad[m0] * rad[m0];
            double RTHL3 = a3 + (b3 - a3) * (zvir1p - ZZon) / dZvir;
            RTHL = pow(RTHL3, 1.0 / 3.0);
        } else RTHL = rad[m0 - 1];
        if (RTHL <= 0.0) { status = DUMP_RTHL_NEG; goto nocol; }
        double RTHL3 = RTHL * RTHL * RTHL;
        double RTHL5 = RTHL3 * RTHL * RTHL;
        double rad3p = rad[m0 - 1] * rad[m0 - 1] * rad[m0 - 1];
        double rad3 = rad[m0] * rad[m0] * rad[m0];
        double drad3 = rad3 - rad3p;
        double dFbar = Fbar[m0] - Fbar[m0 - 1];
        double Fbarx, gradpk_final[3], gradpkf_final[3];
        if (zvir1 > 0.0 && drad3 != 0.0) {
            double frac = (RTHL3 - rad3p) / drad3;
            Fbarx = Fbar[m0 - 1] + frac * dFbar;
            Frhpk = Frhoh + frac * (Frhpk - Frhoh);
            Fnupk = Frho + frac * (Fnupk - Frho);
            Fevpk = e_v + frac * (Fevpk - e_v);
            Fpvpk = p_v + frac * (Fpvpk - p_v);
            for (int i = 0; i < 9; i++) strain_mat[i] = Ebar_last[i] + frac * (strain_mat[i] - Ebar_last[i]);
            for (int i = 0; i < 3; i++) {
                gradpk_final[i] = gradpk_last[i] + frac * (gradpk_prev[i] - gradpk_last[i]);
                gradpkf_final[i] = gradpkf_last[i] + frac * (gradpkf_prev[i] - gradpkf_last[i]);
            }
        } else {
            Fbarx = Fbar[m0 - 1];
            Fnupk = Frho; Fevpk = e_v; Fpvpk = p_v;
            memcpy(strain_mat, Ebar_last, sizeof strain_mat);
            memcpy(gradpk_final, gradpk_prev, sizeof gradpk_final);
            memcpy(gradpkf_final, gradpkf_prev, sizeof gradpkf_final);
        }
        (void)Frhpk; (void)Fnupk; (void)Fevpk; (void)Fpvpk;
        orc_normalize_strain(strain_mat, Fbarx);
        if (nonfinite9(strain_mat)) { status = DUMP_NONFINITE; goto nocol; }
        double lamf[3], Frhoc, e_v_f, p_v_f;
        evals_ep(strain_mat, lamf, &Frhoc, &e_v_f, &p_v_f);
        double smat[9];
        for (int i = 0; i < 9; i++) smat[i] = Ebar_last[i] * (Fbarx / Frhoh);

        /* ---- Phase 7 (RadialShell.jl:756-819) ---- */
        double Sbar[3] = {0, 0, 0}, Sbar2[3] = {0, 0, 0};
        int64_t nSbar = 0;
        for (int m1 = 1; m1 <= m0 - 1; m1++)
            if (nshell[m1] > 0) {
                for (int L = 0; L < 3; L++) { Sbar[L] += P.S[L][m1]; Sbar2[L] += P.S2[L][m1]; }
                nSbar += nshell[m1];
            }
        if (nSbar > 0)
            for (int L = 0; L < 3; L++) { Sbar[L] /= (double)nSbar; Sbar2[L] /= (double)nSbar; }
        double Srb = 0.0, rad5p = 0.0;
        if (m0 > 2)
            for (int mp = 2; mp <= m0 - 1; mp++) {
                double r2v = rad[mp] * rad[mp];
                double rad5 = r2v * r2v * rad[mp];
                Srb += 0.5 * (Fbar[mp - 1] + Fbar[mp]) * (rad5 - rad5p);
                rad5p = rad5;
            }
        if (zvir1 > 0.0 && dZvir != 0.0) Srb += 0.5 * (Fbar[m0 - 1] + Fbarx) * (RTHL5 - rad5p);
        if (Fbarx * RTHL5 > 0.0) Srb /= (Fbarx * RTHL5);

        if (pg->mask != NULL) {
            for (int64_t jp = 1; jp <= npart; jp++) {
                if (sqrt((double)SH_R2(jp)) > RTHL) continue;
                int v1 = ii + S_di[jp - 1], v2 = ij + S_dj[jp - 1], v3 = ik + S_dk[jp - 1];
                if (v1 < nbuff + 1 || v1 > n1 - nbuff || v2 < nbuff + 1 || v2 > n2 - nbuff ||
                    v3 < nbuff + 1 || v3 > n3 - nbuff) continue;
                pg->mask[IDX(v1, v2, v3)] = 1;
            }
        }
        double d2F = 0.0;
        if (pg->lapd != NULL) {
            int64_t navg = 0;
            for (int64_t jp = 2; jp <= npart; jp++) {
                if (sqrt((double)SH_R2(jp)) > RTHL) continue;
                int v1 = ii + S_di[jp - 1], v2 = ij + S_dj[jp - 1], v3 = ik + S_dk[jp - 1];
                if (v1 < 1 || v1 > n1 || v2 < 1 || v2 > n2 || v3 < 1 || v3 > n3) break;
                navg += 1;
                d2F += (double)pg->lapd[IDX(v1, v2, v3)];
            }
            if (navg > 0) d2F /= (double)navg;
        }

        /* ---- zvir_half (RadialShell.jl:821-885) ---- */
        double zvir_half = -1.0;
        if (RTHL / hlatt >= 3.0) {
            for (int jj = 1; jj <= 10; jj++) {
                double rcur = RTHL * pow((double)(jj - 1) / 9.0 * 0.5 + 0.5, 1.0 / 3.0);
                double rupp_rc = rcur + 2.0 * hlatt;
                int mupp_rc = 1;
                for (int m1 = 2; m1 <= m - 1; m1++) {
                    if (rad[m1] > rupp_rc) { mupp_rc = m1; break; }
                    else if (rad[m1] == 0.0) { mupp_rc = m1 - 1; break; }
                    mupp_rc = m1;
                }
                double rlow_rc = rcur - 2.0 * hlatt > 0.0 ? rcur - 2.0 * hlatt : 0.0;
                int mlow_rc = 1;
                for (int m1 = 1; m1 <= m - 1; m1++) if (rad[m1] > rlow_rc) { mlow_rc = m1; break; }
                int m0_rc = 1;
                for (int m1 = 1; m1 <= m - 1; m1++) {
                    if (rad[m1] > rcur) { m0_rc = m1; break; }
                    m0_rc = m1;
                }
                double Erc[9], g1[3], g2[3], Frho_rc, e_rc, p_rc;
                kernel_strain(&P, m0_rc, mlow_rc, mupp_rc, akk, wRnor, aRnor, 1.0, 1.0, Erc, g1, g2);
                if (nonfinite9(Erc)) { status = DUMP_NONFINITE; goto nocol; }
                evals_ep(Erc, lam, &Frho_rc, &e_rc, &p_rc);
                double Frhoc_rc = Fbar[m0_rc];
                double poe_rc = e_rc < 1e-5 ? 0.0 : p_rc / e_rc;
                double zvir_rc = -1.0;
                if (Frhoc_rc > 0.0) zvir_rc = orc_interp(ct, log10(Frhoc_rc), e_rc, poe_rc);
                double zz = zvir_rc - 1.0;
                zvir_half = zz > zvir_half ? zz : zvir_half;
            }
        }

        res->RTHL = RTHL; res->Srb = Srb;
        for (int L = 0; L < 3; L++) {
            res->Sbar[L] = Sbar[L]; res->Sbar2[L] = Sbar2[L];
            res->gradpk[L] = gradpk_final[L]; res->gradpkrf[L] = gradpkrf[L];
            res->gradpkf[L] = gradpkf_final[L];
        }
        memcpy(res->strain, smat, sizeof smat);
        res->Fbarx = Fbarx; res->e_v = e_v_f; res->p_v = p_v_f;
        res->zvir_half = zvir_half; res->d2F = d2F;
        if (pout) pout->m0_final = m0;
    }
    free(buf); free(nsh_arr);
    return DUMP_NONE;
nocol:
    no_collapse(res);
    free(buf); free(nsh_arr);
    return status;
#undef IDX
#undef SH_R2
}

/* Public single-peak entry.  ipp is the 1-based column-major flat index. */
int orc_analyse_peak(const orc_grid *pg, int64_t ipp, double alatt, int64_t ir2min,
                     double ZZon, double Rfclvi, const orc_table *ct,
                     const orc_shells *sh, int nbuff, double rmax2rs, double fcrit,
                     orc_result *res, orc_profile_out *pout)
{
    double akk[1101];
    orc_atab4(akk);
    return analyse_peak_impl(pg, ipp, alatt, ir2min, ZZon, Rfclvi, ct, sh, nbuff, rmax2rs,
                             fcrit, akk, res, pout);
}

/* Batched entry, OpenMP dynamic schedule over peaks (the reference's
 * Threads.@threads over analyse_peak, Pipeline.jl:233-241).  Per-peak
 * arrays: ipp (1-based), ir2min, ZZon, Rf, fcrit.  status[p] = dump reason. */
int orc_analyse_peaks(const orc_grid *pg, int64_t npk, const int64_t *ipp,
                      const int64_t *ir2min, const double *ZZon, const double *Rf,
                      const double *fcrit, double alatt, const orc_table *ct,
                      const orc_shells *sh, int nbuff, double rmax2rs, int nthreads,
                      orc_result *res, int32_t *status)
{
    double akk[1101];
    orc_atab4(akk);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < npk; p++)
        status[p] = analyse_peak_impl(pg, ipp[p], alatt, ir2min[p], ZZon[p], Rf[p], ct, sh,
                                      nbuff, rmax2rs, fcrit[p], akk, &res[p], NULL);
    return 0;
}

size_t orc_result_size(void) { return sizeof(orc_result); }

/* ------------------------------------------------------------------------ */
/* Ellipsoidal collapse, Fortran-compatible RK4 path                        */
/* (src/EllipsoidalCollapse/EllipsoidalCollapse.jl:60-361, 510-624) and the  */
/* table grid (src/EllipsoidalCollapse/CollapseTable.jl:9-45).               */
/* Parameters fixed at the reference's defaults: iforce_strat 4, ivir_strat */
/* 2, fcoll (0.01, 0.171, 0.171), tfac 0.01, nstepmax 10000, zinit_fac 20,  */
/* iwant_rd 1.                                                               */
/* ------------------------------------------------------------------------ */
double orc_dfnofa(const orc_dlin *T, double ai)
{   /* Dfnofa (Cosmology.jl:296-316): the D column of the same 4-point rule */
    double Da;
    (void)orc_dlinear_ab(T, ai, NULL, &Da);
    return ai * Da;
}

double orc_elliptic_rd(double x, double y, double z)
{   /* _elliptic_rd (EllipsoidalCollapse.jl:63-107), Carlson's duplication */
    const double ERRTOL = 0.05, C1 = 3.0 / 14.0, C2 = 1.0 / 6.0, C3 = 9.0 / 22.0, C4 = 3.0 / 26.0;
    const double C5 = 0.25 * C3, C6 = 1.5 * C4;
    double xt = x, yt = y, zt = z, sum = 0.0, fac = 1.0, delx = 0, dely = 0, delz = 0, ave = 0;
    for (;;) {
        double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        double alamb = sx * (sy + sz) + sy * sz;
        sum += fac / (sz * (zt + alamb));
        fac *= 0.25;
        xt = 0.25 * (xt + alamb);
        yt = 0.25 * (yt + alamb);
        zt = 0.25 * (zt + alamb);
        ave = 0.2 * (xt + yt + 3.0 * zt);
        delx = (ave - xt) / ave;
        dely = (ave - yt) / ave;
        delz = (ave - zt) / ave;
        double mx = fabs(delx);
        if (fabs(dely) > mx) mx = fabs(dely);
        if (fabs(delz) > mx) mx = fabs(delz);
        if (mx <= ERRTOL) break;
    }
    double ea = delx * dely, eb = delz * delz, ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
    return 3.0 * sum + fac * (1.0 + ed * (-C1 + C5 * ed - C6 * delz * ee) +
                              delz * (C2 * ee + delz * (-C3 * ec + delz * C4 * ea))) / (ave * sqrt(ave));
}

void orc_get_b_2(double a1, double a2, double a3, double *b)
{   /* get_b_2 (EllipsoidalCollapse.jl:113-146), iwant_rd = 1 */
    const double third = 1.0 / 3.0;
    double r2 = a2 / a1, r3 = a3 / a1;
    if (r3 >= r2) {
        double b3;
        if (r3 >= 0.999) { b[0] = b[1] = b[2] = third; return; }
        if (r3 <= 0.001) {
            double ecc2 = 1.0 - r3 * r3, one_ecc2 = r3 * r3, one_ecc = 0.5 * r3 * r3, ecc = sqrt(ecc2);
            b3 = 0.5 * (one_ecc2 / ecc2 * (1.0 / one_ecc2 - log((1.0 + ecc) / one_ecc) * 0.5 / ecc));
        } else {
            double ecc2 = 1.0 - r3 * r3, ecc = sqrt(ecc2);
            b3 = 0.5 * ((1.0 - ecc2) / ecc2 * (1.0 / (1.0 - ecc2) - log((1.0 + ecc) / (1.0 - ecc)) * 0.5 / ecc));
        }
        b[0] = 1.0 - 2.0 * b3; b[1] = b3; b[2] = b3;
        return;
    }
    double r22 = r2 * r2, r32 = r3 * r3, q = third * r2 * r3;
    double b3 = q * orc_elliptic_rd(r22, 1.0, r32);
    double b2 = q * orc_elliptic_rd(r32, 1.0, r22);
    b[0] = 1.0 - (b2 + b3); b[1] = b2; b[2] = b3;
}

typedef struct {
    const orc_dlin *T;
    double Rvac, Rcurv;             /* Omvac / Omnr, Omcurv / Omnr */
    int lvirv[3];
    double a_eq[3], a_3eq2, a_3eq1; /* frozen axis values; thresholds of axes 2 and 1 */
    double Ha;                      /* last Ha_b_nr */
    double aLam[3], Frho_3;
} orc_ell;

static double ha_b_nr(double a, const orc_ell *s)
{
    return sqrt((1.0 + s->Rvac * a * a * a + s->Rcurv * a) / a);
}

static void ell_derivs(double t, double *y, double *dy, orc_ell *s)
{   /* get_derivs! (EllipsoidalCollapse.jl:213-361), iforce 4, ivir 2; y[0..5] */
    const double fcoll[3] = {0.01, 0.171, 0.171};
    for (int i = 0; i < 6; i++) dy[i] = 0.0;
    double a_b = t, a_b3 = a_b * a_b * a_b;
    double Ha = ha_b_nr(a_b, s);
    s->Ha = Ha;
    double Hi = 1.0 / Ha;
    if (s->lvirv[2] == 0 && y[2] <= fcoll[2] * a_b) { s->lvirv[2] = 1; s->a_eq[2] = fcoll[2] * a_b; }
    if (s->lvirv[1] == 0) s->a_3eq2 = fcoll[1] * a_b;
    if (s->lvirv[0] == 0) s->a_3eq1 = fcoll[0] * a_b;
    if (s->lvirv[2] == 1 && y[1] <= s->a_3eq2 && s->lvirv[1] == 0) { s->lvirv[1] = 1; s->a_eq[1] = s->a_3eq2; }
    if (s->lvirv[2] == 1 && y[0] <= s->a_3eq1 && s->lvirv[0] == 0) { s->lvirv[0] = 1; s->a_eq[0] = s->a_3eq1; }
    for (int ax = 2; ax >= 0; ax--)
        if (s->lvirv[ax] == 1) { y[ax] = s->a_eq[ax]; y[ax + 3] = 0.0; dy[ax] = 0.0; dy[ax + 3] = 0.0; }
    double a1 = y[0], a2 = y[1], a3 = y[2];
    double delta1_e = a_b3 / (a1 * a2 * a3);
    double d0 = (0.5 / a_b3 - s->Rvac) * Hi;
    double d1_int = 1.5 * (delta1_e - 1.0) / a_b3 * Hi;
    double dlin = orc_dlinear_ab(s->T, a_b, NULL, NULL);
    double d1_ext = 1.5 / a_b3 * Hi * dlin;
    double b[3];
    orc_get_b_2(a1, a2, a3, b);
    for (int ax = 2; ax >= 0; ax--)
        if (s->lvirv[ax] == 0) {
            dy[ax] = y[ax + 3] * Hi;
            dy[ax + 3] = -y[ax] * (d1_int * b[ax] + d0 + d1_ext * (s->aLam[ax] - s->Frho_3));
        }
}

static void ell_rk4(double *y, const double *dy, double t, double h, double *yout, orc_ell *s)
{   /* rk4_homel! (EllipsoidalCollapse.jl:151-178) */
    double yt[6], dyt[6], dym[6], hh = h * 0.5, h6 = h / 6.0, xh = t + hh;
    for (int i = 0; i < 6; i++) yt[i] = y[i] + hh * dy[i];
    ell_derivs(xh, yt, dyt, s);
    for (int i = 0; i < 6; i++) yt[i] = y[i] + hh * dyt[i];
    ell_derivs(xh, yt, dym, s);
    for (int i = 0; i < 6; i++) { yt[i] = y[i] + h * dym[i]; dym[i] = dyt[i] + dym[i]; }
    ell_derivs(t + h, yt, dyt, s);
    for (int i = 0; i < 6; i++) yout[i] = y[i] + h6 * (dy[i] + dyt[i] + 2.0 * dym[i]);
}

/* evolve_ellipse_full, solver :rk4, idynax = 1 (EllipsoidalCollapse.jl:527-624):
 * returns zvir = 1/a at the first axis's virialisation (1 + z), or -1. */
double orc_evolve_ellipse(const orc_dlin *T, double Om, double OL, double Frho, double e_v, double p_v,
                          int *nsteps)
{
    orc_ell s;
    memset(&s, 0, sizeof s);
    s.T = T;
    s.Rvac = OL / Om;
    s.Rcurv = (1.0 - Om - OL) / Om;
    s.aLam[2] = Frho / 3.0 * (1.0 + 3.0 * e_v + p_v);
    s.aLam[1] = Frho / 3.0 * (1.0 - 2.0 * p_v);
    s.aLam[0] = Frho / 3.0 * (1.0 - 3.0 * e_v + p_v);
    s.Frho_3 = Frho / 3.0;
    double zinit = 20.0 * Frho, a_b = 1.0 / zinit, t = a_b;
    double HD_Ha, D_a;
    double dlin = orc_dlinear_ab(T, a_b, &HD_Ha, &D_a);
    double Ha = ha_b_nr(a_b, &s);
    s.Ha = Ha;
    double y[6], dy[6], yout[6];
    for (int ax = 0; ax < 3; ax++) {
        y[ax] = a_b * (1.0 - dlin * s.aLam[ax]);
        y[ax + 3] = Ha * (1.0 - (1.0 + HD_Ha) * dlin * s.aLam[ax]);
    }
    double zdyn1 = -1.0;
    int ldyn1 = 0, istep = 0;
    for (;;) {
        istep++;
        double a3 = y[2];
        double dt = 0.01 * a3 * sqrt(a3) * s.Ha;
        ell_derivs(t, y, dy, &s);
        ell_rk4(y, dy, t, dt, yout, &s);
        memcpy(y, yout, sizeof y);
        t += dt;
        if (s.lvirv[0] != ldyn1) { zdyn1 = 1.0 / t; ldyn1 = 1; }
        if (s.lvirv[0] == 1 || istep >= 10000 || t >= 2.0) break;
    }
    if (nsteps) *nsteps = istep;
    return zdyn1;
}

/* make_table (CollapseTable.jl:33-55): table[i + nx (j + ny k)] = Float32(zvir)
 * at log10 Frho = X1 + i dX, e = Y1 + j dY, p/e = Z1 + k dZ. */
void orc_make_collapse_table(const orc_dlin *T, double Om, double OL, int nx, int ny, int nz, double X1, double X2,
                             double Y1, double Y2, double Z1, double Z2, float *out)
{
    double dX = (X2 - X1) / (nx - 1), dY = (Y2 - Y1) / (ny - 1), dZ = (Z2 - Z1) / (nz - 1);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double x = X1 + i * dX, yv = Y1 + j * dY, z = Z1 + k * dZ;
                double Frho = pow(10.0, x);
                out[i + (size_t)nx * (j + (size_t)ny * k)] = (float)orc_evolve_ellipse(T, Om, OL, Frho, yv, z * yv, NULL);
            }
}
