// pp_post.h -- device-side per-peak analysis after the shell gather: the
// Fbar recursion, fcrit crossing, inward virialisation walk, R_TH
// interpolation and the final catalogue quantities of analyse_peak
// (src/HaloFinder/RadialShell.jl:379-891).
//
// The whole wave executes one peak's analysis with wave-uniform control
// flow; the radial sums (kernel_strain, Sbar, Srb, d2F) are split across the
// 64 lanes and combined with a fixed butterfly, so results are deterministic.
// Everything here is Float64, including the strain eigenvalues and the
// collapse-table interpolation that decide zvir >= ZZon.
//
// Where the reference scans a monotone radius table linearly (mlow / mupp /
// m0_rc searches) a binary search is used; rad[] is strictly increasing so
// the index found is the same.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pp {

// Profile field order (per shell): count, F, S(3), S2(3), G(3), Gf(3), SR(9, L*3+K), lapd
enum : int { PF_N = 0, PF_F = 1, PF_S = 2, PF_S2 = 5, PF_G = 8, PF_GF = 11, PF_SR = 14, PF_L = 23, PF_NFIELD = 24 };

// Result record (doubles per peak)
enum : int {
    R_RTHL = 0, R_SRB = 1, R_SBAR = 2, R_SBAR2 = 5, R_STRAIN = 8, R_GRADPK = 17, R_GRADPKRF = 20,
    R_FBARX = 23, R_EV = 24, R_PV = 25, R_ZVIRH = 26, R_GRADPKF = 27, R_D2F = 30, R_NREC = 32
};

enum : int { DUMP_NONE = 0, DUMP_PREMASK = 1, DUMP_NO_FCRIT = 2, DUMP_M0_ONE = 3, DUMP_NO_COLLAPSE = 4,
             DUMP_RTHL_NEG = 5, DUMP_NONFINITE = 6,
             POST_DEFER = -1 /* not a status: the walk hit its step budget, nothing was written */ };

struct PeakParam {
    int32_t ci, cj, ck;  // 0-based centre
    int32_t m;           // finalised shells (1-based count, centre included)
    int32_t m0;          // initial m0 from the gather (first shell with r2 > ir2min)
    int32_t mrf;         // Phase-3 shell (with the reference's one-shell lag)
    int32_t npart;       // cells considered (rmax2rs cut)
    int32_t flags;       // bit0: gather sphere fully inside the grid; bit1: premask dump
    double ZZon, fcrit;
    int64_t order;       // caller's index of this peak
};

struct Table3 {
    const float *data;
    int nx, ny, nz;
    double X1, X2, Y1, Y2, Z1, Z2;
};

// Result arrays in the caller's layout (PeakResult as SoA, RadialShell.jl:58-72);
// all device pointers, index = the caller's peak index.
struct ResultView {
    double *RTHL, *Srb, *Sbar, *Sbar2, *strain, *gradpk, *gradpkrf, *Fbarx, *e_v, *p_v, *zvir_half, *gradpkf, *d2F;
    int32_t *status, *nshells;
    int32_t *m0;  // aux: final m0 per peak (0 if no collapse), for the deferred finish pass
    int32_t *q;   // aux: last shell with rad <= RTHL
};

struct PostConst {
    const double *rad;    // rad[s-1] = sqrt(r2 of shell s)
    const double *inv_rad; // 1 / rad[s-1] (inf for the centre)
    const double *akk;    // atab4, 1101 entries
    const int4 *wrow;     // per mp: {offset into wtab, first m1, last m1, 0}
    const double *wtab;   // radial-kernel weight * 1/rad[m1] per (mp, m1)
    const int32_t *cell;  // packed offsets, r^2 sorted
    const int32_t *shell_start;
    const int32_t *shell_ge; // [r2max + 2]: first shell (1-based) with r2 >= k
    const int32_t *shell_r2; // r2 of shell s at [s - 1]
    int r2max, nshells;
    Table3 tab;
    double wRnor, aRnor;
    int n1, n2, n3, nbuff;
    const float *lapd;    // for the rare out-of-bounds d2F walk
    int8_t *mask;         // NULL: no mask output
};

// ----------------------------------------------------------- wave helpers
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// hRinteg (RadialShell.jl:100-127)
__device__ __forceinline__ double d_hRinteg(double u0, double u12)
{
    double x0 = fabs(u0);
    if (x0 >= 2.0) return 0.0;
    const double one7 = 1.0 / 7.0;
    double result = 1.6 * u12 - 6.4 * one7;
    double x02 = x0 * x0, x04 = x02 * x02;
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

// Symmetric 3x3 eigenvalues, ascending, closed form (trigonometric solution
// of the characteristic cubic; Smith 1961).  a = {a00, a11, a22, a01, a02, a12}.
__device__ __forceinline__ void d_eig3(const double a[6], double lam[3])
{
    double p1 = a[3] * a[3] + a[4] * a[4] + a[5] * a[5];
    if (p1 == 0.0) {
        double x = a[0], y = a[1], z = a[2], t;
        if (x > y) { t = x; x = y; y = t; }
        if (y > z) { t = y; y = z; z = t; }
        if (x > y) { t = x; x = y; y = t; }
        lam[0] = x; lam[1] = y; lam[2] = z;
        return;
    }
    double q = (a[0] + a[1] + a[2]) / 3.0;
    double b0 = a[0] - q, b1 = a[1] - q, b2 = a[2] - q;
    double p2 = b0 * b0 + b1 * b1 + b2 * b2 + 2.0 * p1;
    double p = sqrt(p2 / 6.0);
    double det = b0 * (b1 * b2 - a[5] * a[5]) - a[3] * (a[3] * b2 - a[5] * a[4]) + a[4] * (a[3] * a[5] - b1 * a[4]);
    double r = det / (2.0 * p * p * p);
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    double phi = acos(r) / 3.0;
    double l3 = q + 2.0 * p * cos(phi);
    double l1 = q + 2.0 * p * cos(phi + 2.0943951023931954923);
    double l2 = 3.0 * q - l1 - l3;
    lam[0] = l1; lam[1] = l2; lam[2] = l3;
}

// interpolate(ct, x, y, z) (CollapseTable.jl:128-145): trilinear on the
// evenly spaced table, -1 outside [X1,X2]x[Y1,Y2]x[Z1,Z2].
__device__ __forceinline__ void d_lin_pos(double x, double lo, double hi, int n, int &i0, double &f)
{
    double step = (hi - lo) / (double)(n - 1);
    double p = (x - lo) / step + 1.0;
    double fl = floor(p);
    if (p == (double)n) fl -= 1.0;
    if (fl < 1.0) fl = 1.0;
    if (fl > (double)(n - 1)) fl = (double)(n - 1);
    i0 = (int)fl - 1;
    f = p - fl;
}

__device__ __forceinline__ double d_interp(const Table3 &t, double x, double y, double z)
{
    if (!(x >= t.X1 && x <= t.X2 && y >= t.Y1 && y <= t.Y2 && z >= t.Z1 && z <= t.Z2)) return -1.0;
    int ix, iy, iz;
    double fx, fy, fz;
    d_lin_pos(x, t.X1, t.X2, t.nx, ix, fx);
    d_lin_pos(y, t.Y1, t.Y2, t.ny, iy, fy);
    d_lin_pos(z, t.Z1, t.Z2, t.nz, iz, fz);
    const long nx = t.nx, nxy = (long)t.nx * t.ny;
    const float *d = t.data + ix + iy * nx + iz * nxy;
    double c00 = (double)d[0] * (1.0 - fx) + (double)d[1] * fx;
    double c10 = (double)d[nx] * (1.0 - fx) + (double)d[nx + 1] * fx;
    double c01 = (double)d[nxy] * (1.0 - fx) + (double)d[nxy + 1] * fx;
    double c11 = (double)d[nxy + nx] * (1.0 - fx) + (double)d[nxy + nx + 1] * fx;
    double c0 = c00 * (1.0 - fy) + c10 * fy;
    double c1 = c01 * (1.0 - fy) + c11 * fy;
    return c0 * (1.0 - fz) + c1 * fz;
}

// Radius searches over the shell table.  rad(s) = sqrt(r2(s)) is strictly
// increasing and r2 is an integer, so shell_ge[k] (the first shell with
// r2 >= k) taken at k = floor(x^2) - 1 starts at or below the answer and a
// few exact comparisons on rad finish it; the index found is the one of the
// reference's linear scans.
// First s in [lo, hi] (1-based) with rad(s) > x, or hi+1.
__device__ __forceinline__ int first_rad_gt(const PostConst &C, int lo, int hi, double x)
{
    if (lo > hi) return lo;
    if (x != x) return hi + 1;
    int s = 1;
    if (x > 0.0) {
        const double xx = x * x;
        const int kk = xx >= (double)C.r2max ? C.r2max + 1 : (int)xx;
        s = C.shell_ge[kk > 0 ? kk - 1 : 0];
    }
    while (s <= C.nshells && C.rad[s - 1] <= x) s++;
    return s < lo ? lo : (s > hi ? hi + 1 : s);
}
// Last s in [lo, hi] with rad(s) < x, or lo-1.
__device__ __forceinline__ int last_rad_lt(const PostConst &C, int lo, int hi, double x)
{
    if (lo > hi) return hi;
    if (x != x) return lo - 1;
    int s = 1;  // first shell with rad >= x
    if (x > 0.0) {
        const double xx = x * x;
        const int kk = xx >= (double)C.r2max ? C.r2max + 1 : (int)xx;
        s = C.shell_ge[kk > 0 ? kk - 1 : 0];
        while (s <= C.nshells && C.rad[s - 1] < x) s++;
    }
    const int t = s - 1;
    return t > hi ? hi : (t < lo ? lo - 1 : t);
}

struct Strain {
    double E[6];  // Ebar a00, a11, a22, a01, a02, a12
    double g[3], gf[3];
};

// Execution widths of the post-process: W lanes work on one peak (radial
// sums split over the W lanes of an aligned lane group, combined by a fixed
// butterfly); W = 1: every lane runs its own peak serially, W = 64: one peak
// per wave.  All lanes of a group take the same branches (they hold the same
// peak), so the group's shuffles and ballots see all of its lanes.
template <int W> struct Exec {
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16 || W == 32 || W == 64, "lane-group width");
    static constexpr bool COOP = W > 1;
    __device__ static int lane() { return W == 1 ? 0 : (int)(threadIdx.x & (W - 1)); }
    __device__ static int base() { return (int)(threadIdx.x & 63) & ~(W - 1); }
    __device__ static double sum(double v)
    {
#pragma unroll
        for (int o = W / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    }
    // bits of the group's lanes (bit i: lane i of the group)
    __device__ static unsigned long long ballot(bool hit)
    {
        const unsigned long long b = __ballot(hit);
        if constexpr (W == 64) return b;
        else return (b >> base()) & ((1ull << W) - 1);
    }
    __device__ static bool leader() { return lane() == 0; }
};

// kernel_strain (RadialShell.jl:228-277).  The weight of shell m1 around
// shell mp depends on the two radii only, so it comes from a table built once
// per shell table (k_wtab); the row of mp holds exactly the shells with
// (rad(mp) - rad(m1))^2 < 4, the reference's cut, so clamping [mlow, mupp]
// to the row is the reference's skip.  KS selects the sums formed (and the
// profile inputs read): KS_E the strain Ebar, KS_G the gradients g and gf;
// a sum not selected is returned as zero.  Each sum is formed in the same
// order whatever else is selected, so its bits do not depend on KS.
enum : int { KS_E = 1, KS_G = 2 };
template <int W, int KS, class PV>
__device__ Strain kernel_strain(const PV &P, const PostConst &C, int mp, int mlow, int mupp)
{
    using X = Exec<W>;
    double E[6] = {0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0}, gf[3] = {0, 0, 0};
    const int4 row = C.wrow[mp - 1];
    const int a = max(mlow, row.y), b = min(mupp, row.z);
    // one shell's terms; Es = {SR00+SR00, SR11+SR11, SR22+SR22, SR01+SR10,
    // SR02+SR20, SR12+SR21}
    auto add_shell = [&](double w, const float *G, const float *Gf, const float *Es) {
        if constexpr ((KS & KS_G) != 0) {
#pragma unroll
            for (int L = 0; L < 3; L++) {
                g[L] += w * (double)G[L];
                gf[L] += w * (double)Gf[L];
            }
        }
        if constexpr ((KS & KS_E) != 0) {
            const double hw = 0.5 * w;
#pragma unroll
            for (int q = 0; q < 6; q++) E[q] -= hw * (double)Es[q];
        }
    };
    int m1 = a + X::lane();
    if constexpr (W == 1) {
        // One lane per peak: the lanes of a wave read 64 different peaks'
        // records, so each load instruction touches up to 64 cache lines.
        // Two shells per iteration halve the instructions: their weights
        // and records are adjacent (16 bytes and 2 x 24 bytes).  The shells
        // are still added in order.
#pragma unroll 2  // two pairs: four shells' loads in flight per lane
        for (; m1 < b; m1 += 2) {
            double w2[2];
            __builtin_memcpy(w2, C.wtab + row.x + (m1 - row.y), sizeof w2);
            float G[6], Gf[6], Es[12];
            if constexpr ((KS & KS_G) != 0) P.strain_G2(m1, G, Gf);
            if constexpr ((KS & KS_E) != 0) P.strain_E2(m1, Es);
            add_shell(w2[0], G, Gf, Es);
            add_shell(w2[1], G + 3, Gf + 3, Es + 6);
        }
    }
#pragma unroll 1
    for (; m1 <= b; m1 += W) {
        const double w = C.wtab[row.x + (m1 - row.y)];
        float G[3], Gf[3], Es[6];
        if constexpr ((KS & KS_G) != 0) P.strain_G(m1, G, Gf);
        if constexpr ((KS & KS_E) != 0) P.strain_E(m1, Es);
        add_shell(w, G, Gf, Es);
    }
    Strain s;
    const double scale = mp > 1 ? C.wRnor / P.rad(mp) : C.aRnor;
#pragma unroll
    for (int i = 0; i < 6; i++) s.E[i] = (KS & KS_E) ? X::sum(E[i]) * scale : 0.0;
#pragma unroll
    for (int L = 0; L < 3; L++) {
        s.g[L] = (KS & KS_G) ? X::sum(g[L]) * scale : 0.0;
        s.gf[L] = (KS & KS_G) ? X::sum(gf[L]) * scale : 0.0;
    }
    return s;
}

// One weight of kernel_strain: hRinteg over the pair's overlap for mp > 1,
// the atab4 lookup at the centre (mp = 1); the shell sum's 1/rad(m1)
// (RadialShell.jl:412-418) is folded in.
__device__ __forceinline__ double strain_weight(const double *rad, const double *inv_rad, const double *akk, int mp,
                                                int m1)
{
    const double rmp = rad[mp - 1], r1 = rad[m1 - 1];
    double wt;
    if (mp > 1) {
        const double u0 = rmp - r1;
        const double u02 = u0 * u0;
        const double u12 = rmp * rmp + r1 * r1;
        const double den = u12 - u02;
        wt = d_hRinteg(u0, u12) / (den * den);
    } else {
        const double u0 = r1;
        const double con = 100.0 * u0 * u0;
        const int icon = (int)floor(con);
        const double diff = con - (double)icon;
        const double aww = akk[icon] + diff * (akk[icon + 1] - akk[icon]);
        wt = 0.5 * aww / r1;
    }
    return wt * inv_rad[m1 - 1];
}

__device__ __forceinline__ bool finite6(const double E[6])
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; i++) ok = ok && isfinite(E[i]);
    return ok;
}

__device__ __forceinline__ void evals_ep(const double E[6], double &Frho, double &e_v, double &p_v)
{
    double lam[3];
    d_eig3(E, lam);
    Frho = lam[0] + lam[1] + lam[2];
    e_v = 0.0;
    p_v = 0.0;
    if (Frho > 0.0) {
        e_v = 0.5 * (lam[2] - lam[0]) / Frho;
        p_v = 0.5 * (lam[2] + lam[0] - 2.0 * lam[1]) / Frho;
    }
}

// One walk step computed ahead of the walk (the persistent form's walk
// table): kernel_strain's scaled Ebar at shell mp over the band
// [mlow, mupp & kWalkBand] it was formed on, and the step's zvir (the
// collapse-table value the walk compares with ZZon; Fbar(mp) until the
// table's second pass replaces it).  Bit 30 of mupp is set
// when shell mupp + 1 lies at exactly rad(mp) + 2: a walk that arrives with
// that shell as its band top keeps it (the reference's test is rad > rupp).
// 64 bytes: four 16-byte loads.
constexpr int32_t kWalkBand = (1 << 30) - 1;
struct WalkRec {
    double E[6];
    double zvir;
    int32_t mlow, mupp;
};

// The step's eigenvalue measures and zvir from its Ebar and Fbar(mp)
// (RadialShell.jl:600-640).
__device__ __forceinline__ double step_zvir(const PostConst &C, const double E[6], double Fb, double &Frhoh,
                                            double &e_v, double &p_v)
{
    evals_ep(E, Frhoh, e_v, p_v);
    double z = -1.0;
    if (Fb > 0.0) {
        const double poe = e_v < 1e-5 ? 0.0 : p_v / e_v;
        z = d_interp(C.tab, log10(Fb), e_v, poe);
    }
    return z;
}

// The band a walk step at mp takes when the walk reaches mp from outside:
// shells with rad in (rad(mp) - 2, rad(mp) + 2), at most m (Phase 5,
// RadialShell.jl:540-600).  The walk's own lower edge is always this one (at
// mp = 1 it keeps shell 2, the edge of every step above it: the centre is
// never inside rad > 0); its upper edge differs only where a shell lies at
// exactly rad(mp) + 2.
__device__ __forceinline__ void walk_band(const PostConst &C, int mp, int m, int &mlow, int &mupp)
{
    const double r = C.rad[mp - 1];
    mlow = mp > 1 ? mp : 2;
    if (mp > 1) {
        const int a = first_rad_gt(C, 1, mp - 1, fmax(r - 2.0, 0.0));
        if (a <= mp - 1) mlow = a;
    }
    const int b = last_rad_lt(C, mp + 1, m, r + 2.0);
    mupp = b >= mp + 1 ? b : mp;
}

template <class PV>
__device__ __forceinline__ const WalkRec &tab_rec(const PV &P, int mp)
{
    return P.wt[mp - 1];
}
template <class PV>
__device__ __forceinline__ int tab_crossing(const PV &P)
{
    return P.wmc;
}

// Ebar of a walk step: the peak's walk-table record for mp when it was formed
// on the same band (identical to kernel_strain's, which computed it), else
// kernel_strain itself.  Only the one-lane-per-peak post has a table.
template <int W, class PV>
__device__ __forceinline__ Strain walk_strain(const PV &P, const PostConst &C, int mp, int mlow, int mupp)
{
    if constexpr (W == 1) {
        if (P.wt && mp <= P.wmc) {
            const WalkRec *r = P.wt + (mp - 1);
            const int2 band = *(const int2 *)&r->mlow;
            if (band.x == mlow && (band.y & kWalkBand) == mupp) {
                Strain s;
                const double2 *e = (const double2 *)r->E;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const double2 v = e[q];
                    s.E[2 * q] = v.x;
                    s.E[2 * q + 1] = v.y;
                }
#pragma unroll
                for (int L = 0; L < 3; L++) s.g[L] = s.gf[L] = 0.0;
                return s;
            }
        }
    }
    return kernel_strain<W, KS_E>(P, C, mp, mlow, mupp);
}

__device__ __forceinline__ void write_nocollapse(const ResultView &R, int64_t o, int status)
{
    R.RTHL[o] = -1.0; R.Srb[o] = 0.0; R.Fbarx[o] = 0.0; R.e_v[o] = 0.0; R.p_v[o] = 0.0;
    R.zvir_half[o] = -1.0; R.d2F[o] = 0.0;
    for (int L = 0; L < 3; L++) {
        R.Sbar[o * 3 + L] = 0.0; R.Sbar2[o * 3 + L] = 0.0; R.gradpk[o * 3 + L] = 0.0;
        R.gradpkrf[o * 3 + L] = 0.0; R.gradpkf[o * 3 + L] = 0.0;
    }
    for (int q = 0; q < 9; q++) R.strain[o * 9 + q] = 0.0;
    R.status[o] = status;
    R.m0[o] = 0;
    R.q[o] = 0;
}

// Sum of S, S2 over shells 1..m0-1 and of lapd over cells 2..(r <= RTHL)
// with the stop at the first out-of-bounds cell, plus the mask writes
// (RadialShell.jl:756-819), from the cell table.  Used by the finish pass;
// COOP splits the cells over the lanes.
template <bool COOP>
__device__ void finish_cells(const PostConst &C, const PeakParam &pk, int m0, int q,
                             const float *const ex[3], const float *const e2[3], const ResultView &R, int64_t o)
{
    using X = Exec<COOP ? 64 : 1>;
    const int cS = C.shell_start[m0 - 1];          // cells of shells 1..m0-1
    const int cend = min(C.shell_start[q], pk.npart);
    const long n1 = C.n1, n12 = (long)C.n1 * C.n2;
    double Sb[3] = {0, 0, 0}, Sb2[3] = {0, 0, 0}, nS = 0.0;
    const bool inside = pk.flags & 1;             // the whole gather sphere is inside the grid
    double sL = 0.0, nL = 0.0;
    int firstoob = 0x7fffffff;
    const int lo1 = C.nbuff, hi1 = C.n1 - C.nbuff, lo2 = C.nbuff, hi2 = C.n2 - C.nbuff;
    const int lo3 = C.nbuff, hi3 = C.n3 - C.nbuff;
    const int cmax = max(cS, cend);
    for (int c = X::lane(); c < cmax; c += (COOP ? 64 : 1)) {
        const int pc = C.cell[c];
        const int x = pk.ci + (pc & 255) - 128, y = pk.cj + ((pc >> 8) & 255) - 128, z = pk.ck + ((pc >> 16) & 255) - 128;
        const bool inb = inside || (x >= 0 && x < C.n1 && y >= 0 && y < C.n2 && z >= 0 && z < C.n3);
        const long a = x + y * n1 + z * n12;
        if (inb && c < cS) {
            for (int L = 0; L < 3; L++) {
                Sb[L] += (double)ex[L][a];
                if (e2[0]) Sb2[L] += (double)e2[L][a];
            }
            nS += 1.0;
        }
        if (c < cend) {
            if (c >= 1) {
                if (!inb) firstoob = min(firstoob, c);
                else if (C.lapd) { sL += (double)C.lapd[a]; nL += 1.0; }
            }
            if (C.mask && x >= lo1 && x < hi1 && y >= lo2 && y < hi2 && z >= lo3 && z < hi3)
                C.mask[a] = 1;
        }
    }
    for (int L = 0; L < 3; L++) { Sb[L] = X::sum(Sb[L]); Sb2[L] = X::sum(Sb2[L]); }
    nS = X::sum(nS);
    if constexpr (COOP) {
        for (int o2 = 32; o2 >= 1; o2 >>= 1) firstoob = min(firstoob, __shfl_xor(firstoob, o2, 64));
    }
    double d2F = 0.0;
    if (C.lapd) {
        if (firstoob == 0x7fffffff) {
            sL = X::sum(sL);
            nL = X::sum(nL);
        } else {
            // stop at the first out-of-bounds cell: recount cells [1, firstoob)
            sL = 0.0; nL = 0.0;
            for (int c = 1 + X::lane(); c < firstoob; c += (COOP ? 64 : 1)) {
                const int pc = C.cell[c];
                const int x = pk.ci + (pc & 255) - 128, y = pk.cj + ((pc >> 8) & 255) - 128, z = pk.ck + ((pc >> 16) & 255) - 128;
                sL += (double)C.lapd[x + y * n1 + z * n12];
                nL += 1.0;
            }
            sL = X::sum(sL);
            nL = X::sum(nL);
        }
        if (nL > 0.0) d2F = sL / nL;
    }
    if (X::leader()) {
        for (int L = 0; L < 3; L++) {
            R.Sbar[o * 3 + L] = nS > 0.0 ? Sb[L] / nS : 0.0;
            R.Sbar2[o * 3 + L] = nS > 0.0 ? Sb2[L] / nS : 0.0;
        }
        R.d2F[o] = d2F;
    }
}

// Phase 7's sums over the shells inside the collapse (RadialShell.jl:756-819)
// from a peak's shell records: Sbar and Sbar2 over shells 1..m0-1, d2F over
// the cells of shells 2..q with the stop at the first out-of-bounds cell,
// and the mask over the cells of shells 1..q.  Inline in the post when its
// profile carries the S, S2 and lapd sums; the persistent form's small peaks
// run it in a pass of its own after their post (k_small_finish).
template <int W, class PV>
__device__ void finish_records(const PV &P, const PeakParam &pk, const PostConst &C, const ResultView &R,
                               int64_t o, int m0, int q)
{
    using X = Exec<W>;
    const int lane = X::lane();
    double Sb[3] = {0, 0, 0}, Sb2[3] = {0, 0, 0}, nS = 0.0;
    for (int m1 = 1 + lane; m1 <= m0 - 1; m1 += W) {
        double cnt, Sv[3], S2v[3];
        P.sums(m1, cnt, Sv, S2v);
        if (cnt > 0.0) {
            for (int L = 0; L < 3; L++) { Sb[L] += Sv[L]; Sb2[L] += S2v[L]; }
            nS += cnt;
        }
    }
    for (int L = 0; L < 3; L++) { Sb[L] = X::sum(Sb[L]); Sb2[L] = X::sum(Sb2[L]); }
    nS = X::sum(nS);
    if (nS > 0.0) for (int L = 0; L < 3; L++) { Sb[L] /= nS; Sb2[L] /= nS; }
    const int cend = min(C.shell_start[q], pk.npart);
    const int Rq = (int)floor(P.rad(q));
    const bool inside_q = pk.ci - Rq >= 0 && pk.ci + Rq < C.n1 && pk.cj - Rq >= 0 && pk.cj + Rq < C.n2 &&
                          pk.ck - Rq >= 0 && pk.ck + Rq < C.n3;
    if (C.mask) {
        const int lo1 = C.nbuff, hi1 = C.n1 - C.nbuff, lo2 = C.nbuff, hi2 = C.n2 - C.nbuff;
        const int lo3 = C.nbuff, hi3 = C.n3 - C.nbuff;
        for (int c = lane; c < cend; c += W) {
            int pc = C.cell[c];
            int x = pk.ci + (pc & 255) - 128, y = pk.cj + ((pc >> 8) & 255) - 128, z = pk.ck + ((pc >> 16) & 255) - 128;
            if (x < lo1 || x >= hi1 || y < lo2 || y >= hi2 || z < lo3 || z >= hi3) continue;
            C.mask[x + (long)y * C.n1 + (long)z * C.n1 * C.n2] = 1;
        }
    }
    double d2F = 0.0;
    if (P.has_lapd()) {
        if (inside_q) {
            double sL = 0.0, sn = 0.0;
            for (int m1 = 2 + lane; m1 <= q; m1 += W) { sL += P.lap(m1); sn += P.count(m1); }
            sL = X::sum(sL);
            sn = X::sum(sn);
            if (sn > 0.0) d2F = sL / sn;
        } else if (X::leader()) {
            // RadialShell.jl:804-819: stop at the first out-of-bounds cell
            double sacc = 0.0;
            long nav = 0;
            for (int c = 1; c < cend; c++) {
                int pc = C.cell[c];
                int x = pk.ci + (pc & 255) - 128, y = pk.cj + ((pc >> 8) & 255) - 128, z = pk.ck + ((pc >> 16) & 255) - 128;
                if (x < 0 || x >= C.n1 || y < 0 || y >= C.n2 || z < 0 || z >= C.n3) break;
                nav++;
                sacc += (double)C.lapd[x + (long)y * C.n1 + (long)z * C.n1 * C.n2];
            }
            d2F = nav > 0 ? sacc / (double)nav : 0.0;
        }
        if constexpr (W > 1) d2F = __shfl(d2F, X::base(), 64);
    }
    if (X::leader()) {
        for (int L = 0; L < 3; L++) { R.Sbar[o * 3 + L] = Sb[L]; R.Sbar2[o * 3 + L] = Sb2[L]; }
        R.d2F[o] = d2F;
    }
}

// The fcrit crossing of one peak, serially (RadialShell.jl:492-538): the m0
// the inward walk starts from, as analyse_post<1> finds it; 0 when there is
// none (a DUMP_NO_FCRIT peak).  Used to order the lane post's peaks.
template <class PV>
__device__ int crossing_m0(const PV &P, const PeakParam &pk)
{
    const double fcrit = pk.fcrit;
    const int m = pk.m;
    int m0 = pk.m0;
    if (P.Fbar(m0) >= fcrit) {
        for (int mm = m0; mm <= m; mm++)
            if (P.Fbar(mm) < fcrit) return mm;
        return m0;
    }
    const int mstart = m0 - 1 > 1 ? m0 - 1 : 1;
    for (int mp = mstart; mp >= 1; mp--) {
        if (P.Fbar(mp) >= fcrit) return m0;
        m0 = mp;
    }
    return 0;
}

// The per-peak analysis after the gather.  P provides rad(s), Fbar(s),
// count(s) and strain_inputs(s); with FINISH it also provides S / S2 / lap
// sums per shell and computes Sbar, Sbar2, d2F and the mask here; without
// FINISH those are left to finish_cells() and the final m0 is recorded.
// With a step budget, a walk longer than max_steps returns POST_DEFER
// before anything is written.
template <int W, bool FINISH, class PV>
__device__ int analyse_post(const PV &P, const PeakParam &pk, const PostConst &C, const ResultView &R, int64_t o,
                            int max_steps = 0x7fffffff)
{
    using X = Exec<W>;
    const int lane = X::lane();
    const int m = pk.m;
    int status = DUMP_NONE;
    double RTHL = -1.0;
    if (pk.flags & 2) { status = DUMP_PREMASK; goto nocol; }
    {
        // ---- Phase 3: gradient at Rf (RadialShell.jl:456-490) ----
        const int mrf = pk.mrf;
        double rlow_rf = fmax(P.rad(mrf) - 2.0, 0.0);
        int mlow_rf = mrf;
        if (mrf > 1) { int a = first_rad_gt(C, 1, mrf - 1, rlow_rf); if (a <= mrf - 1) mlow_rf = a; }
        int mupp_rf = mrf;
        { int b = last_rad_lt(C, mrf + 1, m, P.rad(mrf) + 2.0); if (b >= mrf + 1) mupp_rf = b; }
        Strain srf = kernel_strain<W, KS_G>(P, C, mrf, mlow_rf, mupp_rf);
        const double gradpkrf[3] = {srf.g[0], srf.g[1], srf.g[2]};

        // ---- Phase 4: fcrit crossing (RadialShell.jl:492-538) ----
        const double fcrit = pk.fcrit;
        int m0 = pk.m0, mupp;
        bool tab = false;  // the walk table (one lane per peak): crossing, bands and steps from it
        if constexpr (W == 1) tab = P.wt != nullptr;
        if (tab) {
            // the crossing crossing_m0 found ahead of the walk (0: none); an
            // outward search that found none (the crossing is pk.m0 itself)
            // keeps the band top at m
            const int mc = tab_crossing(P);
            if (mc == 0) { status = DUMP_NO_FCRIT; goto nocol; }
            const bool outward_none = mc == m0 && P.Fbar(m0) >= fcrit;
            m0 = mc;
            mupp = outward_none ? m : (tab_rec(P, m0).mupp & kWalkBand);
        } else if (P.Fbar(m0) >= fcrit) {
            int found = 0;
            if constexpr (W > 1) {
                for (int base = m0; base <= m; base += W) {
                    int mm = base + lane;
                    bool hit = mm <= m && P.Fbar(mm) < fcrit;
                    unsigned long long b = X::ballot(hit);
                    if (b) { m0 = base + __ffsll((long long)b) - 1; found = 1; break; }
                }
            } else {
                for (int mm = m0; mm <= m; mm++)
                    if (P.Fbar(mm) < fcrit) { m0 = mm; found = 1; break; }
            }
            if (!found) mupp = m;
            else {
                int b = last_rad_lt(C, m0 + 1, m, P.rad(m0) + 2.0);
                mupp = b >= m0 + 1 ? b : m0;
            }
        } else {
            int mstart = m0 - 1 > 1 ? m0 - 1 : 1;
            int found = 0;
            if constexpr (W > 1) {
                for (int top = mstart; top >= 1; top -= W) {
                    int mp = top - lane;
                    bool hit = mp >= 1 && P.Fbar(mp) >= fcrit;
                    unsigned long long b = X::ballot(hit);
                    if (b) { m0 = top - (__ffsll((long long)b) - 1) + 1; found = 1; break; }
                }
            } else {
                for (int mp = mstart; mp >= 1; mp--) {
                    if (P.Fbar(mp) >= fcrit) { found = 1; break; }
                    m0 = mp;
                }
            }
            if (!found) { status = DUMP_NO_FCRIT; goto nocol; }
            int b = last_rad_lt(C, m0 + 1, m, P.rad(m0) + 2.0);
            mupp = b >= m0 + 1 ? b : m0;
        }

        // ---- Phase 5: inward walk (RadialShell.jl:540-704) ----
        int mlow = m0;
        if (tab) {
            // walk_band's edge, except at m0 = 1, where the start keeps the centre
            if (m0 > 1) mlow = tab_rec(P, m0).mlow;
        } else {
            const double rlow = fmax(P.rad(m0) - 2.0, 0.0);
            if (m0 > 1) { int a = first_rad_gt(C, 1, m0 - 1, rlow); if (a <= m0 - 1) mlow = a; }
        }
        const double ZZon = pk.ZZon;
        Strain s0 = walk_strain<W>(P, C, m0, mlow, mupp);
        if (!finite6(s0.E)) { status = DUMP_NONFINITE; goto nocol; }
        double Frho_m0, e_v_m0, p_v_m0;
        evals_ep(s0.E, Frho_m0, e_v_m0, p_v_m0);
        double zvir1 = -1.0, zvir1p = -1.0;
        // The walk forms Ebar only.  The gradients the result keeps are those
        // of the collapse step (gradpk_last) and of the step before it
        // (gradpk_prev), formed after the walk over those two steps' bands
        // (mp, mlow, mupp), as the steps would have formed them.
        double lastE[6];                // Ebar at the collapse step
        double strain_mat[6];           // the previous step's Ebar
        int band_last[3] = {m0, mlow, mupp}, band_prev[3] = {m0, mlow, mupp};
        double Frhoh = Frho_m0, Frho = P.Fbar(m0), e_v = e_v_m0, p_v = p_v_m0;
        for (int i = 0; i < 6; i++) { strain_mat[i] = s0.E[i]; lastE[i] = s0.E[i]; }
        bool collapsed = false;
        // zvir1p at m0 is -1 (RadialShell.jl:581-583): collapsed at m0 only if ZZon <= -1
        if (-1.0 >= ZZon) {
            collapsed = true;
        } else {
            if (m0 == 1) { status = DUMP_M0_ONE; goto nocol; }
            for (int mp = m0 - 1; mp >= 1; mp--) {
                // a caller with a step budget re-runs long walks elsewhere
                if (m0 - 1 - mp >= max_steps) return POST_DEFER;
                double E[6], zv, Frhoh_mp = 0.0, e_v_mp = 0.0, p_v_mp = 0.0;
                bool from_tab = false;
                if (tab) {
                    // The table's step at mp was formed on the band this
                    // walk takes at mp (walk_band): the lower edge always,
                    // the upper edge unless the walk arrives with the shell
                    // at exactly rad(mp) + 2 as its band top, which it then
                    // keeps.  The table holds the steps down to its own first
                    // collapse, so after such a step the walk goes on without
                    // it.
                    const WalkRec &r = tab_rec(P, mp);
                    const int mu = r.mupp & kWalkBand;
                    mupp = ((r.mupp >> 30) & 1) && mupp == mu + 1 ? mupp : mu;
                    mlow = r.mlow;
                    if (mupp == mu) {
#pragma unroll
                        for (int i = 0; i < 6; i++) E[i] = r.E[i];
                        zv = r.zvir;
                        from_tab = true;
                    } else {
                        tab = false;
                    }
                } else {
                    const double rupp_mp = P.rad(mp) + 2.0;
                    const double rlow_mp = fmax(P.rad(mp) - 2.0, 0.0);
                    if (P.rad(mupp) > rupp_mp) {
                        int b = last_rad_lt(C, mp + 1, mupp, rupp_mp);
                        mupp = b >= mp + 1 ? b : mp;
                    }
                    if (mlow > 1 && P.rad(mlow - 1) > rlow_mp) {
                        int a = first_rad_gt(C, 1, mlow - 1, rlow_mp);
                        mlow = a <= mlow - 1 ? a : mlow;
                    }
                }
                if (!from_tab) {
                    const Strain smp = kernel_strain<W, KS_E>(P, C, mp, mlow, mupp);
#pragma unroll
                    for (int i = 0; i < 6; i++) E[i] = smp.E[i];
                    if (!finite6(E)) { status = DUMP_NONFINITE; goto nocol; }
                    zv = step_zvir(C, E, P.Fbar(mp), Frhoh_mp, e_v_mp, p_v_mp);
                }
                if (!finite6(E)) { status = DUMP_NONFINITE; goto nocol; }
                if (zv >= ZZon) {
                    m0 = mp + 1;
                    zvir1p = zv;
                    for (int i = 0; i < 6; i++) lastE[i] = E[i];
                    band_last[0] = mp; band_last[1] = mlow; band_last[2] = mupp;
                    if (from_tab) evals_ep(E, Frhoh_mp, e_v_mp, p_v_mp);
                    Frhoh = Frhoh_mp; Frho = P.Fbar(mp); e_v = e_v_mp; p_v = p_v_mp;
                    collapsed = true;
                    break;
                }
                zvir1 = zv;
                band_prev[0] = mp; band_prev[1] = mlow; band_prev[2] = mupp;
                for (int i = 0; i < 6; i++) strain_mat[i] = E[i];
            }
        }
        (void)Frho;
        if (!collapsed) { status = DUMP_NO_COLLAPSE; goto nocol; }
        // only reachable with ZZon <= -1 (collapse flagged at m0 itself): the
        // reference would index rad[0]; report it as the m0 == 1 dump.
        if (m0 < 2) { status = DUMP_M0_ONE; goto nocol; }

        // ---- Phase 6 (RadialShell.jl:706-754) ----
        const double r0 = P.rad(m0 - 1), r1 = P.rad(m0);
        const double rad3p = r0 * r0 * r0, rad3 = r1 * r1 * r1;
        double dZvir = zvir1p - zvir1;
        if (zvir1 > 0.0 && dZvir != 0.0) {
            double RTHL3 = rad3p + (rad3 - rad3p) * (zvir1p - ZZon) / dZvir;
            RTHL = pow(RTHL3, 1.0 / 3.0);
        } else {
            RTHL = r0;
        }
        if (RTHL <= 0.0) { status = DUMP_RTHL_NEG; goto nocol; }
        const double RTHL3 = RTHL * RTHL * RTHL, RTHL5 = RTHL3 * RTHL * RTHL;
        const double drad3 = rad3 - rad3p;
        const double dFbar = P.Fbar(m0) - P.Fbar(m0 - 1);
        const Strain last_g = kernel_strain<W, KS_G>(P, C, band_last[0], band_last[1], band_last[2]);
        double gradpk_prev[3], gradpkf_prev[3];
        if (band_prev[0] == band_last[0] && band_prev[1] == band_last[1] && band_prev[2] == band_last[2]) {
            for (int L = 0; L < 3; L++) { gradpk_prev[L] = last_g.g[L]; gradpkf_prev[L] = last_g.gf[L]; }
        } else {
            const Strain t = kernel_strain<W, KS_G>(P, C, band_prev[0], band_prev[1], band_prev[2]);
            for (int L = 0; L < 3; L++) { gradpk_prev[L] = t.g[L]; gradpkf_prev[L] = t.gf[L]; }
        }
        double Fbarx, smat[6], gpk[3], gpkf[3];
        if (zvir1 > 0.0 && drad3 != 0.0) {
            double frac = (RTHL3 - rad3p) / drad3;
            Fbarx = P.Fbar(m0 - 1) + frac * dFbar;
            for (int i = 0; i < 6; i++) smat[i] = lastE[i] + frac * (strain_mat[i] - lastE[i]);
            for (int L = 0; L < 3; L++) {
                gpk[L] = last_g.g[L] + frac * (gradpk_prev[L] - last_g.g[L]);
                gpkf[L] = last_g.gf[L] + frac * (gradpkf_prev[L] - last_g.gf[L]);
            }
        } else {
            Fbarx = P.Fbar(m0 - 1);
            for (int i = 0; i < 6; i++) smat[i] = lastE[i];
            for (int L = 0; L < 3; L++) { gpk[L] = gradpk_prev[L]; gpkf[L] = gradpkf_prev[L]; }
        }
        {   // normalize_strain (RadialShell.jl:209-212)
            double tr = smat[0] + smat[1] + smat[2];
            if (tr != 0.0) { double f = Fbarx / tr; for (int i = 0; i < 6; i++) smat[i] *= f; }
        }
        if (!finite6(smat)) { status = DUMP_NONFINITE; goto nocol; }
        double Frhoc, e_v_f, p_v_f;
        evals_ep(smat, Frhoc, e_v_f, p_v_f);
        const double sfac = Fbarx / Frhoh;

        // ---- Phase 7 (RadialShell.jl:756-819) ----
        double Srb = 0.0;
        if constexpr (W == 1) {
            // serially, the inner edge's Fbar and rad^5 are the previous shell's outer ones
            double fa = P.Fbar(1), r5a = 0.0;
            for (int m1 = 2; m1 <= m0 - 1; m1++) {
                const double rb = P.rad(m1), fb = P.Fbar(m1);
                const double r5b = rb * rb * rb * rb * rb;
                Srb += 0.5 * (fa + fb) * (r5b - r5a);
                fa = fb;
                r5a = r5b;
            }
        } else {
            for (int m1 = 2 + lane; m1 <= m0 - 1; m1 += W) {
                double ra = P.rad(m1 - 1), rb = P.rad(m1);
                double r5a = m1 == 2 ? 0.0 : ra * ra * ra * ra * ra, r5b = rb * rb * rb * rb * rb;
                Srb += 0.5 * (P.Fbar(m1 - 1) + P.Fbar(m1)) * (r5b - r5a);
            }
            Srb = X::sum(Srb);
        }
        if (zvir1 > 0.0 && dZvir != 0.0) {
            double rp = m0 - 1 >= 2 ? P.rad(m0 - 1) : 0.0;
            double rad5p = rp * rp * rp * rp * rp;
            Srb += 0.5 * (P.Fbar(m0 - 1) + Fbarx) * (RTHL5 - rad5p);
        }
        if (Fbarx * RTHL5 > 0.0) Srb /= (Fbarx * RTHL5);
        // cells with r <= RTHL: shells 1..q
        const int q = last_rad_lt(C, 1, m, nextafter(RTHL, 1e300));

        if constexpr (FINISH) finish_records<W>(P, pk, C, R, o, m0, q);

        // ---- zvir_half (RadialShell.jl:821-885) ----
        double zvir_half = -1.0;
        if (RTHL >= 3.0) {
            // neighbouring radii often land on the same (m0_rc, mlow_rc,
            // mupp_rc) band; its zvir_rc is then the one already in the max
            int band_prev[3] = {-1, -1, -1};
            for (int jj = 1; jj <= 10; jj++) {
                double rcur = RTHL * pow((double)(jj - 1) / 9.0 * 0.5 + 0.5, 1.0 / 3.0);
                double rupp_rc = rcur + 2.0;
                int mupp_rc = 1;
                if (m - 1 >= 2) {
                    int f = first_rad_gt(C, 2, m - 1, rupp_rc);
                    mupp_rc = f <= m - 1 ? f : m - 1;
                }
                double rlow_rc = fmax(rcur - 2.0, 0.0);
                int mlow_rc = 1;
                if (m - 1 >= 1) { int f = first_rad_gt(C, 1, m - 1, rlow_rc); if (f <= m - 1) mlow_rc = f; }
                int m0_rc = 1;
                if (m - 1 >= 1) {
                    int f = first_rad_gt(C, 1, m - 1, rcur);
                    m0_rc = f <= m - 1 ? f : m - 1;
                }
                if (m0_rc == band_prev[0] && mlow_rc == band_prev[1] && mupp_rc == band_prev[2]) continue;
                band_prev[0] = m0_rc; band_prev[1] = mlow_rc; band_prev[2] = mupp_rc;
                Strain src = kernel_strain<W, KS_E>(P, C, m0_rc, mlow_rc, mupp_rc);
                if (!finite6(src.E)) { status = DUMP_NONFINITE; goto nocol; }
                double Frho_rc, e_rc, p_rc;
                evals_ep(src.E, Frho_rc, e_rc, p_rc);
                double Frhoc_rc = P.Fbar(m0_rc);
                double poe_rc = e_rc < 1e-5 ? 0.0 : p_rc / e_rc;
                double zvir_rc = -1.0;
                if (Frhoc_rc > 0.0) zvir_rc = d_interp(C.tab, log10(Frhoc_rc), e_rc, poe_rc);
                zvir_half = fmax(zvir_rc - 1.0, zvir_half);
            }
        }

        if (X::leader()) {
            R.RTHL[o] = RTHL;
            R.Srb[o] = Srb;
            for (int L = 0; L < 3; L++) {
                R.gradpk[o * 3 + L] = gpk[L];
                R.gradpkrf[o * 3 + L] = gradpkrf[L];
                R.gradpkf[o * 3 + L] = gpkf[L];
            }
            // strain_mat = Ebar_last * (Fbarx / Frhoh), column-major 3x3
            const double *E = lastE;
            double *S = R.strain + o * 9;
            S[0] = E[0] * sfac; S[4] = E[1] * sfac; S[8] = E[2] * sfac;
            S[1] = S[3] = E[3] * sfac;
            S[2] = S[6] = E[4] * sfac;
            S[5] = S[7] = E[5] * sfac;
            R.Fbarx[o] = Fbarx;
            R.e_v[o] = e_v_f;
            R.p_v[o] = p_v_f;
            R.zvir_half[o] = zvir_half;
            R.status[o] = DUMP_NONE;
            R.m0[o] = m0;
            R.q[o] = q;
        }
        return DUMP_NONE;
    }
nocol:
    if (X::leader()) write_nocollapse(R, o, status);
    return status;
}

}  // namespace pp
