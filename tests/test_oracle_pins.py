"""Pins of the FP64 oracle against the reference's own known answers, golden
vectors and closed forms (CPU only)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---- hRinteg (test/test_radialshell.jl:3-33) ----
def test_hRinteg_closed_forms():
    assert O.hRinteg(0.0, 0.0) == pytest.approx(-31.0 / 35.0, abs=1e-12)
    assert O.hRinteg(0.0, 1.5) == pytest.approx(1.4 * 1.5 - 31.0 / 35.0, abs=1e-12)
    assert O.hRinteg(2.0, 0.0) == 0.0
    assert O.hRinteg(3.0, 1.0) == 0.0
    assert O.hRinteg(1.0, 0.0) == pytest.approx(-3.2 / 7.0, abs=1e-12)
    assert O.hRinteg(0.5, 2.0) == pytest.approx(O.hRinteg(-0.5, 2.0))
    assert O.hRinteg(0.5, 1.0) != O.hRinteg(0.5, 2.0)


def test_hRinteg_is_the_kernel_integral():
    # hRinteg(u0, u12) = int_{|u0|}^2 4 pi W(x) x (u12 - x^2) dx with the M4
    # kernel W = (1/pi)(1 - 1.5x^2 + 0.75x^3) (x<1), (1/4pi)(2-x)^3 (1<=x<2):
    # brute-force quadrature pins every polynomial coefficient.
    def W(x):
        return (1 - 1.5 * x * x + 0.75 * x ** 3) / math.pi if x < 1 else (2 - x) ** 3 / (4 * math.pi)
    for u0 in (0.0, 0.3, 0.99, 1.0, 1.4, 1.95):
        for u12 in (0.0, 0.7, 3.0):
            xs = np.linspace(abs(u0), 2.0, 200001)
            f = np.array([4 * math.pi * W(x) * x * (u12 - x * x) for x in xs])
            quad = np.trapezoid(f, xs)
            assert O.hRinteg(u0, u12) == pytest.approx(quad, abs=2e-9)


# ---- atab4 (test/test_radialshell.jl:36-52) ----
def test_atab4():
    akk = O.atab4()
    assert len(akk) == 1101
    ca = 3.0 / (2.0 * math.pi)
    assert akk[0] == pytest.approx(ca * 2.0 / 3.0, abs=1e-14)
    assert akk[400] == 0.0 and akk[499] == 0.0 and akk[1099] == 0.0
    cb = 1.0 / (4.0 * math.pi)
    assert ca * (2.0 / 3.0 - 1.0 + 0.5) == pytest.approx(cb * 1.0, abs=1e-14)
    assert akk[100] == pytest.approx(cb, abs=1e-14)


# ---- precompute_shells (test/test_radialshell.jl:54-71) + brute force ----
def test_precompute_shells_counts():
    s1 = O.Shells(1)
    assert len(s1) == 7 and s1.r2[0] == 0 and (s1.r2 <= 1).all()
    s2 = O.Shells(2)
    assert len(s2) == 33 and (np.diff(s2.r2) >= 0).all()


@pytest.mark.parametrize("rmax,cells,shells", [(23, 50883, 443), (26, 73525, 566), (44, 356637, 1616)])
def test_precompute_shells_brute_force(rmax, cells, shells):
    s = O.Shells(rmax)
    r = np.arange(-rmax, rmax + 1)
    q = (r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2).ravel()
    assert len(s) == (q <= rmax * rmax).sum() == cells
    assert len(np.unique(s.r2)) == shells
    # stable: inside one r^2 the generation order (dk, dj, di) is kept
    key = (s.dk.astype(np.int64) * 1000 + s.dj) * 1000 + s.di
    for v in np.unique(s.r2)[:50]:
        sel = key[s.r2 == v]
        assert (np.diff(sel) > 0).all()
    assert ((s.di ** 2 + s.dj ** 2 + s.dk ** 2) == s.r2).all()


# ---- get_evals / normalize_strain (test/test_radialshell.jl:73-93) ----
def test_get_evals_known():
    assert np.allclose(O.get_evals(np.eye(3)), [1, 1, 1], atol=1e-12)
    assert np.allclose(O.get_evals(np.diag([1.0, 2.0, 3.0])), [1, 2, 3], atol=1e-12)


def test_get_evals_vs_lapack():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = rng.normal(size=(3, 3)) * 10 ** rng.uniform(-3, 3)
        a = a + a.T
        assert np.allclose(O.get_evals(a), np.linalg.eigvalsh(a), rtol=1e-12, atol=1e-12 * np.abs(a).max())
    # near-degenerate and exactly degenerate cases
    for a in (np.diag([2.0, 2.0, 2.0 + 1e-12]), np.ones((3, 3)), np.array([[4, 1e-9, 0], [1e-9, 4, 0], [0, 0, -1.0]])):
        assert np.allclose(O.get_evals(a), np.linalg.eigvalsh(a), atol=1e-11)


def test_normalize_strain():
    S = O.normalize_strain(np.diag([1.0, 2.0, 3.0]), 12.0)
    assert np.trace(S) == pytest.approx(12.0) and np.allclose(np.diag(S), [2, 4, 6])


# ---- windows (test/test_filters.jl:2-38) ----
def test_windows():
    assert O.gaussian_window(0.0) == pytest.approx(1.0)
    for kR in (0.1, 0.5, 1.0, 2.0, 5.0):
        assert O.gaussian_window(kR) == pytest.approx(math.exp(-kR ** 2 / 2), rel=1e-10)
    assert O.gaussian_window(10.0) < 1e-20
    for kR in (0.0, 0.1, 0.5, 1.0, 2.0, 5.0):
        assert O.gaussian_window_fortran(kR) == pytest.approx(math.exp(-kR ** 2 / 8), rel=1e-10)
    assert O.tophat_window(0.0) == pytest.approx(1.0)
    for kR in (0.01, 0.1, 0.5, 1.0, 2.0, 5.0):
        assert O.tophat_window(kR) == pytest.approx(3 * (math.sin(kR) - kR * math.cos(kR)) / kR ** 3, rel=1e-10)
    assert O.tophat_window(4.0) > 0 and O.tophat_window(5.0) < 0
    assert O.tophat_window(0.01) > 0.999


def test_window_multiply_smooth_field_invariants():
    # test_pipeline.jl:66-99: a delta spike smoothed by a top-hat peaks at the
    # spike and is isotropic; the DC mode is untouched.
    n, box = 32, 32.0
    d = np.zeros((n, n, n), np.float32)
    d[16, 16, 16] = 1.0
    dk = np.fft.rfftn(d).astype(np.complex64)
    s = O.smooth_field(dk, n, box, 3.0, 1)
    assert np.unravel_index(np.argmax(s), s.shape) == (16, 16, 16)
    assert s[16, 16, 19] == pytest.approx(s[16, 19, 16], rel=1e-6) == pytest.approx(s[19, 16, 16], rel=1e-6)
    assert s.sum() == pytest.approx(1.0, rel=1e-5)
    w = O.window_multiply(dk, n, box, 3.0, 1)
    assert w[0, 0, 0] == dk[0, 0, 0]


# ---- find_peaks known answers (test/test_peakfind.jl:4-118) ----
def _fp(delta, mask=None, nbuff=1, fcrit=1.0, max_peaks=2_000_000):
    mask = np.zeros(delta.shape, np.int8) if mask is None else mask
    return O.find_peaks(delta, mask, nbuff, fcrit, max_peaks), mask


def test_find_peaks_known_answers():
    z = lambda n: np.zeros((n, n, n), np.float32)
    d = z(11); d[5, 5, 5] = 2.0
    (i, j, k, v), m = _fp(d)
    assert list(zip(i, j, k)) == [(6, 6, 6)] and v[0] == 2.0 and m[5, 5, 5] == 1
    (i, *_), _ = _fp(np.full((11, 11, 11), 0.5, np.float32))
    assert len(i) == 0
    g = np.fromfunction(lambda k, j, i: i + j + k + 3, (11, 11, 11)).astype(np.float32)
    (i, *_), _ = _fp(g)
    assert len(i) == 0
    d = z(11); d[5, 5, 5] = 2.0; mask = np.zeros_like(d, np.int8); mask[5, 5, 5] = 1
    (i, *_), _ = _fp(d, mask)
    assert len(i) == 0
    d = z(11); d[5, 5, 0] = 2.0
    (i, *_), _ = _fp(d)
    assert len(i) == 0
    d = z(15); d[3, 3, 3] = 2.0; d[11, 11, 11] = 3.0
    (i, j, k, _), _ = _fp(d)
    assert list(zip(i, j, k)) == [(4, 4, 4), (12, 12, 12)]
    d = z(11); d[6, 4, 2] = 2.0          # Julia delta[3,5,7]
    (i, j, k, _), _ = _fp(d)
    assert (i[0] + (j[0] - 1) * 11 + (k[0] - 1) * 121) == 3 + 4 * 11 + 6 * 121
    d = z(15)
    for (a, b, c) in [(4, 4, 4), (4, 4, 12), (12, 12, 4), (12, 12, 12)]:
        d[c - 1, b - 1, a - 1] = 2.0
    (i, *_), _ = _fp(d, max_peaks=2)
    assert len(i) == 2


def test_find_peaks_plateau_ties_allowed():
    # strict '<' test: equal neighbours do not veto (PeakFind.jl:52)
    d = np.zeros((9, 9, 9), np.float32)
    d[4, 4, 4] = d[4, 4, 5] = 2.0
    (i, j, k, _), _ = _fp(d)
    assert list(zip(i, j, k)) == [(5, 5, 5), (6, 5, 5)]


def test_find_peaks_brute_force():
    rng = np.random.default_rng(3)
    d = rng.normal(size=(12, 13, 14)).astype(np.float32)
    (i, j, k, v), _ = _fp(d, nbuff=1, fcrit=0.5)
    exp = []
    n3, n2, n1 = d.shape
    for kk in range(1, n3 - 1):
        for jj in range(1, n2 - 1):
            for ii in range(1, n1 - 1):
                f = d[kk, jj, ii]
                if f >= 0.5 and f >= d[kk - 1:kk + 2, jj - 1:jj + 2, ii - 1:ii + 2].max():
                    exp.append((ii + 1, jj + 1, kk + 1))
    assert list(zip(i, j, k)) == exp


# ---- collapse table (test/test_ellipsoidalcollapse.jl:127-152) ----
def test_trilinear_exact_on_linear_table():
    g = np.arange(11) * 0.1
    t = (g[:, None, None] + g[None, :, None] + g[None, None, :]).astype(np.float32)
    tab = O.Table(t, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    assert tab.interp(0, 0, 0) == pytest.approx(0.0, abs=1e-5)
    assert tab.interp(0.5, 0, 0) == pytest.approx(0.5, abs=1e-5)
    assert tab.interp(0, 0.5, 0) == pytest.approx(0.5, abs=1e-5)
    assert tab.interp(0.25, 0.25, 0.25) == pytest.approx(0.75, abs=1e-4)
    assert tab.interp(1.0, 1.0, 1.0) == pytest.approx(3.0, abs=1e-5)
    assert tab.interp(-0.1, 0, 0) == -1.0 and tab.interp(1.1, 0, 0) == -1.0


def test_homeltab_fixture_layout(homeltab_path):
    tab = O.Table.read_homeltab(homeltab_path)
    assert (tab.nx, tab.ny, tab.nz) == (50, 20, 20)
    assert tab.bounds[0] == pytest.approx(math.log10(1.5), rel=1e-6)
    assert tab.bounds[1] == pytest.approx(math.log10(8.0), rel=1e-6)
    assert (tab.flat == -1).sum() > 1000 and tab.flat.max() > 1


# ---- analyse_peak: Fortran get_homel golden vector (test_radialshell_fortran.jl) ----
def fortran_grid():
    n = 142
    z = lambda: np.zeros((n, n, n), np.float32)
    delta, ex, ey, ez = z(), z(), z(), z()
    def S(a, i, j, k, v):
        a[k - 1, j - 1, i - 1] = v
    S(delta, 71, 71, 71, 10.0)
    for p in [(70, 71, 71), (72, 71, 71), (71, 70, 71), (71, 72, 71), (71, 71, 70), (71, 71, 72)]:
        S(delta, *p, 5.0)
    S(ex, 71, 71, 71, .1); S(ex, 70, 71, 71, .05); S(ex, 72, 71, 71, .15)
    for p in [(71, 70, 71), (71, 72, 71), (71, 71, 70), (71, 71, 72)]:
        S(ex, *p, .1)
    S(ey, 71, 71, 71, .2); S(ey, 70, 71, 71, .2); S(ey, 72, 71, 71, .2)
    S(ey, 71, 70, 71, .1); S(ey, 71, 72, 71, .3); S(ey, 71, 71, 70, .2); S(ey, 71, 71, 72, .2)
    S(ez, 71, 71, 71, .3)
    for p in [(70, 71, 71), (72, 71, 71), (71, 70, 71), (71, 72, 71)]:
        S(ez, *p, .3)
    S(ez, 71, 71, 70, .15); S(ez, 71, 71, 72, .45)
    return delta, ex, ey, ez


FORTRAN = dict(fcrit=1.7161525487899780, RTHL=1.7320507764816284, Srb=1.9020472764968872,
               Fbarx=2.3230779170989990, Sbar=(2.5925926864147186e-02, 5.1851853728294373e-02, 7.7777773141860962e-02),
               strain=(3.8717970252037048e-01, 7.7435940504074097e-01, 1.1615388393402100e+00),
               e_v=1.6666662693023682e-01, p_v=-5.1315236504478889e-08, zvir_half=-1.0)


def check_fortran(r, rel=0.01):
    F = FORTRAN
    assert r["RTHL"] > 0 and abs(r["RTHL"] - F["RTHL"]) / F["RTHL"] < rel
    assert abs(r["Srb"] - F["Srb"]) / F["Srb"] < rel
    assert abs(r["Fbarx"] - F["Fbarx"]) / F["Fbarx"] < rel
    for a, b in zip(r["Sbar"], F["Sbar"]):
        assert abs(a - b) < 0.01
    assert np.abs(r["Sbar2"]).max() < 0.01
    st = np.asarray(r["strain"]).reshape(3, 3)
    for L in range(3):
        assert abs(st[L, L] - F["strain"][L]) / F["strain"][L] < rel
    assert np.abs(st - np.diag(np.diag(st))).max() < 0.01
    assert np.abs(r["gradpk"]).max() < 0.01 and np.abs(r["gradpkrf"]).max() < 0.01
    assert abs(r["e_v"] - F["e_v"]) / F["e_v"] < rel
    assert abs(r["p_v"] - F["p_v"]) < 0.01
    assert abs(r["zvir_half"] - F["zvir_half"]) < 0.01


def test_analyse_peak_fortran_golden(homeltab_path):
    d, ex, ey, ez = fortran_grid()
    r, st, _ = O.analyse_peak(O.Grid(d, ex, ey, ez), 71 + 70 * 142 + 70 * 142 * 142, 1.0, 4, 1.0, 4.0,
                              O.Table.read_homeltab(homeltab_path), O.Shells(8), fcrit=FORTRAN["fcrit"])
    assert st == 0
    check_fortran(r)
    # tighter than the reference's 1%: the values are Fortran single-precision
    assert r["RTHL"] == pytest.approx(math.sqrt(3.0), rel=1e-12)
    assert r["Fbarx"] == pytest.approx(FORTRAN["Fbarx"], rel=1e-6)


def test_analyse_peak_synthetic_collapse_and_no_collapse():
    # test_radialshell.jl:95-156
    n = 20
    d = np.zeros((n, n, n), np.float32)
    d[8:11, 8:11, 8:11] = 5.0
    d[9, 9, 9] = 10.0
    z = np.zeros_like(d)
    one = O.Table(np.full((5, 3, 3), 1.0, np.float32), math.log10(1.5), math.log10(8.0), 0.0, 0.5, -0.9999, 0.9999)
    ipp = 10 + 9 * n + 9 * n * n
    r, st, _ = O.analyse_peak(O.Grid(d, z, z, z), ipp, 1.0, 4, 0.0, 4.0, one, O.Shells(8))
    assert st == 0 and r["RTHL"] > 0 and r["Srb"] > 0
    hi = O.Table(np.full((5, 3, 3), 100.0, np.float32), math.log10(1.5), math.log10(8.0), 0.0, 0.5, -0.9999, 0.9999)
    r, st, _ = O.analyse_peak(O.Grid(np.full((n, n, n), 0.5, np.float32), z, z, z), ipp, 1.0, 4, 0.0, 4.0, hi, O.Shells(8))
    assert r["RTHL"] == -1.0 and st != 0


def test_analyse_peak_shell_profile_brute_force():
    # nshell / Fbar of the gather loop against a direct per-r^2 computation
    rng = np.random.default_rng(11)
    n = 24
    d = rng.normal(size=(n, n, n)).astype(np.float32)
    z = np.zeros_like(d)
    sh = O.Shells(6)
    tab = O.Table(np.full((5, 5, 5), 2.0, np.float32), -2, 2, 0, 1, 0, 2)
    for (ci, cj, ck) in [(12, 12, 12), (2, 3, 20)]:
        ipp = ci + 1 + cj * n + ck * n * n
        _, _, prof = O.analyse_peak(O.Grid(d, z, z, z), ipp, 1.0, 4, 1.0, 2.0, tab, sh, profile_cap=200)
        m = prof["m"]
        r2v = np.unique(sh.r2)
        assert m == len(r2v) - 1           # the last shell is never finalised
        fb_prev, dfp = float(d[ck, cj, ci]), float(d[ck, cj, ci])
        assert prof["Fbar"][0] == fb_prev
        for s in range(2, m + 1):
            sel = sh.r2 == r2v[s - 1]
            x, y, zz = ci + sh.di[sel], cj + sh.dj[sel], ck + sh.dk[sel]
            ok = (x >= 0) & (x < n) & (y >= 0) & (y < n) & (zz >= 0) & (zz < n)
            assert prof["nshell"][s - 1] == ok.sum()
            F = d[zz[ok], y[ok], x[ok]].astype(np.float64).sum()
            dF = F / ok.sum() if ok.sum() else 0.0
            ra, rb = math.sqrt(r2v[s - 2]), math.sqrt(r2v[s - 1])
            fb = (ra ** 3 * fb_prev + 0.5 * (dfp + dF) * (rb ** 3 - ra ** 3)) / rb ** 3
            assert prof["Fbar"][s - 1] == pytest.approx(fb, rel=1e-12, abs=1e-12)
            fb_prev, dfp = fb, dF


# ---- cosmology (test/test_cosmology.jl, test_ellipsoidalcollapse.jl:40-69) ----
def test_growth_tables():
    dl = O.Dlinear(0.315, 0.049, 0.685, 0.674)
    assert dl.ab(1.0)[0] == pytest.approx(1.0, abs=2e-4)
    assert dl.fsc_of_z(0.0) == pytest.approx(1.686, abs=4e-4)
    d05 = dl.ab(1 / 1.5)[0]
    assert 0.7 < d05 < 0.8 and dl.fsc_of_z(0.5) == pytest.approx(1.686 * d05)
    # Einstein-de Sitter: D(a) = a exactly
    eds = O.Dlinear(1.0, 0.0, 0.0, 0.7)
    for a in (0.01, 0.1, 0.5, 0.9):
        assert eds.ab(a)[0] == pytest.approx(a, rel=2e-4)
        assert eds.ab(a)[1] == pytest.approx(1.0, rel=1e-4)


def test_chi_and_chi_to_z():
    Om, OL, h = 0.315, 0.685, 0.674
    # EdS closed form chi = 2998 h * 2 (1 - 1/sqrt(1+z))
    assert O.chi(1.0, 1.0, 0.0, 0.7) == pytest.approx(2998 * 0.7 * 2 * (1 - 1 / math.sqrt(2)), rel=1e-10)
    c2z = O.ChiToZ(Om, OL, h, z_max=5.0)
    for z in (0.01, 0.3, 1.0, 2.5, 4.5):
        assert c2z(O.chi(z, Om, OL, h)) == pytest.approx(z, abs=5e-3)
    assert c2z(-1.0) == 0.0 and c2z(1e9) == 5.0
    assert c2z.peak_redshift((0, 0, 0), (0, 0, O.chi(0.5, Om, OL, h))) == pytest.approx(0.5, abs=5e-3)


# ---- analyse_peak outputs no reference vector fixes: analytic cases ----
# (tests/pin_cases.py states each expectation and why it holds)
import pin_cases as PC  # noqa: E402


def _run_case(case, table, bounds, fcrit, ZZon, mask=None):
    g = O.Grid(case.delta, *case.eta, eta2=case.eta2, mask=mask, lapd=case.lapd)
    r, st, prof = O.analyse_peak(g, case.ipp, 1.0, case.ir2min, ZZon, 4.0, O.Table(table, *bounds),
                                 O.Shells(case.rmax), nbuff=case.nbuff, rmax2rs=0.0, fcrit=fcrit,
                                 profile_cap=2000)
    return r, st, prof


def _check_common(case, r, R, Fbarx, inside_r2):
    ex = PC.expect_common(case, R, Fbarx, inside_r2)
    st = np.asarray(r["strain"]).reshape(3, 3, order="F")
    assert np.allclose(st, ex["strain"], rtol=1e-10, atol=1e-12)
    assert r["e_v"] == pytest.approx(ex["e_v"], rel=1e-10)
    assert r["p_v"] == pytest.approx(ex["p_v"], rel=1e-9)
    # kernel-weighted gradient of a linear field: its slope, up to the
    # kernel's discretisation error on the lattice
    assert np.allclose(r["gradpk"], ex["gradpk"], rtol=1e-4)
    # quirk Q6: the permuted read rotates the slopes
    assert np.allclose(r["gradpkf"], ex["gradpkf"], rtol=1e-4)
    assert np.allclose(r["gradpkf"], np.asarray(r["gradpk"])[[1, 2, 0]], rtol=1e-9)
    assert np.allclose(r["Sbar2"], ex["Sbar2"], rtol=1e-12)
    assert r["d2F"] == pytest.approx(ex["d2F"], rel=1e-12)


@pytest.mark.parametrize("fcrit", [1.0, 1.686, 3.0])
def test_pin_constant_table_collapse(fcrit):
    """Every query collapses: R_TH = radius of the last shell with
    Fbar >= fcrit, Fbarx = its Fbar, zvir_half = table value - 1, and the mask
    holds exactly the cells within R_TH inside the nbuff frame."""
    case = PC.bump_case()
    t, b = PC.constant_table(1.8)
    mask = np.zeros(case.delta.shape, np.int8)
    r, st, prof = _run_case(case, t, b, fcrit, 1.5, mask)
    assert st == 0
    rad, fb = PC.fbar_profile(case)
    mc = PC.crossing(fb, fcrit)
    assert r["RTHL"] == rad[mc - 1] and r["RTHL"] >= 3.0
    assert r["Fbarx"] == pytest.approx(fb[mc - 1], rel=1e-12)
    assert r["zvir_half"] == pytest.approx(float(np.float32(1.8)) - 1.0, rel=1e-12)
    _check_common(case, r, r["RTHL"], r["Fbarx"], rad[mc - 1] ** 2)
    sel, _ = PC.inside_cells(case, r["RTHL"], frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel)


@pytest.mark.parametrize("a,b,fcrit,ZZon", [(1.0, 1.0, 1.0, 1.5), (1.2, 0.8, 1.686, 1.7), (0.9, 1.5, 1.0, 2.0)])
def test_pin_linear_table_interpolates_rthl(a, b, fcrit, ZZon):
    """zvir = a + b log10 F: the walk stops at the outermost shell inside the
    fcrit crossing with zvir >= 1 + z, R_TH interpolates zvir linearly in r^3
    to 1 + z, Fbarx interpolates Fbar the same way, and zvir_half is the
    largest zvir - 1 over ten radii from R_TH / 2^(1/3) to R_TH."""
    case = PC.bump_case()
    t, bounds = PC.linear_table(a, b)
    mask = np.zeros(case.delta.shape, np.int8)
    r, st, prof = _run_case(case, t, bounds, fcrit, ZZon, mask)
    assert st == 0
    ex = PC.expect_linear(case, a, b, fcrit, ZZon)
    rad, fb = PC.fbar_profile(case)
    assert prof["m0"] == ex["inside"] + 1
    assert r["RTHL"] == pytest.approx(ex["RTHL"], rel=1e-12) and r["RTHL"] != rad[ex["inside"]]
    assert r["Fbarx"] == pytest.approx(ex["Fbarx"], rel=1e-12)
    assert r["zvir_half"] == pytest.approx(ex["zvir_half"], rel=1e-6, abs=1e-6)
    _check_common(case, r, r["RTHL"], r["Fbarx"], rad[ex["inside"]] ** 2)
    sel, _ = PC.inside_cells(case, r["RTHL"], frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel)


def test_pin_d2F_stops_at_the_grid_edge_and_mask_frame():
    """A peak 6 cells from the x = 0 face with R_TH > 6: d2F averages lapd
    over cells 2.. in shell order and stops at the first cell outside the
    grid (quirk Q7); including the centre, or skipping instead of stopping,
    gives another value.  The mask still holds exactly the cells within
    R_TH inside the frame."""
    case = PC.bump_case(centre=(6, 23, 23))
    t, b = PC.constant_table(1.8)
    mask = np.zeros(case.delta.shape, np.int8)
    r, st, _ = _run_case(case, t, b, 1.0, 1.5, mask)
    assert st == 0 and r["RTHL"] > 7
    sh = O.Shells(case.rmax)
    exp = PC.d2F_walk(case, r["RTHL"], sh.di, sh.dj, sh.dk, sh.r2)
    assert r["d2F"] == pytest.approx(exp, rel=1e-12)
    sel, q = PC.inside_cells(case, r["RTHL"], exclude_centre=True)
    assert abs(case.lapd.astype(np.float64)[sel].mean() - exp) > 1e-3   # skipping would differ
    sel_f, _ = PC.inside_cells(case, r["RTHL"], frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel_f)
