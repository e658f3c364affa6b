"""GPU parity of the shell analysis (analyse_peak) against the oracle."""
import math

import numpy as np
import pytest

from conftest import ipp1
from oracle import oracle as O
from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN, collapse
from test_oracle_pins import FORTRAN, check_fortran, fortran_grid

pytestmark = pytest.mark.gpu
FORMS = [SHELL_PLAIN, SHELL_PERSISTENT]

FIELDS = ("RTHL", "Srb", "Sbar", "Sbar2", "strain", "gradpk", "gradpkrf", "Fbarx", "e_v", "p_v",
          "zvir_half", "gradpkf", "d2F")


def oracle_batch(grid, i, j, k, n, alatt, ir2min, ZZon, Rf, fcrit, table, shells, nbuff=0, rmax2rs=0.0):
    res, st = O.analyse_peaks(grid, ipp1(i, j, k, n), np.broadcast_to(ir2min, len(i)),
                              np.broadcast_to(ZZon, len(i)), np.broadcast_to(Rf, len(i)),
                              np.broadcast_to(fcrit, len(i)), alatt, table, shells, nbuff, rmax2rs)
    return res, st


def upload_table(ctx, ct):
    ctx.upload_collapse_table(ct.table, *ct.bounds)
    return O.Table(ct.table, *ct.bounds)


@pytest.mark.parametrize("form", FORMS)
def test_fortran_golden(ctx, homeltab_path, form):
    d, ex, ey, ez = fortran_grid()
    ct = collapse.read_homeltab(homeltab_path)
    ctx.upload_collapse_table(ct.table, *ct.bounds)
    ctx.build_shell_tables(8)
    r = ctx.analyse_peaks(d, (ex, ey, ez), [70], [70], [70], 4.0, 4, 1.0, FORTRAN["fcrit"], 1.0, kernel=form)
    assert r["status"][0] == 0
    one = {f: (r[f][0].ravel() if np.ndim(r[f][0]) else r[f][0]) for f in FIELDS}
    check_fortran(one)
    assert one["RTHL"] == pytest.approx(math.sqrt(3.0), rel=1e-6)


def bump_field(n, seed, nbump=20, noise=0.05):
    """The reference's realistic GPU test field (test_shell_gpu.jl:701-735)."""
    rng = np.random.default_rng(seed)
    d = (noise * rng.normal(size=(n, n, n))).astype(np.float32)
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for _ in range(nbump):
        c = rng.integers(9, n - 10, size=3)
        amp = 1.0 + 2.0 * rng.random()
        s2 = 15.0 + 10.0 * rng.random()
        d += (amp * np.exp(-((ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2) / s2)).astype(np.float32)
    f = lambda s: (s * rng.normal(size=(n, n, n))).astype(np.float32)
    return d, (f(0.3), f(0.3), f(0.3)), (f(0.15), f(0.15), f(0.15)), f(0.05)


def compare(r, ref, st, rtol=2e-4, atol=2e-5):
    assert np.array_equal(r["status"], st)
    ok = st == 0
    if not ok.any():
        assert (np.asarray(r["RTHL"]) == -1.0).all()
        return
    for f in FIELDS:
        g = np.asarray(r[f])[ok].reshape(ok.sum(), -1)
        o = np.asarray(ref[f])[ok].reshape(ok.sum(), -1)
        # gradpkrf is NaN where the reference's Phase-3 kernel divides by rad[1] = 0
        scale = np.nan_to_num(np.nanmax(np.abs(o), axis=0, keepdims=True)) + 1e-30
        assert np.array_equal(np.isnan(g), np.isnan(o)), f"{f}: NaN pattern differs"
        bad = ~np.isclose(g, o, rtol=rtol, atol=atol * scale) & ~(np.isnan(g) & np.isnan(o))
        assert not bad.any(), (f, np.argwhere(bad)[:5], g[bad][:5], o[bad][:5])


@pytest.mark.parametrize("form", FORMS)
def test_bumps_vs_oracle(ctx, form):
    n, rmax = 64, 10
    d, eta, eta2, lapd = bump_field(n, 31415)
    ct = collapse.sinusoidal_table()
    tab = upload_table(ctx, ct)
    ctx.build_shell_tables(rmax)
    m0 = np.zeros(d.shape, np.int8)
    i, j, k, _ = O.find_peaks(d, m0, 1, 0.8)
    i, j, k = i - 1, j - 1, k - 1
    assert len(i) > 5
    mask_g = np.zeros(d.shape, np.int8)
    r = ctx.analyse_peaks(d, eta, i, j, k, 4.0, 4, 0.0, 1.0, 1.0, eta2=eta2, lapd=lapd, nbuff=0, mask=mask_g,
                          kernel=form)
    mask_c = np.zeros(d.shape, np.int8)
    grid = O.Grid(d, *eta, eta2=eta2, mask=mask_c, lapd=lapd)
    ref, st = oracle_batch(grid, i, j, k, n, 1.0, 4, 0.0, 4.0, 1.0, tab, O.Shells(rmax))
    assert (st == 0).sum() > 3
    compare(r, ref, st)
    assert np.array_equal(mask_g, mask_c)


@pytest.mark.parametrize("form", FORMS)
def test_shell_profile_counts_bitexact(ctx, form):
    n, rmax = 40, 9
    d, eta, eta2, lapd = bump_field(n, 7, nbump=6)
    ctx.upload_collapse_table(*((lambda c: (c.table, *c.bounds))(collapse.constant_table(0.5))))
    ctx.build_shell_tables(rmax)
    # peaks both deep inside and touching the boundary (out-of-bounds cells)
    pts = np.array([[20, 20, 20], [2, 20, 20], [0, 0, 0], [39, 5, 30], [8, 33, 12]])
    sh = O.Shells(rmax)
    tab = O.Table(np.full((5, 5, 5), 0.5, np.float32), -2, 2, 0, 1, 0, 2)
    m, ns, fb = ctx.shell_profiles(d, eta, pts[:, 0], pts[:, 1], pts[:, 2], 3.0, 1.0, eta2=eta2, lapd=lapd,
                                   max_shells=sh.r2.size, kernel=form)
    for q, (a, b, c) in enumerate(pts):
        _, _, prof = O.analyse_peak(O.Grid(d, *eta, eta2=eta2, lapd=lapd), ipp1(a, b, c, n), 1.0, 4, 1.0, 3.0,
                                    tab, sh, profile_cap=sh.r2.size)
        assert m[q] == prof["m"]
        assert np.array_equal(ns[q, :m[q]], prof["nshell"][:m[q]])
        assert np.allclose(fb[q, :m[q]], prof["Fbar"][:m[q]], rtol=1e-6, atol=1e-6)


def mixed_case(ctx):
    """A shuffled batch mixing small, wave-gathered and boundary peaks on a
    48^3 bump field, with its oracle results and mask."""
    n, rmax = 48, 12
    d, eta, eta2, lapd = bump_field(n, 99, nbump=25, noise=0.3)
    ct = collapse.sinusoidal_table()
    tab = upload_table(ctx, ct)
    ctx.build_shell_tables(rmax)
    m0 = np.zeros(d.shape, np.int8)
    i, j, k, _ = O.find_peaks(d, m0, 1, 0.6)
    i, j, k = i - 1, j - 1, k - 1
    rng = np.random.default_rng(5)
    npk = len(i)
    Rf = rng.choice([0.6, 1.0, 1.7, 2.9, 4.4, 6.0], npk)
    ir2 = np.minimum(np.floor((1.75 * Rf) ** 2), np.floor(39.0 ** 2)).astype(np.int64)
    ZZ = 0.3 + 0.5 * rng.random(npk)
    fc = 0.6 + 0.3 * rng.random(npk)
    perm = rng.permutation(npk)
    i, j, k, Rf, ir2, ZZ, fc = (a[perm] for a in (i, j, k, Rf, ir2, ZZ, fc))
    mask_c = np.zeros(d.shape, np.int8)
    ref, st = O.analyse_peaks(O.Grid(d, *eta, eta2=eta2, mask=mask_c, lapd=lapd), ipp1(i, j, k, n), ir2, ZZ, Rf, fc,
                              1.0, tab, O.Shells(rmax), 3, 2.0)

    def run(form, device_out=False):
        mask = np.zeros(d.shape, np.int8)
        r = ctx.analyse_peaks(d, eta, i, j, k, Rf, ir2, ZZ, fc, 1.0, eta2=eta2, lapd=lapd, rmax2rs=2.0, nbuff=3,
                              mask=mask, kernel=form, device_out=device_out)
        if device_out:
            r = {f: (v.cpu().numpy() if hasattr(v, "cpu") else v) for f, v in r.items()}
        return r, mask
    run.autotune = lambda: ctx.autotune_shell(d, eta, i, j, k, Rf, ir2, ZZ, fc, 1.0, eta2=eta2, lapd=lapd,
                                              rmax2rs=2.0)
    return run, ref, st, mask_c


def test_forms_agree_on_unsorted_mixed_batch(ctx):
    """Plain and persistent forms on a shuffled batch mixing small, gathered
    and boundary peaks (exercises the radix sort, the small-peak waves, the
    chunked gather and both posts), against each other and the oracle."""
    run, ref, st, mask_c = mixed_case(ctx)
    assert (st == 0).sum() > 100 and (st == 6).sum() > 0
    out = {}
    for form in FORMS:
        out[form], mask = run(form)
        compare(out[form], ref, st)
        assert np.array_equal(mask, mask_c)
        assert np.array_equal(out[form]["nshells"], out[FORMS[0]]["nshells"])


@pytest.mark.parametrize("form", FORMS)
def test_device_results_equal_host_results(ctx, form):
    """Device result arrays are written by the kernels directly; host arrays
    are filled by a copy at the end.  Both hold the same bits."""
    run, ref, st, mask_c = mixed_case(ctx)
    host, mh = run(form)
    dev, md = run(form, device_out=True)
    assert host["dump_counts"] == dev["dump_counts"]
    for f in host:
        if f == "dump_counts":
            continue
        a, b = np.ascontiguousarray(host[f]), np.ascontiguousarray(dev[f])
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), f
    assert np.array_equal(mh, md)
    compare(dev, ref, st)


TUNINGS = [
    # everything through the chunked gather
    dict(small_m=0, chunk_shells=8, coop_m=4, gather_waves_cu=4, post_width=1, lane_walk=0),
    # lane posts only, walks beyond 3 shells re-run one wave per peak
    dict(small_m=16, chunk_shells=16, coop_m=1 << 20, gather_waves_cu=32, post_width=1, lane_walk=3),
    # one chunk, wave posts only
    dict(small_m=6, chunk_shells=1024, coop_m=1, gather_waves_cu=1, post_width=1, lane_walk=0),
    # 2-lane group posts
    dict(small_m=4, chunk_shells=64, coop_m=1 << 20, gather_waves_cu=16, post_width=2, lane_walk=0),
    # 4-lane group posts, the cell-based small gather up to 64 shells
    dict(small_m=64, chunk_shells=32, coop_m=1 << 20, gather_waves_cu=8, post_width=4, lane_walk=0),
    # 16-lane group posts
    dict(small_m=8, chunk_shells=128, coop_m=40, gather_waves_cu=16, post_width=16, lane_walk=0),
]


@pytest.mark.parametrize("tune", TUNINGS, ids=["gather-all", "lane-post-deferred", "wave-post", "group2-post",
                                               "group4-post", "group16-post"])
def test_persistent_tunables_keep_results(ctx, tune):
    """The persistent form's launch parameters move peaks between code paths
    (small waves / chunk tables / lane, lane-group or wave post); every
    setting must still match the oracle."""
    run, ref, st, mask_c = mixed_case(ctx)
    before = ctx.tunables()
    try:
        assert ctx.tunables(**tune) == tune
        r, mask = run(SHELL_PERSISTENT)
    finally:
        ctx.tunables(**before)
    assert ctx.tunables() == before
    compare(r, ref, st)
    assert np.array_equal(mask, mask_c)


def test_autotune_keeps_results(ctx):
    """The autotuner picks one of its candidates and results stay put."""
    run, ref, st, mask_c = mixed_case(ctx)
    before = ctx.tunables()
    try:
        chosen = run.autotune()
        assert chosen["chunk_shells"] in (32, 64, 128, 256) and chosen["small_m"] in (0, 8, 12, 16, 24, 32, 48, 64)
        assert chosen["post_width"] in (1, 2, 4, 8, 16) and chosen["lane_walk"] in (0, 8, 16, 32, 64)
        r, mask = run(SHELL_PERSISTENT)
    finally:
        ctx.tunables(**before)
    compare(r, ref, st)
    assert np.array_equal(mask, mask_c)


@pytest.mark.parametrize("form", FORMS)
def test_edge_cases(ctx, form):
    n = 24
    d, eta, eta2, lapd = bump_field(n, 3, nbump=3)
    ct = collapse.sinusoidal_table()
    tab = upload_table(ctx, ct)
    ctx.build_shell_tables(6)
    # empty batch
    r = ctx.analyse_peaks(d, eta, [], [], [], 2.0, 4, 0.0, 1.0, 1.0, kernel=form)
    assert len(r["RTHL"]) == 0
    # near-boundary peaks with the d2F stop at the first out-of-bounds cell,
    # no eta2, tiny filter (npart < 2) and the full table (rmax2rs = 0)
    pts = np.array([[1, 1, 1], [12, 12, 12], [22, 3, 11], [0, 23, 5]])
    for Rf, rmax2rs in ((2.5, 0.0), (2.5, 2.0), (0.3, 2.0)):
        mask_g = np.zeros(d.shape, np.int8)
        r = ctx.analyse_peaks(d, eta, pts[:, 0], pts[:, 1], pts[:, 2], Rf, 4, 0.0, 1.0, 1.0, lapd=lapd,
                              rmax2rs=rmax2rs, nbuff=2, mask=mask_g, kernel=form)
        mask_c = np.zeros(d.shape, np.int8)
        ref, st = oracle_batch(O.Grid(d, *eta, mask=mask_c, lapd=lapd), pts[:, 0], pts[:, 1], pts[:, 2], n,
                               1.0, 4, 0.0, Rf, 1.0, tab, O.Shells(6), nbuff=2, rmax2rs=rmax2rs)
        compare(r, ref, st)
        assert np.array_equal(mask_g, mask_c)
