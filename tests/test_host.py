"""Host-side code of the path (CPU only): cosmology bookkeeping, synthetic
inputs, collapse-table I/O and the tile configurations, against the oracle
and closed forms."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from peakpatch_hip import collapse, cosmology, synth
from peakpatch_hip.pipeline import CONFIGS, TileDriver, gather_extent, shell_counts

COSMO = dict(Om=0.315, Ob=0.049, OL=0.685, h=0.674)


def test_growth_tables_match_oracle():
    """fsc_of_z = 1.686 D(z) (RadialShell.jl:191-196, Cosmology.jl:148-293)."""
    g = cosmology.GrowthTables(cosmology.CosmologyParams())
    d = O.Dlinear(COSMO["Om"], COSMO["Ob"], COSMO["OL"], COSMO["h"])
    z = np.array([0.0, 0.05, 0.3, 0.5, 1.0, 2.0, 4.6])
    ours = g.fsc_of_z(z)
    ref = np.array([d.fsc_of_z(v) for v in z])
    np.testing.assert_allclose(ours, ref, rtol=1e-10)
    # D is monotone in a and 1 today
    assert ours[0] == pytest.approx(1.686, abs=4e-4)
    assert (np.diff(ours) < 0).all()


def test_chi_to_z_matches_oracle():
    """peak_redshift via the chi(z) table + bisection (Cosmology.jl:21-24, 333-372)."""
    c = cosmology.CosmologyParams()
    ours = cosmology.ChiToZ(c, z_max=5.6)
    ref = O.ChiToZ(COSMO["Om"], COSMO["OL"], COSMO["h"], z_max=5.6)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-300, 900, size=(200, 3))
    obs = (-300.0, -300.0, -300.0)
    zo = ours.peak_redshift(obs, pts[:, 0], pts[:, 1], pts[:, 2])
    zr = np.array([ref.peak_redshift(obs, p) for p in pts])
    np.testing.assert_allclose(zo, zr, atol=1e-6)
    # the h factor of chi (quirk Q2): chi(z) = 2998 h int dz/E, EdS closed form
    eds = cosmology.CosmologyParams(Om=1.0, OB=0.0, OL=0.0, h=0.7)
    assert cosmology.chi(1.0, eds) == pytest.approx(2998 * 0.7 * 2 * (1 - 1 / math.sqrt(2)), rel=1e-8)


def test_ir2min():
    """ir2min = min(floor((1.75 Rf/alatt)^2), floor((40/alatt - 1)^2)) (Pipeline.jl:222-230)."""
    Rf = np.array([1.0, 4.0, 15.0, 30.0])
    for alatt in (600 / 512, 0.326, 3.0):
        want = [min(math.floor((1.75 * r / alatt) ** 2), math.floor((40.0 / alatt - 1) ** 2)) for r in Rf]
        assert list(cosmology.ir2min(Rf, alatt)) == want


def test_threefry_known_answer_and_moments():
    """Threefry-2x64-20: Random123 known-answer vector (zero counter, zero
    key); the Box-Muller normals have unit variance."""
    from oracle import fields as OF
    x0, x1 = OF.threefry2x64_20(np.uint64(0), np.uint64(0), np.uint64(0), np.uint64(0))
    assert (int(x0), int(x1)) == (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)
    g1, g2 = OF.threefry_gaussian(42, np.arange(200000, dtype=np.uint64))
    g = np.concatenate([g1, g2])
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01
    # counter mode: the same seed and index give the same number
    a = OF.fill_noise_threefry(8, 7)
    b = OF.fill_noise_threefry(8, 7)
    assert np.array_equal(a, b) and not np.array_equal(a, OF.fill_noise_threefry(8, 8))


def test_pk_table_interpolation():
    """load_pk's log-log interpolation (PowerSpectrum.jl:11-23) of the EH98
    table reproduces the closed form between rows and extrapolates linearly
    in log-log beyond them."""
    from oracle import fields as OF
    k, P = synth.pk_table()
    f = OF.PkTable(k, P)
    kk = np.exp(np.linspace(np.log(2e-4), np.log(50.0), 777))
    assert np.allclose(f(kk), synth.EH98NoWiggle()(kk), rtol=2e-5)
    lo = f(np.array([k[0] / 10]))[0]
    slope = (np.log(P[1]) - np.log(P[0])) / (np.log(k[1]) - np.log(k[0]))
    assert lo == pytest.approx(P[0] * 10.0 ** -slope, rel=1e-9)


def test_power_spectrum_normalisation():
    """EH98 no-wiggle P(k) normalised to sigma8 (MassFunction.jl:29-42)."""
    pk = synth.EH98NoWiggle()
    assert pk.sigma_R(8.0) == pytest.approx(0.808, rel=1e-4)
    # field-generator convention: P / (2 pi)^3 (RandomField.jl:136-175, quirk Q18)
    k = np.array([0.05, 0.1, 0.5])
    np.testing.assert_allclose(pk(k), pk.p_std(k) / (2 * math.pi) ** 3, rtol=1e-12)


def test_filter_bank():
    f = synth.filter_bank(25, 1.0, 30.0)
    assert len(f) == 25 and f[0] == pytest.approx(30.0) and f[-1] == pytest.approx(1.0)
    assert (np.diff(f) < 0).all()
    np.testing.assert_allclose(np.diff(np.log(f)), np.log(1.0 / 30.0) / 24)


def test_collapse_table_roundtrip(tmp_path, homeltab_path):
    """read_homeltab / write_homeltab layout (CollapseTable.jl:81-122)."""
    ct = collapse.read_homeltab(homeltab_path)
    p = os.path.join(tmp_path, "t.dat")
    collapse.write_homeltab(p, ct)
    assert open(p, "rb").read() == open(homeltab_path, "rb").read()
    assert ct.table.dtype == np.float32 and ct.table.ndim == 3


def test_tile_configs():
    """nbuff = cap + 1 with cap = floor(1.75 Rf_max / alatt) (Pipeline.jl:96-98)."""
    assert CONFIGS["c3"].alatt == pytest.approx(600 / 512)
    assert [CONFIGS[c].cap for c in ("c1", "c2", "c3", "c4")] == [23, 26, 44, 80]
    for c in ("c1", "c2", "c3", "c4"):
        cfg = CONFIGS[c]
        assert cfg.nbuff_eff == cfg.cap + 1 and cfg.nhunt == cfg.cap


def test_shell_counts_match_oracle():
    r2, cnt = shell_counts(9)
    sh = O.Shells(9)
    v, c = np.unique(sh.r2, return_counts=True)
    assert np.array_equal(r2, v) and np.array_equal(cnt, c)


def test_gather_extent_matches_oracle_profile():
    """The finalised shell count m used for the byte count equals the
    oracle's (npart cut, last shell unfinalised)."""
    n, rmax, alatt = 40, 9, 1.0
    rng = np.random.default_rng(1)
    d = rng.normal(size=(n, n, n)).astype(np.float32)
    z = np.zeros_like(d)
    sh = O.Shells(rmax)
    tab = O.Table(np.full((5, 5, 5), 0.5, np.float32), -2, 2, 0, 1, 0, 2)
    Rf = np.array([0.7, 1.0, 1.3, 2.0, 3.0])
    cells, m = gather_extent(rmax, Rf, alatt, 2.0)
    for q, rf in enumerate(Rf):
        _, _, prof = O.analyse_peak(O.Grid(d, z, z, z), 20 + 20 * n + 20 * n * n + 1, alatt, 4, 1.0, rf, tab, sh,
                                    rmax2rs=2.0, profile_cap=sh.r2.size)
        assert m[q] == prof["m"]


def test_tile_driver_bookkeeping():
    """Per-peak bookkeeping of run_multitile_split (MultiResolution.jl:959-988):
    positions, ZZon from the light-cone distance, fcrit = fsc_of_z(z_pk), z_max cut."""
    cfg = CONFIGS["c3"]
    drv = TileDriver(cfg)

    class P:
        pass
    p = P()
    p.i = np.array([0, 255, 511], np.int32)
    p.j = np.array([0, 255, 511], np.int32)
    p.k = np.array([0, 300, 511], np.int32)
    p.filt = np.array([24, 10, 0], np.int32)
    p.delta_s = np.ones(3, np.float32)
    p.lapd_s = np.zeros(3, np.float32)
    b = drv.batch(p)
    cen = 0.5 * (cfg.n + 1)
    np.testing.assert_allclose(b.x, cfg.alatt * (p.i + 1 - cen))
    zpk = drv.chi2z.peak_redshift(cfg.cen, b.x, b.y, b.z)
    np.testing.assert_allclose(b.ZZon, 1.0 + zpk)
    np.testing.assert_allclose(b.fcrit, drv.growth.fsc_of_z(zpk))
    assert list(b.Rf) == list(cfg.filters[p.filt])
