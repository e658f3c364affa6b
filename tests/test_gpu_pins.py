"""The analytic cases of tests/pin_cases.py through both GPU forms: R_TH
interpolation, Fbarx, strain, e_v / p_v, gradpk and the permuted gradpkf,
Sbar2, d2F with its stop rule, zvir_half and the exclusion mask against the
values the mathematics fixes (FP32 shell sums: relative tolerances of a few
float32 ulps of the summed quantities)."""
import math

import numpy as np
import pytest

import pin_cases as PC
from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN

pytestmark = pytest.mark.gpu
FORMS = [SHELL_PLAIN, SHELL_PERSISTENT]


def run(ctx, case, table, bounds, fcrit, ZZon, form):
    ctx.upload_collapse_table(table, *bounds)
    ctx.build_shell_tables(case.rmax)
    mask = np.zeros(case.delta.shape, np.int8)
    i, j, k = ([c] for c in case.centre)
    r = ctx.analyse_peaks(case.delta, case.eta, i, j, k, 4.0, case.ir2min, ZZon, fcrit, 1.0, eta2=case.eta2,
                          lapd=case.lapd, rmax2rs=0.0, nbuff=case.nbuff, mask=mask, kernel=form)
    one = {f: np.asarray(v[0]) for f, v in r.items() if f != "dump_counts"}
    return one, mask


def check_common(one, case, R, Fbarx, inside_r2):
    ex = PC.expect_common(case, R, Fbarx, inside_r2)
    assert np.allclose(np.asarray(one["strain"]).reshape(3, 3), ex["strain"], rtol=1e-5, atol=1e-7)
    assert one["e_v"] == pytest.approx(ex["e_v"], rel=1e-5)
    assert one["p_v"] == pytest.approx(ex["p_v"], rel=1e-4)
    assert np.allclose(one["gradpk"], ex["gradpk"], rtol=2e-4)
    assert np.allclose(one["gradpkf"], ex["gradpkf"], rtol=2e-4)
    assert np.allclose(one["gradpkf"], one["gradpk"][[1, 2, 0]], rtol=2e-4)
    assert np.allclose(one["Sbar2"], ex["Sbar2"], rtol=1e-6)
    assert one["d2F"] == pytest.approx(ex["d2F"], rel=1e-6)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("fcrit", [1.0, 3.0])
def test_constant_table(ctx, form, fcrit):
    case = PC.bump_case()
    t, b = PC.constant_table(1.8)
    one, mask = run(ctx, case, t, b, fcrit, 1.5, form)
    assert one["status"] == 0
    rad, fb = PC.fbar_profile(case)
    mc = PC.crossing(fb, fcrit)
    assert one["RTHL"] == pytest.approx(rad[mc - 1], rel=1e-12)
    assert one["Fbarx"] == pytest.approx(fb[mc - 1], rel=1e-5)
    assert one["zvir_half"] == pytest.approx(float(np.float32(1.8)) - 1.0, rel=1e-12)
    check_common(one, case, one["RTHL"], one["Fbarx"], rad[mc - 1] ** 2)
    sel, _ = PC.inside_cells(case, one["RTHL"], frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("a,b,fcrit,ZZon", [(1.0, 1.0, 1.0, 1.5), (0.9, 1.5, 1.0, 2.0)])
def test_linear_table(ctx, form, a, b, fcrit, ZZon):
    case = PC.bump_case()
    t, bounds = PC.linear_table(a, b)
    one, mask = run(ctx, case, t, bounds, fcrit, ZZon, form)
    assert one["status"] == 0
    ex = PC.expect_linear(case, a, b, fcrit, ZZon)
    rad, _ = PC.fbar_profile(case)
    assert one["RTHL"] == pytest.approx(ex["RTHL"], rel=1e-5)
    assert rad[ex["inside"]] < one["RTHL"] < rad[ex["inside"] + 1]
    assert one["Fbarx"] == pytest.approx(ex["Fbarx"], rel=1e-5)
    assert one["zvir_half"] == pytest.approx(ex["zvir_half"], rel=1e-5, abs=1e-5)
    check_common(one, case, one["RTHL"], one["Fbarx"], rad[ex["inside"]] ** 2)
    sel, _ = PC.inside_cells(case, one["RTHL"], frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel)


@pytest.mark.parametrize("form", FORMS)
def test_d2F_stop_at_grid_edge(ctx, form):
    from oracle import oracle as O
    case = PC.bump_case(centre=(6, 23, 23))
    t, b = PC.constant_table(1.8)
    one, mask = run(ctx, case, t, b, 1.0, 1.5, form)
    assert one["status"] == 0 and one["RTHL"] > 7
    sh = O.Shells(case.rmax)
    exp = PC.d2F_walk(case, float(one["RTHL"]), sh.di, sh.dj, sh.dk, sh.r2)
    assert one["d2F"] == pytest.approx(exp, rel=1e-6)
    sel, _ = PC.inside_cells(case, float(one["RTHL"]), frame=case.nbuff)
    assert np.array_equal(mask.astype(bool), sel)
    assert math.isfinite(one["zvir_half"])
