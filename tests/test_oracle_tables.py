"""Pins for the oracle's table ingest and access (read_pot_table,
init_threepoint, PAIR_INT2/VAL_FUNC2/DERIV_FUNC2).

The 3-point interpolation reproduces any quadratic exactly, and its guard
rows are the last parabola continued, so tables sampled from quadratics
have closed-form answers: a dropped second-difference term, an index off by
one, a wrong fraction or a missing factor 2 in the derivative all fail here.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from imd_amd import imdg as G
from imd_amd import workloads as W


def quad_table(a, b, c, begin, end, npts, radial=True):
    """f(x) = a + b x + c x^2 sampled on [begin, end] (x = r^2)."""
    step, xs = W._grid(begin, end, npts)
    vals = a + b * xs + c * xs * xs
    return W.Table(np.array([begin]), np.array([end]), np.array([step]),
                   [vals], radial)


def test_format2_info_and_guard_rows():
    t = quad_table(0.7, -0.3, 0.05, 1.0, 25.0, 500)
    ot = O.OracleTable.from_table(t)
    info = ot.info()
    assert info["len"][0] == 500 and info["maxsteps"] == 500
    assert info["begin"][0] == 1.0 and info["end"][0] == 25.0
    assert info["invstep"][0] == 1.0 / info["step"][0]
    f = lambda x: 0.7 - 0.3 * x + 0.05 * x * x
    step = info["step"][0]
    shift = f(25.0)
    # radial shift: last tabulated value becomes zero
    assert abs(info["table"][499, 0]) < 1e-15
    # guard rows continue the parabola (exact for a quadratic)
    for k in (500, 501):
        x = 1.0 + k * step
        assert abs(info["table"][k, 0] - (f(x) - shift)) < 1e-12
    assert ot.cellsz.value == 25.0


@pytest.mark.parametrize("coef", [(0.0, 1.0, 0.0), (2.0, -0.5, 0.02),
                                  (-1.0, 0.3, -0.04)])
def test_pair_int2_exact_on_quadratics(coef):
    a, b, c = coef
    t = quad_table(a, b, c, 1.0, 25.0, 1000)
    ot = O.OracleTable.from_table(t)
    f = lambda x: a + b * x + c * x * x
    df = lambda x: b + 2 * c * x
    rng = np.random.default_rng(0)
    for x in np.concatenate([rng.uniform(1.0, 25.0, 200), [1.0, 25.0]]):
        pot, grad, short = ot.pair_int2(0, float(x))
        assert short == 0
        assert abs(pot - (f(x) - f(25.0))) < 1e-11 * max(1.0, abs(f(x)))
        # twice the derivative with respect to the argument
        assert abs(grad - 2 * df(x)) < 1e-9 * max(1.0, abs(df(x)))
        assert abs(ot.val_func2(0, float(x)) - pot) == 0.0
        assert abs(ot.deriv_func2(0, float(x)) - grad) == 0.0


def test_clamping_and_is_short():
    t = quad_table(1.0, 0.5, 0.1, 2.0, 9.0, 200)
    ot = O.OracleTable.from_table(t)
    p_lo, g_lo, s_lo = ot.pair_int2(0, 0.5)
    p_b, g_b, s_b = ot.pair_int2(0, 2.0)
    assert s_lo == 1 and s_b == 0
    assert p_lo == p_b and g_lo == g_b
    p_hi, g_hi, s_hi = ot.pair_int2(0, 50.0)
    p_e, g_e, _ = ot.pair_int2(0, 9.0)
    assert s_hi == 0 and p_hi == p_e and g_hi == g_e
    assert abs(p_e) < 1e-15  # shifted to zero at the end


def test_nonradial_table_not_shifted():
    t = quad_table(-2.0, 0.1, 0.3, 0.0, 3.0, 300, radial=False)
    ot = O.OracleTable.from_table(t, radial=False)
    assert ot.cellsz.value == 0.0
    pot, grad, _ = ot.pair_int2(0, 1.5)
    f = lambda x: -2.0 + 0.1 * x + 0.3 * x * x
    assert abs(pot - f(1.5)) < 1e-12
    assert abs(grad - 2 * (0.1 + 0.6 * 1.5)) < 1e-10


def _write(text):
    fd, path = tempfile.mkstemp(suffix=".pt")
    os.write(fd, text.encode())
    os.close(fd)
    return path


def test_format1_reader_end_len_and_shift():
    # r2 rows with two columns; column 1 ends with zeros
    rows = []
    for k in range(10):
        r2 = 1.0 + 0.5 * k
        v0 = 10.0 - r2
        v1 = (4.0 - r2) if k < 6 else 0.0
        rows.append("%.17g %.17g %.17g" % (r2, v0, v1))
    path = _write("#F 1 2\n#E\n" + "\n".join(rows) + "\n")
    try:
        ot = O.OracleTable(path, 2, radial=True)
        info = ot.info()
        assert info["len"].tolist() == [10, 6]
        assert info["end"].tolist() == [5.5, 3.5]
        assert info["begin"].tolist() == [1.0, 1.0]
        assert abs(info["step"][0] - 0.5) < 1e-15
        # column 0 shifted by its last row (10 - 5.5), column 1 not shifted
        assert abs(info["table"][9, 0]) < 1e-15
        assert abs(info["table"][0, 0] - (9.0 - 4.5)) < 1e-15
        assert info["table"][0, 1] == 3.0
        assert ot.cellsz.value == 5.5
        # maxsteps grows in steps of 50 (PSTEP)
        assert info["maxsteps"] == 50
    finally:
        os.unlink(path)


def test_headerless_file_uses_default_format():
    path = _write("1.0 3.0 0.5\n" + "\n".join("%.1f" % (5 - 0.5 * k) for k in range(5)) + "\n")
    try:
        ot = O.OracleTable(path, 1, radial=True, default_format=2)
        assert ot.info()["len"][0] == 5
    finally:
        os.unlink(path)


@pytest.mark.parametrize("text", ["#F 2 3\n#E\n1 2 0.5\n",
                                  "#F 7 1\n#E\n1 2 0.5\n",
                                  "#C no format\n#E\n1 2 0.5\n",
                                  "#F 2 1\n1 2 0.5\n1\n2\n3\n"])
def test_corrupt_headers_rejected(text):
    path = _write(text)
    try:
        with pytest.raises(ValueError):
            O.OracleTable(path, 1, radial=True)
    finally:
        os.unlink(path)


def test_library_table_relayout_matches_oracle_reader():
    """TableArrays.from_format2 (what the tests hand the C ABI) is the same
    table, bit for bit, as the oracle's file reader produces."""
    eam = W.johnson_eam((W.CU, W.NI), rc=5.5, npts=700)
    for tab, radial in ((eam.pair, True), (eam.rho, True), (eam.embed, False)):
        ta = G.TableArrays.from_format2(tab)
        info = O.OracleTable.from_table(tab, radial=radial).info()
        assert np.array_equal(ta.len, info["len"])
        assert np.array_equal(ta.begin, info["begin"])
        assert np.array_equal(ta.invstep, info["invstep"])
        assert ta.maxsteps == info["maxsteps"]
        tab2 = ta.table.reshape(-1, ta.ncols)
        assert np.array_equal(tab2, info["table"])


# ---------------------------------------------------------------------------
# 4-point and spline interpolation (IMD's FOURPOINT / SPLINE build options)
# ---------------------------------------------------------------------------

def cubic_table(c, begin, end, npts, radial=False):
    step, xs = W._grid(begin, end, npts)
    vals = c[0] + xs * (c[1] + xs * (c[2] + xs * c[3]))
    return W.Table(np.array([begin]), np.array([end]), np.array([step]),
                   [vals], radial)


@pytest.mark.parametrize("c", [(0.5, -0.2, 0.03, -0.001), (1.0, 0.0, 0.0, 0.002)])
def test_fourpoint_exact_on_cubics(c):
    """PAIR_INT3 is the Lagrange polynomial through rows k-1..k+2 with
    k = max(trunc(t), 1), and init_fourpoint continues the last cubic: both
    exact on cubics, including the first segment (read through k = 1) and
    the end (guard rows)."""
    t = cubic_table(c, 1.0, 25.0, 400)
    ot = O.OracleTable.from_table(t, interp=O.INTERP_4POINT)
    info = ot.info()
    f = lambda x: c[0] + x * (c[1] + x * (c[2] + x * c[3]))
    df = lambda x: c[1] + x * (2 * c[2] + 3 * x * c[3])
    step = info["step"][0]
    for k in (400, 401):
        x = 1.0 + k * step
        assert abs(info["table"][k, 0] - f(x)) < 1e-10 * max(1.0, abs(f(x)))
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(1.0, 25.0, 200),
                         [1.0, 1.0 + 0.3 * step, 25.0 - 0.2 * step, 25.0]])
    for x in xs:
        pot, grad, short = ot.pair_int2(0, float(x))
        assert abs(pot - f(x)) < 1e-10 * max(1.0, abs(f(x))), x
        assert abs(grad - 2 * df(x)) < 1e-8 * max(1.0, abs(df(x))), x
        assert ot.val_func2(0, float(x)) == pot
        assert ot.deriv_func2(0, float(x)) == grad


@pytest.mark.parametrize("radial", [False, True])
def test_spline_matches_scipy(radial):
    """init_spline + PAIR_INT_SP against scipy's cubic spline through the
    same rows: natural at the start; at the end the first derivative is
    zero for radial tables and the spline is natural otherwise."""
    from scipy.interpolate import CubicSpline
    begin, end, n = 1.0, 16.0, 300
    step, xs = W._grid(begin, end, n)
    g = lambda x: np.exp(-0.3 * x) * np.cos(0.7 * x) + 0.1 * x
    t = W.Table(np.array([begin]), np.array([end]), np.array([step]),
                [g(xs)], radial)
    ot = O.OracleTable.from_table(t, interp=O.INTERP_SPLINE)
    info = ot.info()
    x = begin + step * np.arange(n)
    y = info["table"][:n, 0]               # after the radial shift
    bc = ((2, 0.0), (1, 0.0)) if radial else "natural"
    cs = CubicSpline(x, y, bc_type=bc)
    assert np.allclose(info["table2"][:n, 0], cs(x, 2), rtol=1e-8,
                       atol=1e-10 * np.abs(cs(x, 2)).max())
    rng = np.random.default_rng(3)
    for r2 in np.concatenate([rng.uniform(begin, end, 300), [begin, end]]):
        pot, grad, short = ot.pair_int2(0, float(r2))
        want_v = float(cs(r2))
        want_g = 2 * float(cs(r2, 1))
        assert abs(pot - want_v) < 1e-11 * max(1.0, abs(want_v)), r2
        assert abs(grad - want_g) < 1e-9 * max(1.0, abs(want_g)), r2


def test_spline_and_fourpoint_differ_from_threepoint():
    """The three interpolations are distinct on a function that is not a
    polynomial of low degree (a mode that silently fell back would pass the
    closed forms above only for its own polynomial degree)."""
    begin, end, n = 1.0, 16.0, 60
    step, xs = W._grid(begin, end, n)
    t = W.Table(np.array([begin]), np.array([end]), np.array([step]),
                [np.sin(xs)], False)
    vals = []
    for m in (O.INTERP_3POINT, O.INTERP_4POINT, O.INTERP_SPLINE):
        ot = O.OracleTable.from_table(t, interp=m)
        vals.append(ot.pair_int2(0, 5.37)[0])
    assert len({round(v, 12) for v in vals}) == 3
    assert all(abs(v - np.sin(5.37)) < 5e-3 for v in vals)
