"""Pins the oracle's restatement of the property-table path
(src/input_thermo.f90:121-693) with the reference's own supercritical-water
table (fixture tests/golden/nist_water_23.5mpa.npy) and closed forms."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
T0 = 645.15  # ref_T0 of example/channel_periodic_scw/input_chapsim.ini


def _table():
    return np.load(os.path.join(HERE, "golden", "nist_water_23.5mpa.npy"))


def test_reference_state_is_a_table_row(po):
    raw = _table()
    i = 400
    p = po.Props(raw, raw[i, 2])
    ref = p.s.ref[:]
    assert ref[0] == raw[i, 2]
    # d, m, k, h, b, cp of the row itself (w1 = 0 or 1 at a tabulated T)
    for got, col in zip(ref[1:], (3, 4, 5, 1, 7, 6)):
        assert abs(got - raw[i, col]) <= 1e-15 * abs(raw[i, col])
    # scaled table: that row becomes the unit state with h = 0
    assert abs(p.column("t")[i] - 1.0) < 1e-15 and abs(p.column("h")[i]) < 1e-15


def test_rows_reproduced_at_their_own_rho_h(po):
    raw = _table()
    p = po.Props(raw, T0)
    rh = p.column("rhoh")
    # rho*h rises to a maximum and falls slightly after it (the reference
    # warns about this, ftplist_check_monotonicity_DH_of_HorT): only values
    # below the falling tail occur once, and only those map back to their row
    top = int(np.argmax(rh))
    idx = np.arange(5, top - 5, 37)
    idx = idx[rh[idx] < rh[top:].min()]
    assert len(idx) > 10
    res, flags = p.from_dh(rh[idx])
    assert not flags.any()
    for j, name in enumerate(("h", "d", "m", "k", "t", "b", "cp")):
        ref = p.column(name)[idx]
        assert np.abs(res[:, j] - ref).max() <= 1e-12 * np.abs(ref).max(), name


def test_linear_table_interpolates_exactly(po):
    """A table whose columns are affine in T is reproduced exactly between
    rows (h linear in T, d constant, so rho*h is linear too)."""
    T = np.linspace(300.0, 700.0, 41)
    raw = np.zeros((41, 8))
    raw[:, 0] = 1.0
    raw[:, 1] = 4000.0 * T - 1.0e5
    raw[:, 2] = T
    raw[:, 3] = 800.0
    raw[:, 4] = 1e-4 + 1e-7 * T
    raw[:, 5] = 0.5 + 1e-4 * T
    raw[:, 6] = 4000.0
    raw[:, 7] = 1e-3 + 1e-6 * T
    p = po.Props(raw, 500.0)
    Tq = np.array([312.5, 401.0, 555.55, 699.0])
    hq = (4000.0 * Tq - 1.0e5 - p.s.ref[4]) / 500.0 / 4000.0
    res, flags = p.from_dh(1.0 * hq)
    assert not flags.any()
    assert np.abs(res[:, 0] - hq).max() < 1e-13
    assert np.abs(res[:, 4] - Tq / 500.0).max() < 1e-13
    assert np.abs(res[:, 2] - (1e-4 + 1e-7 * Tq) / (1e-4 + 1e-7 * 500.0)).max() < 1e-13


def test_out_of_range_is_clamped_and_flagged(po):
    p = po.Props(_table(), T0)
    res, flags = p.from_dh([p.dhmin - 1.0, p.dhmax + 1.0, 0.5 * (p.dhmin + p.dhmax)])
    assert list(flags) == [True, True, False]
    r2, _ = p.from_dh([p.dhmin, p.dhmax])
    assert np.array_equal(res[:2], r2)
