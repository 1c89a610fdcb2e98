"""The oracle's Heaviside projection pinned on the CPU.

Known answers follow the reference's tests/test_projection.cpp: end points,
monotonicity, the derivative against central differences, the identity and
step limits (:105-135), the beta schedule's stages (:137-162), gradients
through the projection against finite differences (:164-212), and a projected
run ending crisper at beta_max than the plain run (:260-328).
"""
import numpy as np
import pytest

from oracle import refcpu as R
from sparlab_amd import model as M


def test_projection_end_points_monotone_and_slope():
    for eta in (0.3, 0.5, 0.7):
        for beta in (0.5, 1.0, 8.0, 64.0):
            assert abs(R.heaviside_project(0.0, beta, eta)) <= 1e-15
            assert R.heaviside_project(1.0, beta, eta) == pytest.approx(1.0, rel=1e-15)
            prev = -1.0
            for i in range(101):
                r = 0.01 * i
                v = R.heaviside_project(r, beta, eta)
                assert v >= prev
                prev = v
                h = 1e-6
                fd = (R.heaviside_project(r + h, beta, eta) - R.heaviside_project(r - h, beta, eta)) / (2 * h)
                assert R.heaviside_derivative(r, beta, eta) == pytest.approx(fd, rel=1e-6, abs=1e-9)
    for i in range(11):
        assert abs(R.heaviside_project(0.1 * i, 1e-3, 0.5) - 0.1 * i) < 1e-6
    assert R.heaviside_project(0.45, 256.0, 0.5) < 1e-5
    assert R.heaviside_project(0.55, 256.0, 0.5) > 1.0 - 1e-5
    assert R.heaviside_project(0.5, 4.0, 0.5) == pytest.approx(0.5)
    # symmetric about eta = 1/2
    for r in np.linspace(0, 1, 9):
        assert R.heaviside_project(r, 6.0, 0.5) + R.heaviside_project(1 - r, 6.0, 0.5) == pytest.approx(1.0, rel=1e-14)


def test_beta_schedule_stages():
    assert R.projection_stages(1.0, 32.0, 2.0, 0) == (6, 1.0)  # 1, 2, 4, 8, 16, 32
    assert R.projection_stages(1.0, 32.0, 2.0, 3)[1] == 8.0
    assert R.projection_stages(1.0, 32.0, 2.0, 9)[1] == 32.0
    assert R.projection_stages(1.0, 20.0, 2.0, 0)[0] == 6
    assert R.projection_stages(1.0, 20.0, 2.0, 5)[1] == 20.0
    assert R.projection_stages(1.0, 16.0, 2.0, 4) == (5, 16.0)


def test_gradients_through_the_projection_match_finite_differences():
    m = M.mbb_beam(24, 8)
    o = R.Oracle(m, tol=1e-13)
    o.set_projection(4.0, 0.5)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.25, 0.75, m.nelem)
    ev = o.evaluate(x)
    h = 1e-6
    for e in (3, 50, 101, 170):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        cp, cm = o.evaluate(xp, need_gradients=False), o.evaluate(xm, need_gradients=False)
        assert ev["dc_dx"][e] == pytest.approx((cp["compliance"] - cm["compliance"]) / (2 * h), rel=2e-5)
        assert ev["dv_dx"][e] == pytest.approx((cp["volume"] - cm["volume"]) / (2 * h), rel=2e-5)
    # without the projection the volume gradient is the filter's pull-back of the volumes
    o.set_projection(0.0)
    plain = o.evaluate(x)
    assert not np.allclose(plain["dv_dx"], ev["dv_dx"])


def test_projected_run_ends_crisper_at_beta_max():
    m = M.mbb_beam(60, 20)
    x0 = m.initial_design()
    base = R.Oracle(m)
    xb, hb, nb, _ = base.run_oc(x0, 160, change_tol=0.01)
    sharp = R.Oracle(m)
    xs, hs, ns, _ = sharp.run_oc(x0, 160, change_tol=0.01,
                                 projection=dict(beta_start=1.0, beta_max=8.0, beta_interval=25))
    assert sharp.last_final_beta == 8.0
    assert np.all(np.diff(hs[:, 12]) >= 0) and hs[-1, 12] == 8.0 and hs[0, 12] == 1.0
    assert np.all(hb[:, 12] == 0.0)
    assert hs[-1, 7] < 0.5 * hb[-1, 7]
    assert hs[-1, 4] == pytest.approx(0.5, rel=1e-3)
    # a stage lasts at most beta_interval iterations
    changes = np.flatnonzero(np.diff(hs[:, 12])) + 1
    assert np.all(np.diff(np.concatenate([[0], changes])) <= 25)


def test_robust_run_orders_the_three_designs():
    """The robust formulation (TopologyOptimizer.cpp:60-80,566-640,347-390): the eroded
    design is the least stiff and lightest, the dilated the stiffest and heaviest; the
    dilated target is rescaled only every robust_volume_interval iterations."""
    m = M.mbb_beam(60, 20)
    o = R.Oracle(m)
    x, h, n, c = o.run_oc(m.initial_design(), 120, change_tol=0.01,
                          projection=dict(beta_start=1.0, beta_max=8.0, beta_interval=25, robust=True,
                                          robust_delta=0.1, robust_volume_interval=5))
    rb = o.last_robust
    assert rb["compliance_eroded"] > c > rb["compliance_dilated"]
    assert rb["volume_fraction_eroded"] < h[-1, 4] < rb["volume_fraction_dilated"]
    assert np.all(h[:, 13] < h[:, 4]) and np.all(h[:, 4] < h[:, 14])
    changes = np.flatnonzero(np.diff(h[:, 15]) != 0) + 1  # 0-based rows where the target moved
    assert np.all(changes % 5 == 0)
    assert rb["dilated_target_fraction"] == h[-1, 15]
    # the dilated design meets its rescaled target
    assert h[-1, 14] == pytest.approx(h[-1, 15], rel=2e-2)
