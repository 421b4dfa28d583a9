"""Binding statistics and interaction entropy (host functions of the library,
no device needed), pinned by closed forms and numpy."""
import math

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools._abi import FrameTermsC


def _frames(dmm, dpol=None, dsa=None):
    out = []
    for k, v in enumerate(dmm):
        t = FrameTermsC()
        t.delta_mm = v
        t.delta_polar = 0.0 if dpol is None else dpol[k]
        t.delta_sa = 0.0 if dsa is None else dsa[k]
        t.delta_total = t.delta_mm + t.delta_polar + t.delta_sa
        out.append(rst.FrameEnergy(t))
    return out


def test_binding_stats_match_numpy():
    rng = np.random.default_rng(3)
    dmm, dpol, dsa = rng.normal(-40, 5, 37), rng.normal(20, 3, 37), rng.normal(-3, 0.2, 37)
    r = rst.BindingResult(_frames(dmm, dpol, dsa))
    tot = dmm + dpol + dsa
    assert r.mean_delta_mm == pytest.approx(np.mean(dmm), rel=1e-14)
    assert r.mean_delta_total == pytest.approx(np.mean(tot), rel=1e-14)
    assert r.std_delta_polar == pytest.approx(np.std(dpol, ddof=1), rel=1e-12)
    assert r.std_delta_total == pytest.approx(np.std(tot, ddof=1), rel=1e-12)
    assert r.sem_delta_total == pytest.approx(np.std(tot, ddof=1) / math.sqrt(37), rel=1e-12)
    assert r.mean_delta_gb == r.mean_delta_polar
    one = rst.BindingResult(_frames([1.5]))
    assert one.std_delta_total == 0.0 and one.sem_delta_total == 0.0
    with pytest.raises(ValueError, match="No frames"):
        rst.BindingResult([])


def test_interaction_entropy_closed_forms():
    T = 298.15
    kt = 0.00198688 * T
    assert rst.interaction_entropy(_frames([-10.0]), T) is None
    # constant interaction energy: no fluctuation, no entropy
    assert rst.interaction_entropy(_frames([-7.0] * 5), T).minus_tds == pytest.approx(0.0, abs=1e-15)
    # two frames at mean +- a: kT ln cosh(a / kT)
    for a in (0.1, 1.0, 3.0):
        e = rst.interaction_entropy(_frames([-20.0 + a, -20.0 - a]), T)
        assert e.method == "interaction_entropy"
        assert e.minus_tds == pytest.approx(kt * math.log(math.cosh(a / kt)), rel=1e-12)
    # far beyond exp's range: log-sum-exp keeps it finite, -> a - kT ln 2
    a = 2000.0
    e = rst.interaction_entropy(_frames([a, -a]), T)
    assert e.minus_tds == pytest.approx(a - kt * math.log(2.0), rel=1e-12)
    # Gaussian fluctuations: -T dS -> sigma^2 / (2 kT) for many samples
    rng = np.random.default_rng(11)
    sig = 0.3
    e = rst.interaction_entropy(_frames(rng.normal(0.0, sig, 200000)), T)
    assert e.minus_tds == pytest.approx(sig * sig / (2 * kt), rel=0.02)
