"""Quasi-harmonic entropy on the GPU (entropy.rs:97-217; the PyO3
quasi_harmonic_entropy of lib.rs:2226-2237) against the FP64 oracle, which
restates the reference's cyclic Jacobi eigensolver.  The covariance is the
reference's to the bit; the eigenvalues come from rocSOLVER's
divide-and-conquer solver, so -T dS agrees to rounding: 1e-9 relative.
Sizes cover fewer frames than coordinates (a rank-deficient covariance, most
modes zero), more frames than coordinates, and several covariance tiles.  A
1,200-coordinate case is checked against LAPACK's eigenvalues of the
covariance numpy builds (too large for the oracle's Jacobi in test time)."""
import math

import numpy as np
import pytest

import rust_simulation_tools as rst

pytestmark = pytest.mark.gpu

T = 298.15
MASSES = np.array([1.008, 12.011, 14.007, 15.999, 32.06])


def _system(na, nf, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    m = rng.choice(MASSES, na)
    base = rng.normal(scale=6.0, size=(na, 3))
    # correlated motions: a few collective modes plus independent noise
    modes = rng.normal(size=(4, na, 3))
    amp = rng.normal(size=(nf, 4))
    traj = base + scale * np.einsum("fk,kac->fac", amp, modes) + 0.05 * rng.normal(size=(nf, na, 3))
    return traj, m


@pytest.mark.parametrize("na,nf", [(5, 50), (40, 30), (70, 300), (130, 64)])
def test_matches_oracle(oracle, na, nf):
    traj, m = _system(na, nf, seed=na * 1000 + nf)
    got = rst.quasi_harmonic_entropy(traj, m, T)
    want = oracle.quasi_harmonic_entropy(traj, m, T)
    print(f"na {na} nf {nf}: gpu {got.minus_tds!r} oracle {want!r}")
    assert got.method == "quasi_harmonic"
    assert got.minus_tds == pytest.approx(want, rel=1e-9)


def test_single_mode_and_none(oracle):
    # the reference's test system: one 12 amu atom at +-1 A along x
    traj = np.array([[[1.0, 0.0, 0.0]], [[-1.0, 0.0, 0.0]]])
    got = rst.quasi_harmonic_entropy(traj, [12.0], T).minus_tds
    x = 1.054571817e-34 / math.sqrt(1.380649e-23 * T * 12.0 * 1e-20 * 1.66054e-27)
    want = -T * 0.00198688 * (x / math.expm1(x) - math.log(-math.expm1(-x)))
    assert got == pytest.approx(want, rel=1e-12)
    assert got == pytest.approx(oracle.quasi_harmonic_entropy(traj, [12.0], T), rel=1e-12)
    assert rst.quasi_harmonic_entropy(traj[:1], [12.0], T) is None
    assert rst.quasi_harmonic_entropy(np.zeros((3, 0, 3)), [], T) is None
    # a negative mass gives NaN in the reference (square root of it), here too
    bad = rst.quasi_harmonic_entropy(np.stack([traj[0], traj[1], traj[0]]), [-1.0], T)
    assert math.isnan(bad.minus_tds)
    assert math.isnan(oracle.quasi_harmonic_entropy(np.stack([traj[0], traj[1], traj[0]]),
                                                    [-1.0], T))


def test_large_against_lapack():
    na, nf = 400, 120
    traj, m = _system(na, nf, seed=7)
    got = rst.quasi_harmonic_entropy(traj, m, T).minus_tds
    # the reference's covariance in numpy, LAPACK's eigenvalues, the same
    # mode entropy
    w = np.sqrt(m)[None, :, None]
    x = (traj * w).reshape(nf, -1)
    d = x - x.mean(axis=0)
    cov = d.T @ d / nf
    lam = np.sort(np.linalg.eigvalsh(cov))[::-1]
    s = 0.0
    for v in lam:
        if v < 1e-10:
            continue
        a = 1.054571817e-34 / math.sqrt(1.380649e-23 * T * v * 1e-20 * 1.66054e-27)
        if a > 500:
            continue
        s += a / math.expm1(a) - math.log(-math.expm1(-a))
    want = -T * 0.00198688 * s
    print(f"gpu {got!r} lapack {want!r}")
    assert got == pytest.approx(want, rel=1e-8)


def test_reference_python_test_shape(cfg0):
    # tests/test_mmpbsa.py:372-385 of the reference: 50 noisy frames of a
    # subsystem's coordinates, masses 12: a finite estimate of that method
    top, x, _, lig, _ = cfg0
    sel = top.build_selection(lig)
    coords = np.asarray(x)[sel]
    rng = np.random.default_rng(42)
    traj = np.stack([coords + rng.normal(0, 0.1, coords.shape) for _ in range(50)])
    qh = rst.quasi_harmonic_entropy(traj, np.ones(len(sel)) * 12.0, temperature=298.15)
    assert qh is not None and np.isfinite(qh.minus_tds) and qh.method == "quasi_harmonic"
