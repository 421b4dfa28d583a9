"""The solver's secondary paths on the GPU against the FP64 oracle.

- min dimension <= 8: Chebyshev-ramped SOR instead of multigrid
  (pb_solver.rs:1042-1048, solve_lpbe :339-447), on both the fine grid and
  the coarse focusing grid;
- a coarse focusing solve that does not converge within max_iterations:
  the fine solve then takes Debye-Hueckel boundaries (pb_energy.rs:196-235);
- multigrid divergence: Chebyshev SOR continues from the partial potential
  with the boundary condition re-applied (pb_solver.rs:1110-1138).  The
  divergence test is made to fire at cycle 3 from the host
  (MMPBSA_FORCE_DIVERGENCE_CYCLE in the library, orc_set_force_divergence in
  the oracle); every kernel still runs to its normal end.

Bars: PB energies within 0.01 kcal/mol, identical convergence and focusing
flags, SOR sweep and V-cycle counts equal (within 2 sweeps after the forced
divergence, where the SOR starts from an FP32-corrected partial potential).
"""
import os

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

pytestmark = pytest.mark.gpu

TOL = 0.01


def _compare(r, o, sweeps_slack=0):
    print(f"dG {r.total:.6f} vs {o.total:.6f}; E_solv {r.e_solv:.6f} vs {o.solv.energy:.6f}; "
          f"E_ref {r.e_ref:.6f} vs {o.ref.energy:.6f}; it {r.iterations} vs "
          f"({o.solv.iterations}, {o.ref.iterations}); focused {r.focused} vs "
          f"({o.solv.focused}, {o.ref.focused}); coarse it {r.coarse_iterations} vs "
          f"({o.solv.coarse_iterations}, {o.ref.coarse_iterations})")
    assert r.fine_dims == tuple(o.fine_dims) and r.coarse_dims == tuple(o.coarse_dims)
    assert r.focused == (bool(o.solv.focused), bool(o.ref.focused))
    assert r.converged == bool(o.solv.converged and o.ref.converged)
    for a, b in zip(r.iterations, (o.solv.iterations, o.ref.iterations)):
        assert abs(a - b) <= sweeps_slack, (r.iterations, o.solv.iterations, o.ref.iterations)
    for a, b in zip(r.coarse_iterations, (o.solv.coarse_iterations, o.ref.coarse_iterations)):
        assert abs(a - b) <= sweeps_slack
    assert abs(r.e_solv - o.solv.energy) < TOL
    assert abs(r.e_ref - o.ref.energy) < TOL
    assert abs(r.total - o.total) < TOL


def test_sor_on_small_grids(oracle):
    """Fine grid 8 x 9 x 7 and a 5^3 coarse focusing grid: both solves are
    Chebyshev SOR (min dimension <= 8)."""
    top, x, _, _ = synthetic.build_complex(64, 16, seed=101)
    x = (x / 10.0) * 10.0
    kw = dict(grid_spacing=1.5, solute_dielectric=4.0, fine_dims=(8, 9, 7))
    r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    o = oracle.pb_energy(top, x, oracle.pb_params(**kw))
    assert max(o.coarse_dims) <= 8 and min(o.fine_dims) <= 8
    _compare(r, o)


def test_coarse_focusing_failure_falls_back(oracle):
    """max_iterations 18 allows 3 V-cycles: the coarse solve (which needs 4)
    fails, the fine solves run from Debye-Hueckel boundaries and stop
    unconverged after 3 cycles, in the library as in the reference."""
    top, x, _, _ = synthetic.build_complex(300, 40, seed=102)
    x = (x / 10.0) * 10.0
    kw = dict(grid_spacing=1.0, grid_buffer=4.0, solute_dielectric=4.0, max_iterations=18)
    r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    o = oracle.pb_energy(top, x, oracle.pb_params(**kw))
    assert not o.solv.focused and not o.ref.focused and not o.solv.converged
    _compare(r, o)


def test_multigrid_divergence_falls_back_to_sor(oracle):
    top, x, _, _ = synthetic.build_complex(300, 40, seed=102)
    x = (x / 10.0) * 10.0
    kw = dict(grid_spacing=1.0, grid_buffer=4.0, solute_dielectric=4.0)
    plain = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    os.environ["MMPBSA_FORCE_DIVERGENCE_CYCLE"] = "3"
    oracle.lib().orc_set_force_divergence(3)
    try:
        r = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
        o = oracle.pb_energy(top, x, oracle.pb_params(**kw))
    finally:
        del os.environ["MMPBSA_FORCE_DIVERGENCE_CYCLE"]
        oracle.lib().orc_set_force_divergence(0)
    # the fallback really ran: SOR sweep counts, not the ~16 V-cycles
    assert min(o.solv.iterations, o.ref.iterations) > 30
    assert min(r.iterations) > 30 and max(plain.iterations) < 30
    _compare(r, o, sweeps_slack=2)
    # with the hook cleared the library is back on the multigrid path
    again = rst.compute_pb_energy(top, x / 10.0, rst.PbParams(**kw))
    assert again.iterations == plain.iterations and again.total == plain.total
