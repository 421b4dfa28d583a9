"""Alternative multigrid kernels that must agree bit for bit, on uniform and
non-uniform maps, focused and unfocused, single and batched:
- the one-pass-per-phase smoother (three red-black sweeps per launch) against
  the one-pass-per-sweep smoother (the same updates in the same order);
- the z-streaming prolongation and the fused residual + restriction pass
  against the per-node kernels (the same expressions, the same fma chains)."""
import os

import numpy as np
import pytest

import rust_simulation_tools as rst
from rust_simulation_tools import synthetic

pytestmark = pytest.mark.gpu

PB_CFG0 = dict(grid_spacing=0.6, solute_dielectric=4.0, solvent_dielectric=80.0,
               fine_dims=(65, 65, 65))


def _with_env(var, value, fn):
    old = os.environ.get(var)
    if value is None:
        os.environ.pop(var, None)
    else:
        os.environ[var] = value
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop(var, None)
        else:
            os.environ[var] = old


def _with_smoother(mode, fn):
    return _with_env("MMPBSA_SMOOTHER", mode, fn)


def _pb_tuple(r):
    return (r.total, r.e_solv, r.e_ref, r.iterations, r.converged)


@pytest.mark.parametrize("kw", [PB_CFG0,
                                dict(grid_spacing=0.7, grid_buffer=6.0, solute_dielectric=2.0,
                                     solvent_dielectric=80.0, salt_concentration=0.0),
                                dict(grid_spacing=0.5, solute_dielectric=4.0,
                                     solvent_dielectric=80.0, fine_dims=(97, 81, 73),
                                     fillratio=0.0)])
def test_pb_identical(cfg0, kw):
    top, x, _, _, _ = cfg0
    p = rst.PbParams(**kw)
    a = _with_smoother("sweep", lambda: rst.compute_pb_energy(top, x / 10.0, p))
    b = _with_smoother(None, lambda: rst.compute_pb_energy(top, x / 10.0, p))
    print("sweep", _pb_tuple(a), "phase", _pb_tuple(b))
    assert _pb_tuple(a) == _pb_tuple(b)


def test_batched_frames_identical(cfg0):
    top, _, rec, lig, seed = cfg0
    base = synthetic.config_system(0)[1]
    frames = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(3)]) / 10.0
    pb = rst.PbParams(**PB_CFG0)

    def run():
        r = rst.compute_binding_energy_frames(top, frames, rec, lig, pb_params=pb, batched=True)
        return [(f.complex_polar, f.receptor_polar, f.ligand_polar, f.delta_total)
                for f in r.frames()]

    assert _with_smoother("sweep", run) == _with_smoother(None, run)


PB_CASES = [PB_CFG0,
            dict(grid_spacing=0.7, grid_buffer=6.0, solute_dielectric=2.0, solvent_dielectric=80.0,
                 salt_concentration=0.0),
            dict(grid_spacing=0.5, solute_dielectric=4.0, solvent_dielectric=80.0,
                 fine_dims=(97, 81, 73), fillratio=0.0)]


@pytest.mark.parametrize("kw", PB_CASES)
def test_streaming_kernels_identical(cfg0, kw):
    top, x, _, _, _ = cfg0
    p = rst.PbParams(**kw)
    a = _with_env("MMPBSA_STREAM", "0", lambda: rst.compute_pb_energy(top, x / 10.0, p))
    b = _with_env("MMPBSA_STREAM", None, lambda: rst.compute_pb_energy(top, x / 10.0, p))
    print("per-node", _pb_tuple(a), "streaming", _pb_tuple(b))
    assert _pb_tuple(a) == _pb_tuple(b)


def test_streaming_batched_identical(cfg0):
    top, _, rec, lig, seed = cfg0
    base = synthetic.config_system(0)[1]
    frames = np.stack([synthetic.noisy_frame(base, seed, f) for f in range(3)]) / 10.0
    pb = rst.PbParams(**PB_CFG0)

    def run():
        r = rst.compute_binding_energy_frames(top, frames, rec, lig, pb_params=pb, batched=True)
        return [(f.complex_polar, f.receptor_polar, f.ligand_polar, f.delta_total)
                for f in r.frames()]

    assert _with_env("MMPBSA_STREAM", "0", run) == _with_env("MMPBSA_STREAM", None, run)
