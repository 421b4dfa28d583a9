"""MI355X (gfx950) MM-PBSA solvation energies behind the reference's Python API.

Names, keyword arguments, defaults, units and result getters follow the PyO3
module of rust-simulation-tools (crates/bindings/src/lib.rs:1450-2280):
coordinates are float64 (N, 3) arrays in nm and are scaled to Angstrom
(lib.rs:83-88); string errors raise ValueError (lib.rs:2136, :2158).
Every energy is computed by libmmpbsa_hip.so on the GPU; importing works
without a GPU, calling a function without one raises RuntimeError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (FrameTermsC, GbParamsC, MmTermsC, PbParamsC, PbResultC, RequestStatsC,
                   ResidueContribC,
                   SaParamsC, load_library)
from .topology import AmberTopology, extract_coords, read_prmtop
from .trajectory import DcdReader, MdcrdReader, read_dcd_header, reimage
from .topology import extract_subtopology as _extract_subtopology

__all__ = [
    "AmberTopology", "GbModel", "GbParams", "PbParams", "SaParams", "MmEnergy", "GbEnergy",
    "SaEnergy", "PbEnergy", "FrameEnergy", "BindingResult", "ResidueContribution",
    "DecompositionResult", "compute_mm_energy", "compute_gb_energy", "compute_sa_energy",
    "compute_pb_energy", "compute_gb_energy_from_radii", "compute_gb_energy_cached", "GbCache",
    "compute_binding_energy_single_frame", "compute_binding_energy_frames",
    "decompose_binding_energy", "decompose_binding_energy_pb", "extract_subtopology",
    "EntropyEstimate", "interaction_entropy", "quasi_harmonic_entropy",
]


class GbModel:
    """lib.rs:1454-1470 (AMBER igb 1 / 2 / 5)."""
    Hct = 0
    ObcI = 1
    ObcII = 2


def _opt(p, **kw):
    for k, v in kw.items():
        if v is not None:
            setattr(p, k, v)
    return p


class GbParams:
    """lib.rs:1472-1540; defaults gb_energy.rs:43-56."""

    def __init__(self, model=None, solute_dielectric=None, solvent_dielectric=None,
                 salt_concentration=None, temperature=None, offset=None, rgbmax=None,
                 cutoff=None):
        self._c = GbParamsC()
        load_library().mmpbsa_gb_params_default(C.byref(self._c))
        _opt(self._c, model=None if model is None else int(model),
             solute_dielectric=solute_dielectric, solvent_dielectric=solvent_dielectric,
             salt_concentration=salt_concentration, temperature=temperature, offset=offset,
             rgbmax=rgbmax, cutoff=cutoff)

    solute_dielectric = property(lambda s: s._c.solute_dielectric)
    solvent_dielectric = property(lambda s: s._c.solvent_dielectric)
    salt_concentration = property(lambda s: s._c.salt_concentration)
    temperature = property(lambda s: s._c.temperature)
    cutoff = property(lambda s: s._c.cutoff)


class PbParams:
    """lib.rs:1542-1626; defaults pb_energy.rs:49-66.  Extensions (no
    counterpart in the reference): ``fine_dims`` fixes the fine grid
    dimensions, centred as auto_grid centres or at ``fine_center``;
    ``coarse_dims`` / ``coarse_spacing`` / ``coarse_center`` give an explicit
    focusing coarse grid.  Explicit grids serve all three subsystems."""

    def __init__(self, grid_spacing=None, grid_buffer=None, solute_dielectric=None,
                 solvent_dielectric=None, salt_concentration=None, temperature=None,
                 probe_radius=None, ion_radius=None, tolerance=None, max_iterations=None,
                 fillratio=None, fscale=None, fine_dims=None, fine_center=None,
                 coarse_dims=None, coarse_spacing=None, coarse_center=None):
        self._c = PbParamsC()
        load_library().mmpbsa_pb_params_default(C.byref(self._c))
        _opt(self._c, grid_spacing=grid_spacing, grid_buffer=grid_buffer,
             solute_dielectric=solute_dielectric, solvent_dielectric=solvent_dielectric,
             salt_concentration=salt_concentration, temperature=temperature,
             probe_radius=probe_radius, ion_radius=ion_radius, tolerance=tolerance,
             max_iterations=None if max_iterations is None else int(max_iterations),
             fillratio=fillratio, fscale=None if fscale is None else int(fscale))
        if fine_dims is not None:
            for i in range(3):
                self._c.fine_dims[i] = int(fine_dims[i])
        if fine_center is not None:
            self._c.fine_centered = 1
            for i in range(3):
                self._c.fine_center[i] = float(fine_center[i])
        if coarse_dims is not None:
            if coarse_spacing is None or coarse_center is None:
                raise ValueError("coarse_dims needs coarse_spacing and coarse_center")
            self._c.coarse_spacing = float(coarse_spacing)
            for i in range(3):
                self._c.coarse_dims[i] = int(coarse_dims[i])
                self._c.coarse_center[i] = float(coarse_center[i])

    grid_spacing = property(lambda s: s._c.grid_spacing)
    solute_dielectric = property(lambda s: s._c.solute_dielectric)
    solvent_dielectric = property(lambda s: s._c.solvent_dielectric)
    salt_concentration = property(lambda s: s._c.salt_concentration)
    temperature = property(lambda s: s._c.temperature)


class SaParams:
    """lib.rs:1628-1676; defaults sa_energy.rs:23-32."""

    def __init__(self, surface_tension=None, offset=None, probe_radius=None,
                 n_sphere_points=None):
        self._c = SaParamsC()
        load_library().mmpbsa_sa_params_default(C.byref(self._c))
        _opt(self._c, surface_tension=surface_tension, offset=offset, probe_radius=probe_radius,
             n_sphere_points=None if n_sphere_points is None else int(n_sphere_points))

    surface_tension = property(lambda s: s._c.surface_tension)
    offset = property(lambda s: s._c.offset)
    probe_radius = property(lambda s: s._c.probe_radius)
    n_sphere_points = property(lambda s: s._c.n_sphere_points)


# ---- results ----------------------------------------------------------------
class MmEnergy:
    def __init__(self, t: MmTermsC):
        for k, _ in MmTermsC._fields_:
            setattr(self, k, getattr(t, k))

    @property
    def total(self):
        return (self.bond + self.angle + self.dihedral + self.vdw + self.elec + self.vdw_14
                + self.elec_14)


class GbEnergy:
    def __init__(self, total, radii):
        self.total = total
        self._r = radii

    def born_radii(self):
        return self._r.copy()


class SaEnergy:
    def __init__(self, total, total_sasa, per_atom):
        self.total = total
        self.total_sasa = total_sasa
        self._p = per_atom

    def per_atom_sasa(self):
        return self._p.copy()


class PbEnergy:
    def __init__(self, r: PbResultC):
        self.total = r.total
        self.e_solv = r.solv.energy
        self.e_ref = r.ref.energy
        self.converged = bool(r.solv.converged and r.ref.converged)
        self.iterations = (r.solv.iterations, r.ref.iterations)
        self.final_residual = (r.solv.final_residual, r.ref.final_residual)
        self.fine_dims = tuple(r.fine_dims)
        self.coarse_dims = tuple(r.coarse_dims)
        # extensions: whether each fine solve took its boundary from the
        # coarse focusing solve, and that solve's cycle count
        self.focused = (bool(r.solv.focused), bool(r.ref.focused))
        self.coarse_iterations = (r.solv.coarse_iterations, r.ref.coarse_iterations)


class FrameEnergy:
    """FrameEnergy getters (lib.rs:1760-1850), *_gb aliases of *_polar."""

    def __init__(self, t: FrameTermsC):
        self._c = FrameTermsC.from_buffer_copy(t)
        for part, src in (("complex", t.complex_), ("receptor", t.receptor), ("ligand", t.ligand)):
            setattr(self, part + "_mm", src.mm)
            setattr(self, part + "_polar", src.polar)
            setattr(self, part + "_gb", src.polar)
            setattr(self, part + "_sa", src.sa)
            setattr(self, part + "_total", src.mm + src.polar + src.sa)
        self.delta_mm = t.delta_mm
        self.delta_polar = t.delta_polar
        self.delta_gb = t.delta_polar
        self.delta_sa = t.delta_sa
        self.delta_total = t.delta_total


def _terms_array(frames):
    arr = (FrameTermsC * max(len(frames), 1))()
    for k, f in enumerate(frames):
        arr[k] = f._c
    return arr


class BindingResult:
    """BindingResult (binding.rs:704-743): per-frame energies with means,
    sample standard deviations and the SEM of delta_total, computed by the
    library (mmpbsa_binding_stats_compute) in the reference's order."""

    def __init__(self, frames):
        self._frames = frames
        lib = load_library()
        st = _abi.BindingStatsC()
        _check(lib.mmpbsa_binding_stats_compute(_terms_array(frames), len(frames), C.byref(st)),
               lib)
        for k, _ in _abi.BindingStatsC._fields_:
            setattr(self, k, getattr(st, k))
        self.mean_delta_gb = self.mean_delta_polar
        self.std_delta_gb = self.std_delta_polar

    def frames(self):
        return list(self._frames)


class EntropyEstimate:
    """EntropyEstimate getters (lib.rs:1993-2011)."""

    def __init__(self, minus_tds, method):
        self.minus_tds = minus_tds
        self.method = method


def interaction_entropy(frames, temperature=298.15):
    """interaction_entropy (entropy.rs:53-78; lib.rs:2216-2224): -T dS from the
    fluctuation of delta_mm over frames, or None with fewer than two frames."""
    frames = list(frames)
    if len(frames) < 2:
        return None
    lib = load_library()
    out = C.c_double()
    _check(lib.mmpbsa_interaction_entropy(_terms_array(frames), len(frames), float(temperature),
                                          C.byref(out)), lib)
    return EntropyEstimate(out.value, "interaction_entropy")


def quasi_harmonic_entropy(trajectory_3d, masses, temperature=298.15, device=0):
    """quasi_harmonic_entropy (entropy.rs:97-217; lib.rs:2226-2237): -T dS
    from the mass-weighted covariance of a subsystem's atomic fluctuations over
    frames, trajectory_3d [F, N, 3] in Angstrom (passed through as the
    reference passes it) with masses [N] in amu, computed on the GPU; None with
    fewer than two frames or no atoms."""
    t = np.ascontiguousarray(np.asarray(trajectory_3d, dtype=np.float64))
    m = np.ascontiguousarray(np.asarray(masses, dtype=np.float64).reshape(-1))
    if t.ndim != 3 or t.shape[2] != 3:
        raise ValueError("trajectory_3d must have shape (frames, atoms, 3)")
    if t.shape[1] != m.shape[0]:
        raise ValueError("masses must have one value per atom")
    if t.shape[0] < 2 or t.shape[1] == 0:
        return None
    lib = load_library()
    out = C.c_double()
    dp = C.POINTER(C.c_double)
    _check(lib.mmpbsa_quasi_harmonic_entropy(int(device), t.ctypes.data_as(dp), t.shape[0],
                                             t.shape[1], m.ctypes.data_as(dp),
                                             float(temperature), C.byref(out)), lib)
    return EntropyEstimate(out.value, "quasi_harmonic")


class ResidueContribution:
    def __init__(self, c: ResidueContribC, label: str):
        self.residue_index = int(c.residue_index)
        self.residue_label = label
        self.vdw = c.vdw
        self.elec = c.elec
        self.polar = c.polar
        self.gb = c.polar
        self.sa = c.sa

    @property
    def total(self):
        return self.vdw + self.elec + self.polar + self.sa


class DecompositionResult:
    def __init__(self, rec, lig):
        self._r = rec
        self._l = lig

    def receptor_residues(self):
        return list(self._r)

    def ligand_residues(self):
        return list(self._l)


# ---- device contexts (cached per topology) ---------------------------------
def _check(rc, lib, ctx=None):
    if rc == _abi.MMPBSA_OK:
        return
    msg = lib.mmpbsa_last_error(ctx)
    msg = msg.decode() if msg else _abi.ERRORS.get(rc, "error")
    if rc in (1, 2):
        raise ValueError(msg)
    raise RuntimeError(f"mmpbsa-hip error {rc}: {msg}")


class _Ctx:
    def __init__(self, top: AmberTopology):
        self.lib = load_library()
        self.view = top.view()
        self.top = top
        self.h = C.c_void_p()
        _check(self.lib.mmpbsa_ctx_create(C.byref(self.view), 0, C.byref(self.h)), self.lib)
        self.splits = {}

    def __del__(self):
        try:
            for s in self.splits.values():
                self.lib.mmpbsa_split_destroy(s)
            if self.h:
                self.lib.mmpbsa_ctx_destroy(self.h)
        except Exception:
            pass

    def split(self, rec, lig):
        key = (tuple(int(r) for r in rec), tuple(int(r) for r in lig))
        s = self.splits.get(key)
        if s is None:
            ra = np.ascontiguousarray(key[0], dtype=np.int32)
            la = np.ascontiguousarray(key[1], dtype=np.int32)
            s = C.c_void_p()
            _check(self.lib.mmpbsa_split_create(self.h, ra.ctypes.data_as(_abi.c_int32_p), len(ra),
                                                la.ctypes.data_as(_abi.c_int32_p), len(la),
                                                C.byref(s)), self.lib, self.h)
            self.splits[key] = s
        return s


def _ctx(top: AmberTopology) -> _Ctx:
    c = getattr(top, "_mmpbsa_ctx", None)
    if c is None:
        c = _Ctx(top)
        object.__setattr__(top, "_mmpbsa_ctx", c)
    return c


def _angstrom(coords, n):
    x = np.asarray(coords, dtype=np.float64)
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError("coords must be an (N, 3) array")
    if x.shape[0] != n:
        raise ValueError(f"coords has {x.shape[0]} atoms, topology has {n}")
    return np.ascontiguousarray(x * 10.0)


def _dp(a):
    return a.ctypes.data_as(_abi.c_double_p)


# ---- functions (lib.rs:2065-2278) -------------------------------------------
def compute_mm_energy(topology: AmberTopology, coords) -> MmEnergy:
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    out = MmTermsC()
    _check(c.lib.mmpbsa_mm_energy(c.h, _dp(x), C.byref(out)), c.lib, c.h)
    return MmEnergy(out)


def compute_gb_energy(topology: AmberTopology, coords, gb_params: GbParams | None = None) -> GbEnergy:
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    p = (gb_params or GbParams())._c
    radii = np.zeros(topology.n_atoms)
    tot = C.c_double()
    _check(c.lib.mmpbsa_gb_energy(c.h, _dp(x), C.byref(p), _dp(radii), C.byref(tot)), c.lib, c.h)
    return GbEnergy(tot.value, radii)


def compute_sa_energy(topology: AmberTopology, coords, sa_params: SaParams | None = None) -> SaEnergy:
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    p = (sa_params or SaParams())._c
    per = np.zeros(topology.n_atoms)
    ts, t = C.c_double(), C.c_double()
    _check(c.lib.mmpbsa_sa_energy(c.h, _dp(x), C.byref(p), _dp(per), C.byref(ts), C.byref(t)),
           c.lib, c.h)
    return SaEnergy(t.value, ts.value, per)


def compute_pb_energy(topology: AmberTopology, coords, pb_params: PbParams | None = None,
                      solute_residues=None) -> PbEnergy:
    """compute_pb_energy_solvated (pb_energy.rs:469-482): with solute_residues
    the topology is stripped to those residues on the library side
    (mmpbsa_pb_energy_solvated) before the PB solve."""
    p = (pb_params or PbParams())._c
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    out = PbResultC()
    if solute_residues is None:
        _check(c.lib.mmpbsa_pb_energy(c.h, _dp(x), C.byref(p), C.byref(out)), c.lib, c.h)
    else:
        res = np.ascontiguousarray([int(r) for r in solute_residues], dtype=np.int32)
        if res.size and res.min() < 0:
            raise ValueError(f"Residue index {int(res.min())} out of range "
                             f"(0-{topology.n_residues - 1})")
        _check(c.lib.mmpbsa_pb_energy_solvated(c.h, _dp(x), C.byref(p),
                                               res.ctypes.data_as(_abi.c_int32_p), len(res),
                                               C.byref(out)), c.lib, c.h)
    return PbEnergy(out)


def compute_gb_energy_from_radii(topology: AmberTopology, coords, born_radii,
                                 gb_params: GbParams | None = None) -> float:
    """compute_gb_energy_from_radii (gb_energy.rs:283-355): the GB energy of
    given Born radii (Angstrom), with no radii computation."""
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    R = np.ascontiguousarray(born_radii, dtype=np.float64)
    if R.shape != (topology.n_atoms,):
        raise ValueError(f"born_radii has shape {R.shape}, topology has {topology.n_atoms} atoms")
    p = (gb_params or GbParams())._c
    tot = C.c_double()
    _check(c.lib.mmpbsa_gb_energy_from_radii(c.h, _dp(x), C.byref(p), _dp(R), C.byref(tot)),
           c.lib, c.h)
    return tot.value


class GbCache:
    """GbCache (gb_energy.rs:66-120): Born radii reused across frames whose
    RMSD (Angstrom) to the last computed frame is below rmsd_threshold.
    last_reused tells whether the last compute_gb_energy_cached call reused
    them."""

    def __init__(self, rmsd_threshold: float = 0.5):
        self.lib = load_library()
        self.rmsd_threshold = float(rmsd_threshold)
        self.h = C.c_void_p()
        _check(self.lib.mmpbsa_gb_cache_create(self.rmsd_threshold, C.byref(self.h)), self.lib)
        self.last_reused = False

    def __del__(self):
        try:
            if self.h:
                self.lib.mmpbsa_gb_cache_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


def compute_gb_energy_cached(topology: AmberTopology, coords, gb_params: GbParams | None,
                             cache: GbCache) -> GbEnergy:
    """compute_gb_energy_cached (gb_energy.rs:372-387)."""
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    p = (gb_params or GbParams())._c
    radii = np.zeros(topology.n_atoms)
    tot = C.c_double()
    hit = C.c_int32()
    _check(c.lib.mmpbsa_gb_energy_cached(c.h, cache.h, _dp(x), C.byref(p), _dp(radii),
                                         C.byref(tot), C.byref(hit)), c.lib, c.h)
    cache.last_reused = bool(hit.value)
    return GbEnergy(tot.value, radii)


def _method(gb_params, pb_params):
    # build_solvation_method (lib.rs:2105-2115): PB wins when both are given
    if pb_params is not None:
        return 1, GbParams()._c, pb_params._c
    return 0, (gb_params or GbParams())._c, PbParams()._c


def compute_binding_energy_single_frame(topology: AmberTopology, coords, receptor_residues,
                                        ligand_residues, gb_params=None, pb_params=None,
                                        sa_params=None) -> FrameEnergy:
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    s = c.split(receptor_residues, ligand_residues)
    m, g, p = _method(gb_params, pb_params)
    a = (sa_params or SaParams())._c
    out = FrameTermsC()
    _check(c.lib.mmpbsa_frame_energy(s, _dp(x), m, C.byref(g), C.byref(p), C.byref(a),
                                     C.byref(out)), c.lib, c.h)
    return FrameEnergy(out)


def compute_binding_energy_frames(topology: AmberTopology, frames, receptor_residues,
                                  ligand_residues, gb_params=None, pb_params=None,
                                  sa_params=None, batched=True) -> BindingResult:
    """compute_binding_energy (binding.rs:461-744) over in-memory frames
    ([F, N, 3] nm) instead of a trajectory file; batched=True overlaps frames
    on concurrent device streams."""
    c = _ctx(topology)
    f = np.asarray(frames, dtype=np.float64)
    if f.ndim != 3 or f.shape[1] != topology.n_atoms or f.shape[2] != 3:
        raise ValueError("frames must be an (F, N, 3) array")
    if f.shape[0] == 0:
        raise ValueError("No frames found in trajectory")
    x = np.ascontiguousarray(f * 10.0)
    s = c.split(receptor_residues, ligand_residues)
    m, g, p = _method(gb_params, pb_params)
    a = (sa_params or SaParams())._c
    out = (FrameTermsC * f.shape[0])()
    _check(c.lib.mmpbsa_frames_energy(s, _dp(x), f.shape[0], m, C.byref(g), C.byref(p),
                                      C.byref(a), 1 if batched else 0, out), c.lib, c.h)
    return BindingResult([FrameEnergy(t) for t in out])


class DeviceArray:
    """A float64 array in HBM, allocated and filled through the library (no
    other GPU runtime involved).  Slicing the first axis gives a view.
    Extension: the reference has no device-resident inputs."""

    def __init__(self, host=None, *, _base=None, _ptr=None, _shape=None, device=0):
        self._base = _base
        if _base is not None:
            self.ptr, self.shape = _ptr, _shape
            return
        a = np.ascontiguousarray(np.asarray(host, dtype=np.float64))
        self.shape = a.shape
        lib = _abi.load_library()
        p = C.c_void_p()
        _check(lib.mmpbsa_device_alloc(device, a.nbytes, C.byref(p)), lib, None)
        self.ptr = p.value
        self._lib = lib
        _check(lib.mmpbsa_device_upload(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                        a.nbytes), lib, None)

    def __getitem__(self, sl):
        if not isinstance(sl, slice) or sl.step not in (None, 1):
            raise TypeError("DeviceArray supports contiguous slices of the first axis")
        a, b, _ = sl.indices(self.shape[0])
        row = int(np.prod(self.shape[1:])) * 8
        return DeviceArray(_base=self, _ptr=self.ptr + a * row,
                           _shape=(max(0, b - a),) + tuple(self.shape[1:]))

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float64)
        lib = _abi.load_library()
        _check(lib.mmpbsa_device_download(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                          out.nbytes), lib, None)
        return out

    def __del__(self):
        if getattr(self, "_base", None) is None and getattr(self, "ptr", None):
            try:
                self._lib.mmpbsa_device_free(C.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None


def compute_binding_energy(topology, trajectory_path, receptor_residues, ligand_residues,
                           gb_params=None, pb_params=None, sa_params=None, trajectory_format=None,
                           has_box=False, stride=1, start_frame=0,
                           end_frame=None) -> BindingResult:
    """compute_binding_energy (lib.rs:2117-2160; binding.rs:461-744): topology
    is an AmberTopology or a prmtop path (OSError if it cannot be read); the
    trajectory is mdcrd (A, has_box) unless trajectory_format == "dcd" (nm);
    frames start_frame, start_frame + stride, ... below end_frame, sliced to
    the complex, reimaged when the file has a box, in batches of 64 through
    the batched device path.  Other failures raise ValueError, as there."""
    top = topology if isinstance(topology, AmberTopology) else read_prmtop(topology)
    c = _ctx(top)
    s = c.split(receptor_residues, ligand_residues)
    m, g, p = _method(gb_params, pb_params)
    a = (sa_params or SaParams())._c
    fmt = 1 if trajectory_format == "dcd" else 0
    n = C.c_int32()
    rc = c.lib.mmpbsa_binding_energy_trajectory(
        s, os.fsencode(str(trajectory_path)), fmt, 1 if has_box else 0, max(1, int(stride)),
        int(start_frame), -1 if end_frame is None else int(end_frame), m, C.byref(g), C.byref(p),
        C.byref(a), C.byref(n))
    if rc == _abi.MMPBSA_ERR_IO:
        msg = c.lib.mmpbsa_last_error(c.h)
        raise ValueError(msg.decode() if msg else "trajectory error")
    _check(rc, c.lib, c.h)
    out = (FrameTermsC * max(n.value, 1))()
    _check(c.lib.mmpbsa_trajectory_frames(s, out, n.value, None), c.lib, c.h)
    return BindingResult([FrameEnergy(out[k]) for k in range(n.value)])


def compute_binding_energy_device_frames(topology: AmberTopology, frames_angstrom: DeviceArray,
                                         receptor_residues, ligand_residues, gb_params=None,
                                         pb_params=None, sa_params=None,
                                         batched=True) -> BindingResult:
    """As compute_binding_energy_frames, with the frames already in HBM: a
    DeviceArray [F, N, 3] in Angstrom (the library's unit; nothing is copied
    or converted).  Extension: the reference reads frames from files only."""
    c = _ctx(topology)
    t = frames_angstrom
    if (not isinstance(t, DeviceArray) or len(t.shape) != 3 or t.shape[1] != topology.n_atoms
            or t.shape[2] != 3):
        raise ValueError("frames must be a DeviceArray of shape (F, N, 3)")
    if t.shape[0] == 0:
        raise ValueError("No frames found in trajectory")
    s = c.split(receptor_residues, ligand_residues)
    m, g, p = _method(gb_params, pb_params)
    a = (sa_params or SaParams())._c
    out = (FrameTermsC * t.shape[0])()
    _check(c.lib.mmpbsa_frames_energy_device(s, C.c_void_p(t.ptr), t.shape[0], m, C.byref(g),
                                             C.byref(p), C.byref(a), 1 if batched else 0, out),
           c.lib, c.h)
    return BindingResult([FrameEnergy(tt) for tt in out])


class RequestStats:
    """Latency of one served request (milliseconds from submission)."""

    def __init__(self, st: RequestStatsC):
        self.queue_ms = st.queue_ms
        self.total_ms = st.total_ms
        self.n_frames = st.n_frames


class RequestSummary:
    """Statistics of a served request computed on the device when its last
    frame finishes: the BindingResult statistics (binding.rs:704-743) as
    attributes of the same names, and interaction_entropy (entropy.rs:53-78)
    at the request's temperature (None for a single frame)."""

    def __init__(self, s: "_abi.RequestSummaryC"):
        for k, _ in _abi.BindingStatsC._fields_:
            setattr(self, k, getattr(s.stats, k))
        self.temperature = s.temperature
        self.entropy = (None if s.minus_tds != s.minus_tds
                        else EntropyEstimate(s.minus_tds, "interaction_entropy"))


class Server:
    """Continuous-batching MM-PBSA server (extension: the reference has no
    served path).  submit() copies the frames ([F, N, 3] nm) and returns a
    ticket at once; a worker thread schedules frames of the waiting requests,
    oldest first, into shared batched rounds under an HBM budget; poll() / wait()
    return (BindingResult, RequestStats).  A frame's result is identical alone
    and in any mix of requests."""

    def __init__(self, hbm_budget_bytes=0, max_round_frames=16, device=0):
        self.lib = load_library()
        p = _abi.ServerParamsC()
        self.lib.mmpbsa_server_params_default(C.byref(p))
        p.hbm_budget_bytes = int(hbm_budget_bytes)
        p.max_round_frames = int(max_round_frames)
        self.h = C.c_void_p()
        _check(self.lib.mmpbsa_server_create(C.byref(p), device, C.byref(self.h)), self.lib)
        self._keep = {}   # ticket -> (context, frame count): contexts outlive requests

    def submit(self, topology: AmberTopology, frames, receptor_residues, ligand_residues,
               gb_params=None, pb_params=None, sa_params=None) -> int:
        c = _ctx(topology)
        f = np.asarray(frames, dtype=np.float64)
        if f.ndim == 2:
            f = f[None]
        if f.ndim != 3 or f.shape[1] != topology.n_atoms or f.shape[2] != 3 or f.shape[0] == 0:
            raise ValueError("frames must be an (F, N, 3) array with F > 0")
        x = np.ascontiguousarray(f * 10.0)
        s = c.split(receptor_residues, ligand_residues)
        m, g, p = _method(gb_params, pb_params)
        a = (sa_params or SaParams())._c
        t = C.c_uint64()
        _check(self.lib.mmpbsa_server_submit(self.h, s, _dp(x), f.shape[0], m, C.byref(g),
                                             C.byref(p), C.byref(a), C.byref(t)), self.lib)
        self._keep[t.value] = (c, f.shape[0])
        return t.value

    def _finish(self, fn, ticket, summary=False):
        c, n = self._keep[ticket]
        out = (FrameTermsC * n)()
        st = RequestStatsC()
        sm = _abi.RequestSummaryC()
        args = (self.h, C.c_uint64(ticket), out, C.byref(st)) + ((C.byref(sm),) if summary else ())
        rc = fn(*args)
        if rc == _abi.MMPBSA_ERR_NOT_READY:
            return None
        del self._keep[ticket]
        _check(rc, self.lib)
        res = (BindingResult([FrameEnergy(t) for t in out]), RequestStats(st))
        return res + (RequestSummary(sm),) if summary else res

    def poll(self, ticket: int):
        """(BindingResult, RequestStats), or None while frames are pending."""
        return self._finish(self.lib.mmpbsa_server_poll, ticket)

    def wait(self, ticket: int):
        return self._finish(self.lib.mmpbsa_server_wait, ticket)

    def poll_summary(self, ticket: int):
        """(BindingResult, RequestStats, RequestSummary): the request's statistics
        and interaction entropy computed on the device; None while pending."""
        return self._finish(self.lib.mmpbsa_server_poll_summary, ticket, True)

    def wait_summary(self, ticket: int):
        return self._finish(self.lib.mmpbsa_server_wait_summary, ticket, True)

    def close(self):
        if self.h:
            self.lib.mmpbsa_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decompose_binding_energy(topology: AmberTopology, coords, receptor_residues,
                             ligand_residues, gb_params=None, sa_params=None) -> DecompositionResult:
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    s = c.split(receptor_residues, ligand_residues)
    g = (gb_params or GbParams())._c
    a = (sa_params or SaParams())._c
    rec = (ResidueContribC * max(len(receptor_residues), 1))()
    lig = (ResidueContribC * max(len(ligand_residues), 1))()
    _check(c.lib.mmpbsa_decompose(s, _dp(x), C.byref(g), C.byref(a), rec, lig), c.lib, c.h)
    lab = topology.residue_labels
    return DecompositionResult(
        [ResidueContribution(rec[k], lab[int(rec[k].residue_index)])
         for k in range(len(receptor_residues))],
        [ResidueContribution(lig[k], lab[int(lig[k].residue_index)])
         for k in range(len(ligand_residues))])


def decompose_binding_energy_pb(topology: AmberTopology, coords, receptor_residues,
                                ligand_residues, pb_params=None, sa_params=None):
    """decompose_binding_energy with per-residue PB polar terms (an extension:
    the reference decomposes the GB term only).  Each residue's polar is the
    sum over its atoms of 0.5 q (dphi_complex - dphi_isolated) at the atom,
    dphi = phi_solvated - phi_reference; the polars of all residues sum to
    dG_PB.  Returns (DecompositionResult, dG_PB)."""
    c = _ctx(topology)
    x = _angstrom(coords, topology.n_atoms)
    s = c.split(receptor_residues, ligand_residues)
    p = (pb_params or PbParams())._c
    a = (sa_params or SaParams())._c
    rec = (ResidueContribC * max(len(receptor_residues), 1))()
    lig = (ResidueContribC * max(len(ligand_residues), 1))()
    d = C.c_double(0.0)
    _check(c.lib.mmpbsa_decompose_pb(s, _dp(x), C.byref(p), C.byref(a), rec, lig, C.byref(d)),
           c.lib, c.h)
    lab = topology.residue_labels
    return DecompositionResult(
        [ResidueContribution(rec[k], lab[int(rec[k].residue_index)])
         for k in range(len(receptor_residues))],
        [ResidueContribution(lig[k], lab[int(lig[k].residue_index)])
         for k in range(len(ligand_residues))]), d.value


def extract_subtopology(topology: AmberTopology, atom_indices) -> AmberTopology:
    return _extract_subtopology(topology, atom_indices)
