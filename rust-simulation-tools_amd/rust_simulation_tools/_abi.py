"""ctypes mirror of include/mmpbsa_hip.h and the loader of the HIP library.

The product library ``libmmpbsa_hip.so`` is built in-tree next to its sources
(``rust-simulation-tools_amd/csrc``).  Loading fails loudly when it is absent:
there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "csrc", "libmmpbsa_hip.so")

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)
c_int64_p = C.POINTER(C.c_int64)

MMPBSA_OK = 0
MMPBSA_ERR_NOT_READY = 6
MMPBSA_ERR_TOO_LARGE = 7
ERRORS = {
    1: "invalid argument",
    2: "residue index out of range",
    3: "device error",
    4: "out of device memory",
    5: "no gfx950 device",
    6: "not ready",
    7: "request too large",
    8: "file missing, unreadable or malformed",
}
MMPBSA_ERR_IO = 8


class TopologyView(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("n_residues", C.c_int32),
        ("n_types", C.c_int32),
        ("charges", c_double_p),
        ("charges_amber", c_double_p),
        ("radii", c_double_p),
        ("screen", c_double_p),
        ("atom_type", c_int32_p),
        ("nb_parm_index", c_int64_p),
        ("lj_acoef", c_double_p),
        ("lj_bcoef", c_double_p),
        ("n_lj", C.c_int32),
        ("residue_pointers", c_int32_p),
        ("n_bonds", C.c_int32),
        ("bonds", c_int32_p),
        ("bond_k", c_double_p),
        ("bond_r0", c_double_p),
        ("n_bond_types", C.c_int32),
        ("n_angles", C.c_int32),
        ("angles", c_int32_p),
        ("angle_k", c_double_p),
        ("angle_theta0", c_double_p),
        ("n_angle_types", C.c_int32),
        ("n_dihedrals", C.c_int32),
        ("dihedrals", c_int32_p),
        ("dihedral_pk", c_double_p),
        ("dihedral_pn", c_double_p),
        ("dihedral_phase", c_double_p),
        ("n_dihedral_types", C.c_int32),
        ("num_excluded", c_int32_p),
        ("excluded_atoms", c_int32_p),
        ("n_excluded", C.c_int64),
        ("scee", C.c_double),
        ("scnb", C.c_double),
    ]


class GbParamsC(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("solute_dielectric", C.c_double),
        ("solvent_dielectric", C.c_double),
        ("salt_concentration", C.c_double),
        ("temperature", C.c_double),
        ("offset", C.c_double),
        ("rgbmax", C.c_double),
        ("cutoff", C.c_double),
    ]


class PbParamsC(C.Structure):
    _fields_ = [
        ("grid_spacing", C.c_double),
        ("grid_buffer", C.c_double),
        ("solute_dielectric", C.c_double),
        ("solvent_dielectric", C.c_double),
        ("salt_concentration", C.c_double),
        ("temperature", C.c_double),
        ("probe_radius", C.c_double),
        ("ion_radius", C.c_double),
        ("tolerance", C.c_double),
        ("max_iterations", C.c_int64),
        ("fillratio", C.c_double),
        ("fscale", C.c_int64),
        ("fine_dims", C.c_int32 * 3),
        ("coarse_dims", C.c_int32 * 3),
        ("coarse_spacing", C.c_double),
        ("coarse_center", C.c_double * 3),
        ("fine_centered", C.c_int32),
        ("fine_center", C.c_double * 3),
    ]


class SaParamsC(C.Structure):
    _fields_ = [
        ("surface_tension", C.c_double),
        ("offset", C.c_double),
        ("probe_radius", C.c_double),
        ("n_sphere_points", C.c_int32),
    ]


class MmTermsC(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("bond", "angle", "dihedral", "vdw", "elec", "vdw_14", "elec_14")]


class PbSolveInfoC(C.Structure):
    _fields_ = [
        ("energy", C.c_double),
        ("final_residual", C.c_double),
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("focused", C.c_int32),
        ("coarse_iterations", C.c_int32),
    ]


class PbResultC(C.Structure):
    _fields_ = [
        ("total", C.c_double),
        ("solv", PbSolveInfoC),
        ("ref", PbSolveInfoC),
        ("fine_dims", C.c_int32 * 3),
        ("coarse_dims", C.c_int32 * 3),
    ]


class SubsystemEnergyC(C.Structure):
    _fields_ = [("mm", C.c_double), ("polar", C.c_double), ("sa", C.c_double)]


class FrameTermsC(C.Structure):
    _fields_ = [
        ("complex_", SubsystemEnergyC),
        ("receptor", SubsystemEnergyC),
        ("ligand", SubsystemEnergyC),
        ("delta_mm", C.c_double),
        ("delta_polar", C.c_double),
        ("delta_sa", C.c_double),
        ("delta_total", C.c_double),
    ]


class ResidueContribC(C.Structure):
    _fields_ = [
        ("residue_index", C.c_int32),
        ("vdw", C.c_double),
        ("elec", C.c_double),
        ("polar", C.c_double),
        ("sa", C.c_double),
    ]


class PrmtopInfoC(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_residues", C.c_int32), ("n_types", C.c_int32),
                ("masses", c_double_p), ("atom_names", C.c_void_p),
                ("residue_labels", C.c_void_p)]


class DcdHeaderC(C.Structure):
    _fields_ = [("n_frames", C.c_int64), ("n_atoms", C.c_int64), ("start_timestep", C.c_int32),
                ("timestep_interval", C.c_int32), ("timestep", C.c_float),
                ("has_unit_cell", C.c_int32), ("is_charmm", C.c_int32), ("is_64bit", C.c_int32),
                ("big_endian", C.c_int32), ("first_frame_offset", C.c_int64),
                ("frame_size", C.c_int64)]


class BindingStatsC(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("mean_delta_mm", "mean_delta_polar", "mean_delta_sa", "mean_delta_total",
                 "std_delta_total", "std_delta_mm", "std_delta_polar", "std_delta_sa",
                 "sem_delta_total")]


class ServerParamsC(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_round_frames", C.c_int32)]


class RequestStatsC(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("total_ms", C.c_double), ("n_frames", C.c_int32),
                ("status", C.c_int32)]


class RequestSummaryC(C.Structure):
    _fields_ = [("stats", BindingStatsC), ("minus_tds", C.c_double), ("temperature", C.c_double)]


class TuneEntryC(C.Structure):
    _fields_ = [("nodes_min", C.c_int64), ("nodes_max", C.c_int64), ("solves_min", C.c_int32),
                ("solves_max", C.c_int32), ("smoother_chunks", C.c_int32),
                ("smoother_region", C.c_int32)]


class TuningC(C.Structure):
    _fields_ = [("check_cycles", C.c_int32), ("gb_chunks", C.c_int32), ("n_entries", C.c_int32),
                ("entries", C.POINTER(TuneEntryC))]


# Every symbol include/mmpbsa_hip.h declares, with its ctypes signature.
P = C.c_void_p
SIGNATURES = {
    "mmpbsa_init": (C.c_int, [c_int32_p]),
    "mmpbsa_version": (C.c_char_p, []),
    "mmpbsa_gb_params_default": (None, [C.POINTER(GbParamsC)]),
    "mmpbsa_pb_params_default": (None, [C.POINTER(PbParamsC)]),
    "mmpbsa_sa_params_default": (None, [C.POINTER(SaParamsC)]),
    "mmpbsa_ctx_create": (C.c_int, [C.POINTER(TopologyView), C.c_int32, C.POINTER(P)]),
    "mmpbsa_ctx_destroy": (C.c_int, [P]),
    "mmpbsa_last_error": (C.c_char_p, [P]),
    "mmpbsa_ctx_n_atoms": (C.c_int32, [P]),
    "mmpbsa_mm_energy": (C.c_int, [P, c_double_p, C.POINTER(MmTermsC)]),
    "mmpbsa_gb_energy": (C.c_int, [P, c_double_p, C.POINTER(GbParamsC), c_double_p, c_double_p]),
    "mmpbsa_sa_energy": (C.c_int, [P, c_double_p, C.POINTER(SaParamsC), c_double_p, c_double_p,
                                   c_double_p]),
    "mmpbsa_pb_energy": (C.c_int, [P, c_double_p, C.POINTER(PbParamsC), C.POINTER(PbResultC)]),
    "mmpbsa_pb_energy_solvated": (C.c_int, [P, c_double_p, C.POINTER(PbParamsC), c_int32_p,
                                            C.c_int32, C.POINTER(PbResultC)]),
    "mmpbsa_gb_energy_from_radii": (C.c_int, [P, c_double_p, C.POINTER(GbParamsC), c_double_p,
                                              c_double_p]),
    "mmpbsa_gb_cache_create": (C.c_int, [C.c_double, C.POINTER(P)]),
    "mmpbsa_gb_cache_destroy": (C.c_int, [P]),
    "mmpbsa_gb_energy_cached": (C.c_int, [P, P, c_double_p, C.POINTER(GbParamsC), c_double_p,
                                          c_double_p, c_int32_p]),
    "mmpbsa_gb_energy_noalloc": (C.c_int, [P, C.c_void_p, C.POINTER(GbParamsC), C.c_void_p,
                                           C.c_void_p]),
    "mmpbsa_split_create": (C.c_int, [P, c_int32_p, C.c_int32, c_int32_p, C.c_int32, C.POINTER(P)]),
    "mmpbsa_split_destroy": (C.c_int, [P]),
    "mmpbsa_split_atoms": (C.c_int, [P, C.c_int32, c_int32_p, C.c_int32, c_int32_p]),
    "mmpbsa_frame_energy": (C.c_int, [P, c_double_p, C.c_int32, C.POINTER(GbParamsC),
                                      C.POINTER(PbParamsC), C.POINTER(SaParamsC),
                                      C.POINTER(FrameTermsC)]),
    "mmpbsa_frames_energy": (C.c_int, [P, c_double_p, C.c_int32, C.c_int32, C.POINTER(GbParamsC),
                                       C.POINTER(PbParamsC), C.POINTER(SaParamsC), C.c_int32,
                                       C.POINTER(FrameTermsC)]),
    "mmpbsa_frames_energy_device": (C.c_int, [P, C.c_void_p, C.c_int32, C.c_int32,
                                              C.POINTER(GbParamsC), C.POINTER(PbParamsC),
                                              C.POINTER(SaParamsC), C.c_int32,
                                              C.POINTER(FrameTermsC)]),
    "mmpbsa_decompose": (C.c_int, [P, c_double_p, C.POINTER(GbParamsC), C.POINTER(SaParamsC),
                                   C.POINTER(ResidueContribC), C.POINTER(ResidueContribC)]),
    "mmpbsa_decompose_pb": (C.c_int, [P, c_double_p, C.POINTER(PbParamsC), C.POINTER(SaParamsC),
                                      C.POINTER(ResidueContribC), C.POINTER(ResidueContribC),
                                      c_double_p]),
    "mmpbsa_prmtop_read": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mmpbsa_prmtop_last_error": (C.c_char_p, []),
    "mmpbsa_prmtop_destroy": (C.c_int, [P]),
    "mmpbsa_prmtop_view": (C.c_int, [P, C.POINTER(TopologyView)]),
    "mmpbsa_prmtop_info_get": (C.c_int, [P, C.POINTER(PrmtopInfoC)]),
    "mmpbsa_traj_last_error": (C.c_char_p, []),
    "mmpbsa_dcd_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mmpbsa_dcd_close": (C.c_int, [P]),
    "mmpbsa_dcd_get_header": (C.c_int, [P, C.POINTER(DcdHeaderC), c_int32_p]),
    "mmpbsa_dcd_title": (C.c_int, [P, C.c_int32, C.c_char_p, C.c_int32]),
    "mmpbsa_dcd_seek": (C.c_int, [P, C.c_int64]),
    "mmpbsa_dcd_current": (C.c_int64, [P]),
    "mmpbsa_dcd_read_frame": (C.c_int, [P, c_double_p, c_double_p, c_int32_p]),
    "mmpbsa_mdcrd_open": (C.c_int, [C.c_char_p, C.c_int64, C.c_int32, C.POINTER(P)]),
    "mmpbsa_mdcrd_close": (C.c_int, [P]),
    "mmpbsa_mdcrd_current": (C.c_int64, [P]),
    "mmpbsa_mdcrd_read_frame": (C.c_int, [P, c_double_p, c_double_p, c_int32_p]),
    "mmpbsa_reimage": (C.c_int, [c_double_p, C.c_int32, c_int32_p, C.c_int32, c_double_p]),
    "mmpbsa_binding_energy_trajectory": (C.c_int, [P, C.c_char_p, C.c_int32, C.c_int32,
                                                   C.c_int64, C.c_int64, C.c_int64, C.c_int32,
                                                   C.POINTER(GbParamsC), C.POINTER(PbParamsC),
                                                   C.POINTER(SaParamsC), c_int32_p]),
    "mmpbsa_trajectory_frames": (C.c_int, [P, C.POINTER(FrameTermsC), C.c_int32, c_double_p]),
    "mmpbsa_binding_stats_compute": (C.c_int, [C.POINTER(FrameTermsC), C.c_int32,
                                               C.POINTER(BindingStatsC)]),
    "mmpbsa_interaction_entropy": (C.c_int, [C.POINTER(FrameTermsC), C.c_int32, C.c_double,
                                             c_double_p]),
    "mmpbsa_quasi_harmonic_entropy": (C.c_int, [C.c_int32, c_double_p, C.c_int32, C.c_int32,
                                                c_double_p, C.c_double, c_double_p]),
    "mmpbsa_server_params_default": (None, [C.POINTER(ServerParamsC)]),
    "mmpbsa_server_create": (C.c_int, [C.POINTER(ServerParamsC), C.c_int32, C.POINTER(P)]),
    "mmpbsa_server_destroy": (C.c_int, [P]),
    "mmpbsa_server_submit": (C.c_int, [P, P, c_double_p, C.c_int32, C.c_int32,
                                       C.POINTER(GbParamsC), C.POINTER(PbParamsC),
                                       C.POINTER(SaParamsC), C.POINTER(C.c_uint64)]),
    "mmpbsa_server_poll": (C.c_int, [P, C.c_uint64, C.POINTER(FrameTermsC),
                                     C.POINTER(RequestStatsC)]),
    "mmpbsa_server_wait": (C.c_int, [P, C.c_uint64, C.POINTER(FrameTermsC),
                                     C.POINTER(RequestStatsC)]),
    "mmpbsa_server_poll_summary": (C.c_int, [P, C.c_uint64, C.POINTER(FrameTermsC),
                                             C.POINTER(RequestStatsC),
                                             C.POINTER(RequestSummaryC)]),
    "mmpbsa_server_wait_summary": (C.c_int, [P, C.c_uint64, C.POINTER(FrameTermsC),
                                             C.POINTER(RequestStatsC),
                                             C.POINTER(RequestSummaryC)]),
    "mmpbsa_device_alloc": (C.c_int, [C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]),
    "mmpbsa_device_free": (C.c_int, [C.c_void_p]),
    "mmpbsa_device_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "mmpbsa_device_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "mmpbsa_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "mmpbsa_tuning_set": (C.c_int, [C.POINTER(TuningC)]),
    "mmpbsa_tuning_get": (C.c_int, [C.POINTER(TuningC), C.POINTER(TuneEntryC), C.c_int32]),
    "mmpbsa_tuning_smoother_zc": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mmpbsa_tuning_smoother_region": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "mmpbsa_tuning_observe": (C.c_int, [C.c_int32]),
    "mmpbsa_tuning_observed": (C.c_int32, [c_int64_p, c_int32_p, c_int64_p, C.c_int32]),
    "mmpbsa_profile_enable": (C.c_int, [C.c_int32]),
    "mmpbsa_profile_read": (C.c_int, [c_double_p, c_int64_p, c_double_p, c_int64_p]),
    "mmpbsa_profile_kernel": (C.c_int, [C.c_char_p, c_double_p, c_int64_p, c_double_p]),
}

_lib = None


def load_library(path: str | None = None):
    """Load libmmpbsa_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(
            f"HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def as_ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))
