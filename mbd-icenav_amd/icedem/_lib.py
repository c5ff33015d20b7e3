"""ctypes binding of ``libicedem.so`` (declarations in ``include/icedem.h``).

The library is built in-tree (``mbd-icenav_amd/csrc/Makefile``) and loaded from
this directory.  There is no fallback: if the library is missing or no GPU is
visible, the product path raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_NAME = "libicedem.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

MAX_VERTS = 64
MAX_EDGE_SAMPLES = 16
SHIP_STATE_LEN = 15

_f = C.c_float
_i = C.c_int32
_p = C.c_void_p


class IceDemStatic(C.Structure):
    _fields_ = [
        ("n", _i), ("n_bonds", _i), ("n_floes", _i), ("max_v", _i), ("n_hull", _i),
        ("h_local_polys", _p), ("h_poly_mask", _p), ("h_masses", _p), ("h_inertia", _p),
        ("h_radii", _p), ("h_areas", _p), ("h_thicknesses", _p), ("h_bound_radius", _p),
        ("h_floe_of", _p), ("h_bond_i", _p), ("h_bond_j", _p), ("h_rest_lengths", _p),
        ("h_bond_strengths", _p), ("h_sigma_bar", _p), ("h_hull_local", _p),
    ]


class IceDemParams(C.Structure):
    _fields_ = [
        ("half_len", _f), ("half_beam", _f),
        ("n_edge_samples", _i),
        ("edge_t", _f * MAX_EDGE_SAMPLES), ("edge_omt", _f * MAX_EDGE_SAMPLES),
        ("ship_bound_radius", _f), ("ship_box_half_len", _f), ("ship_box_half_beam", _f),
        ("w_contact_factor", _f), ("w_contact_cap", _f), ("flexural_base", _f),
        ("pen_crit_ref", _f), ("pen_crit_thk_ref", _f), ("side_push_factor", _f),
        ("has_thrust_cap", _i),
        ("max_ship_ice_force", _f),
        ("submergence_pen_cap", _f), ("submergence_rate", _f), ("inv_rise_time", _f), ("buoyancy_coef", _f),
        ("elig_threshold", _f), ("eff_strength_collision", _f),
        ("youngs_modulus", _f), ("bond_stiffness_factor", _f),
        ("hertz_factor", _f), ("hertz_exponent", _f), ("ice_friction", _f), ("friction_cap", _f),
        ("contact_dilation2", _f),
        ("drag_ref_thickness", _f), ("drag_coef", _f), ("linear_drag_factor", _f), ("rot_drag_coef", _f),
        ("dt", _f), ("dt_ice", _f), ("velocity_clip", _f), ("angular_damping", _f),
        ("ang_cap_lo", _f), ("ang_cap_hi", _f), ("position_clip", _f),
        ("n_ice_steps", _i),
        ("max_fracture_thickness", _f), ("k_n", _f), ("k_s", _f),
        ("bond_cohesion", _f), ("bond_friction", _f), ("G_c_I", _f), ("G_c_II", _f), ("eta_bk", _f),
        ("use_auto_icenav", _i),
        ("ship_mode", _i), ("ctrl_thrust_max", _f), ("ctrl_yaw_rate_max", _f), ("ship_mass", _f), ("ship_Izz", _f),
        ("ice_brake_threshold", _f),
        ("M3inv", _f * 9), ("D3", _f * 9), ("Kp", _f * 9), ("Kd", _f * 9), ("Ki", _f * 9), ("D3pKd", _f * 9),
        ("B_pinv", _f * 18), ("B", _f * 18), ("n_max", _f * 6),
        ("T_n", _f), ("vc_surge", _f), ("vc_sway", _f), ("vc_yaw", _f),
        ("grid_x0", _f), ("grid_y0", _f), ("grid_cell", _f), ("grid_rsmall", _f),
        ("grid_nx", _i), ("grid_ny", _i), ("grid_bucket", _i), ("grid_bucket2", _i),
        ("force_lanes", _i),
        ("list_skin", _f),
        ("near_cand_margin", _f),
    ]


class IceDemState(C.Structure):
    _fields_ = [(name, _p) for name in (
        "d_pos", "d_vel", "d_rot", "d_omega", "d_sub", "d_active", "d_bond_intact", "d_bond_damage",
        "d_bond_dm0", "d_bond_dmmax", "d_bond_ds", "d_floe_elig", "d_ship")]


class IceDemInputs(C.Structure):
    _fields_ = [("d_setpoints", _p), ("d_v_water", _p),
                ("vw_stride_k", C.c_int64), ("vw_stride_c", C.c_int64), ("vw_stride_e", C.c_int64),
                ("d_v_water_ship", _p), ("d_controls", _p), ("n_knots", _i), ("hold", _i)]


class IceDemHistory(C.Structure):
    _fields_ = [(name, _p) for name in (
        "d_positions", "d_rotations", "d_bond_intact", "d_eta", "d_submergence", "d_J", "d_F")]


class IceDemDiag(C.Structure):
    _fields_ = [(name, _p) for name in ("d_pen_max", "d_force_mag", "d_F_ice_on_ship", "d_J", "d_rollout")]


# name -> (restype, argtypes)
BATCH_FRAMES = 0   # ICEDEM_BATCH_FRAMES
BATCH_SCENES = 1   # ICEDEM_BATCH_SCENES

SIGNATURES = {
    "icedem_status_string": (C.c_char_p, [C.c_int]),
    "icedem_last_error": (C.c_char_p, [_p]),
    "icedem_version": (C.c_int, []),
    "icedem_create": (C.c_int, [C.POINTER(IceDemStatic), C.POINTER(IceDemParams), C.c_int, C.POINTER(_p)]),
    "icedem_destroy": (C.c_int, [_p]),
    "icedem_create_like": (C.c_int, [_p, C.POINTER(_p)]),
    "icedem_frames": (C.c_int, [_p, C.POINTER(IceDemState), C.POINTER(IceDemInputs), _i,
                                C.POINTER(IceDemHistory), C.POINTER(IceDemDiag), _i, _p]),
    "icedem_frames_profiled": (C.c_int, [_p, C.POINTER(IceDemState), C.POINTER(IceDemInputs), _i,
                                         C.POINTER(IceDemHistory), C.POINTER(C.c_float), _p]),
    "icedem_batch_create": (C.c_int, [C.POINTER(_p), _i, C.POINTER(_p)]),
    "icedem_batch_rollouts": (C.c_int, [_p, C.POINTER(IceDemState), C.POINTER(IceDemInputs), _i,
                                        C.POINTER(IceDemDiag), _i, _p]),
    "icedem_batch_destroy": (C.c_int, [_p]),
    "icedem_batch_mode": (C.c_int, [_p, _i]),
    # surrogate kernels (icesur.h)
    "icesur_gemm": (C.c_int, [_i, _i, _i, _i, _i, _p, _i, _p, _i, _p, _i, _p, _i, _p, _i, _i, _p, _p]),
    "icesur_conv12": (C.c_int, [_p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p]),
    "icesur_im2col": (C.c_int, [_p, _i, _i, _i, _i, _p, _p]),
    "icesur_col2im": (C.c_int, [_p, _i, _i, _i, _i, _p, _p]),
    "icesur_relu_back": (C.c_int, [_p, _i, _p, _i, _i, _i, _p, _i, _p]),
    "icesur_colsum": (C.c_int, [_p, _i, _i, _i, _p, _i, _p, _p]),
    "icesur_to_bf16": (C.c_int, [_p, C.c_int64, _p, _p]),
    "icesur_adam": (C.c_int, [_p, _p, _p, _p, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, _i, _p, _p]),
    "icesur_ke_costmap": (C.c_int, [_p, _p, _p, _i, _i, _i, C.c_float, C.c_float, C.c_float, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_float, C.c_float, _i, _p, _p, _p]),
    "icedem_ship_rollouts": (C.c_int, [_p, _p, _i, _p, _i, _i, _i, _i, _p, _i, _i, C.c_float, C.c_float,
                                       C.c_float, C.c_float, _p, _i, _i, _p, _p]),
    "icesur_mppi_update": (C.c_int, [_p, _p, _i, _i, C.c_float, _p, _p, _p]),
    "icesur_mbd_update": (C.c_int, [_p, _p, _i, _i, C.c_float, _p, C.c_float, C.c_float, C.c_float, _p, _p, _p]),
    "icesur_conv": (C.c_int, [_p, _i, _i, _i, _i, _p, _i, _p, _i, _p, _p, _p]),
    "icesur_adamw": (C.c_int, [_p, _p, _p, _p, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                               _i, _p, _p]),
    "icesur_mse": (C.c_int, [_p, _p, _p, _i, _i, _p, _p, _p, _p]),
    "icesur_concat": (C.c_int, [_p, _i, _i, _p, _i, _i, _i, _p, _i, _p]),
    "icesur_split": (C.c_int, [_p, _i, _i, _i, _p, _p]),
    "icesur_raster": (C.c_int, [_p, _p, _p, _i, _p, _i, _i, C.c_float, C.c_float, _p, _p, _p]),
    "icesur_raster_poly": (C.c_int, [_p, _p, _p, _i, _i, _p, _i, _i, C.c_float, C.c_float, _p, _p, _p]),
    "icedem_sync": (C.c_int, [_p, _p]),
    "icedem_list_stats": (C.c_int, [_p, _p, C.POINTER(C.c_int32)]),
    "icedem_substep_forces": (C.c_int, [_p, C.POINTER(IceDemState), _p, _p, C.c_int64, C.c_int64, _p, _p, _p]),
    "icedem_contact_pairs": (C.c_int, [_p, _p, _p, _p, C.c_int64, C.POINTER(C.c_int64), _p]),
    "icedem_collisions_combined": (C.c_int, [_p, _p, _p, _p, _i, _i, _f, _f, _f, _f, _f, _p, _i, _i, _p, _p,
                                             _p, _p, _p, _p, _p, _p]),
    "icedem_bond_damage": (C.c_int, [_i, _p, _p, _p, _p, _p, _f, _f, _f, _f, _f, _p, _p, _p, _p, _p, _p, _p, _p,
                                     _p]),
    "icedem_connected_components": (C.c_int, [_i, _i, _p, _p, _p, _i, _p, _p]),
    "icedem_connected_components_noalloc": (C.c_int, [_i, _i, _p, _p, _p, _i, _p, _p]),
    "icedem_component_com_velocities": (C.c_int, [_i, _p, _p, _p, _p, _p]),
    "icedem_ship_step": (C.c_int, [_p, _p, _p, _p, _p, _p]),
}

_LIB = None


class IceDemError(RuntimeError):
    pass


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load the shared library (no compute; safe without a GPU)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise IceDemError(f"{p} is not built; run __graft_entry__.build() or make -C mbd-icenav_amd/csrc")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != 0:
        lib = load()
        msg = lib.icedem_last_error(ctx).decode() if ctx is not None else lib.icedem_last_error(None).decode()
        raise IceDemError(f"{what}: {lib.icedem_status_string(rc).decode()} ({rc}): {msg}")
