"""ctypes binding of libheddle_hip.so (include/heddle_hip.h).

The library is the product: this module only declares its C ABI.  There is
no Python fallback for any entry point -- if the shared library is missing
the import fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheddle_hip.so")

_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_fp = ctypes.POINTER(ctypes.c_float)
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_u64p = ctypes.POINTER(ctypes.c_uint64)

HMD_OK = 0
STATUS_NAMES = {1: "HMD_ERR_GPU", 2: "HMD_ERR_INVALID", 3: "HMD_ERR_NBL_OVERFLOW",
                4: "HMD_ERR_BOX_TOO_SMALL", 5: "HMD_ERR_UNSUPPORTED", 6: "HMD_ERR_NO_DEVICE",
                7: "HMD_ERR_OUTPUT_EXISTS", 8: "HMD_ERR_IO"}
HMD_ERR_NBL_OVERFLOW = 3
HMD_ERR_OUTPUT_EXISTS = 7
UNITS = {"si": 0, "atomic": 1}
# Dimension (src/units/mod.rs), in the order of include/heddle_hip.h
DIMENSIONS = ["dimensionless", "length", "inverse_length", "mass", "charge", "energy", "time",
              "inverse_time", "force", "pressure", "inverse_pressure", "temperature", "velocity",
              "stiffness"]

# Every symbol include/heddle_hip.h declares (checked by the CPU tests).
EXPORTED = [
    "hmd_last_error", "hmd_version", "hmd_ctx_create", "hmd_ctx_create_on_stream",
    "hmd_ctx_destroy",
    "hmd_ctx_synchronize", "hmd_ctx_device_info", "hmd_kernel_count", "hmd_kernel_name",
    "hmd_particles_create", "hmd_particles_destroy", "hmd_particles_download",
    "hmd_particles_upload", "hmd_particles_download_energies", "hmd_particles_download_lo",
    "hmd_particles_kinetic_energy", "hmd_particles_energy_sums",
    "hmd_particles_negate_velocities", "hmd_particles_upload_forces",
    "hmd_particles_upload_images", "hmd_particles_set_box", "hmd_particles_get_box",
    "hmd_vv_kick_drift", "hmd_vv_kick", "hmd_lan_drift_half", "hmd_lan_ou_step",
    "hmd_replicas_create", "hmd_replicas_destroy", "hmd_replicas_step", "hmd_replicas_steps",
    "hmd_replicas_plan",
    "hmd_integrator_timings", "hmd_timings_write",
    "hmd_forcefield_set_pair_fragment", "hmd_pair_fragment_compile",
    "hmd_forcefield_autotune", "hmd_forcefield_set_tuning", "hmd_forcefield_get_tuning",
    "hmd_settle_create", "hmd_settle_destroy", "hmd_settle_before_drift",
    "hmd_settle_after_drift", "hmd_settle_after_kick",
    "hmd_forcefield_create", "hmd_forcefield_destroy", "hmd_forcefield_compute",
    "hmd_forcefield_slots", "hmd_spme_scalars", "hmd_spme_grids", "hmd_spme_forces",
    "hmd_nbl_request_rebuild", "hmd_nbl_pair_snapshot", "hmd_nbl_pair_snapshot_noalloc",
    "hmd_nbl_count_pairs",
    "hmd_nbl_stats", "hmd_nbl_set_capacity",
    "hmd_nbl_pre_step", "hmd_pair_force", "hmd_spme_compute", "hmd_forcefield_check",
    "hmd_integrator_create", "hmd_integrator_destroy", "hmd_integrator_step",
    "hmd_integrator_capture", "hmd_integrator_replay", "hmd_integrator_steps",
    "hmd_integrator_profile", "hmd_integrator_minimize",
    "hmd_integrator_set_forcefield", "hmd_integrator_run_judged",
    "hmd_integrator_thermostat_state", "hmd_integrator_checkpoint", "hmd_integrator_rollback",
    "hmd_integrator_set_barostat", "hmd_integrator_barostat_state",
    "hmd_forcefield_add_bonded", "hmd_bonded_compute",
    "hmd_forcefield_set_fragment_exclusion_scales",
    "hmd_generate_velocities", "hmd_chacha_block",
    "hmd_mlp_param_count", "hmd_mlp_create", "hmd_mlp_destroy", "hmd_mlp_set_params",
    "hmd_mlp_get_params", "hmd_mlp_set_force_scale", "hmd_mlp_descriptors",
    "hmd_mlp_forward", "hmd_mlp_loss_grad", "hmd_mlp_adam_step", "hmd_mlp_attach",
    "hmd_mlp_dataset_reserve", "hmd_mlp_record_frame", "hmd_mlp_dataset_stats",
    "hmd_mlp_train", "hmd_mlp_judge",
    "hmd_traj_open", "hmd_traj_write_frame", "hmd_traj_flush", "hmd_traj_close",
    "hmd_log_open", "hmd_log_write_row", "hmd_log_flush", "hmd_log_close",
]


class ForceFieldDesc(ctypes.Structure):
    _fields_ = [
        ("n_types", ctypes.c_uint32),
        ("lj_on", ctypes.c_int32),
        ("lj_sigma", _c_dp), ("lj_epsilon", _c_dp), ("lj_cutoff", _c_dp), ("lj_switch", _c_dp),
        ("coulomb_on", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("r_cut_real", ctypes.c_double), ("k_coulomb", ctypes.c_double),
        ("spme_on", ctypes.c_int32),
        ("grid", ctypes.c_uint32 * 3), ("spline_order", ctypes.c_uint32),
        ("brick", ctypes.c_uint32 * 3),
        ("r_skin", ctypes.c_double),
        ("excl_offsets", _c_u32p), ("excl_partners", _c_u32p),
        ("excl_lj_scale", _c_fp), ("excl_coul_scale", _c_fp),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("pair_blocks_per_cu", ctypes.c_uint32),
        ("fft_blocks_per_cu", ctypes.c_uint32),
        ("brick", ctypes.c_uint32 * 3),
        ("spme_stream", ctypes.c_uint32),
        ("fft_slab", ctypes.c_uint32),
        ("n_candidates", ctypes.c_uint32),
        ("us_initial", ctypes.c_double),
        ("us_tuned", ctypes.c_double),
    ]


class IntegratorDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_char_p),
        ("dt", ctypes.c_double),
        ("compensated", ctypes.c_int32),
        ("temperature", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("n_groups", ctypes.c_uint32),
        ("group_atoms", _c_u32p),
        ("d_oh", ctypes.c_double), ("d_hh", ctypes.c_double),
        ("m_o", ctypes.c_double), ("m_h", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("thermostat", ctypes.c_uint32), ("tau", ctypes.c_double),
    ]


class BarostatDesc(ctypes.Structure):
    """hmd_barostat_desc (the c-rescale barostat's parameters)."""
    _fields_ = [("kind", ctypes.c_uint32), ("pressure", ctypes.c_double),
                ("temperature", ctypes.c_double), ("tau", ctypes.c_double),
                ("compressibility", ctypes.c_double), ("seed", ctypes.c_uint64)]


class BondedDesc(ctypes.Structure):
    """hmd_bonded_desc (one set of bonded terms)."""
    _fields_ = [("kind", ctypes.c_uint32), ("n_terms", ctypes.c_uint32),
                ("terms", _c_u32p), ("n_types", ctypes.c_uint32), ("params", _c_dp)]


class SlotOutput(ctypes.Structure):
    """hmd_slot_output: device pointers of a caller-owned SlotOutputView."""
    _fields_ = [("force_x", ctypes.c_void_p), ("force_y", ctypes.c_void_p),
                ("force_z", ctypes.c_void_p), ("energy", ctypes.c_void_p),
                ("virial", ctypes.c_void_p)]


class MlpDesc(ctypes.Structure):
    _fields_ = [
        ("r_cut", ctypes.c_double),
        ("force_scale", ctypes.c_double),
        ("n_types", ctypes.c_uint32),
        ("species_of_type", _c_u32p),
        ("seed", ctypes.c_uint64),
        ("lr", ctypes.c_double), ("beta1", ctypes.c_double),
        ("beta2", ctypes.c_double), ("eps", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libheddle_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m heddlemd_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp = ctypes.c_void_p
    pvp = ctypes.POINTER(ctypes.c_void_p)
    st = ctypes.c_int32
    sig = {
        "hmd_last_error": (ctypes.c_char_p, []),
        "hmd_version": (ctypes.c_char_p, []),
        "hmd_ctx_create": (st, [ctypes.c_int, pvp]),
        "hmd_ctx_destroy": (None, [vp]),
        "hmd_ctx_synchronize": (st, [vp]),
        "hmd_ctx_device_info": (st, [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]),
        "hmd_kernel_count": (ctypes.c_int, []),
        "hmd_kernel_name": (ctypes.c_char_p, [ctypes.c_int]),
        "hmd_particles_create": (st, [vp, ctypes.c_uint32, _c_fp, _c_fp, _c_fp, _c_u32p, _c_fp, pvp]),
        "hmd_particles_destroy": (None, [vp]),
        "hmd_particles_download": (st, [vp, _c_fp, _c_fp, _c_fp, _c_i32p]),
        "hmd_particles_upload": (st, [vp, _c_fp, _c_fp]),
        "hmd_particles_download_energies": (st, [vp, _c_fp, _c_fp]),
        "hmd_particles_download_lo": (st, [vp, _c_dp, _c_dp]),
        "hmd_particles_kinetic_energy": (st, [vp, _c_dp]),
        "hmd_particles_energy_sums": (st, [vp, _c_dp, _c_dp]),
        "hmd_particles_negate_velocities": (st, [vp]),
        "hmd_particles_upload_forces": (st, [vp, _c_fp]),
        "hmd_particles_upload_images": (st, [vp, _c_i32p]),
        "hmd_particles_set_box": (st, [vp, _c_fp]),
        "hmd_particles_get_box": (st, [vp, _c_fp]),
        "hmd_vv_kick_drift": (st, [vp, ctypes.c_float, ctypes.c_int32]),
        "hmd_vv_kick": (st, [vp, ctypes.c_float, ctypes.c_int32]),
        "hmd_lan_drift_half": (st, [vp, ctypes.c_float]),
        "hmd_replicas_create": (st, [ctypes.POINTER(vp), ctypes.c_uint32, ctypes.c_uint32, pvp]),
        "hmd_settle_create": (st, [vp, ctypes.c_uint32, _c_u32p] + [ctypes.c_double] * 4
                              + [pvp]),
        "hmd_settle_destroy": (None, [vp]),
        "hmd_forcefield_autotune": (st, [vp, ctypes.c_uint32, ctypes.POINTER(Tuning)]),
        "hmd_forcefield_set_tuning": (st, [vp, ctypes.POINTER(Tuning)]),
        "hmd_forcefield_get_tuning": (st, [vp, ctypes.POINTER(Tuning)]),
        "hmd_forcefield_set_pair_fragment": (st, [vp, ctypes.c_char_p, ctypes.c_char_p, _c_fp,
                                                  ctypes.c_uint32]),
        "hmd_pair_fragment_compile": (st, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_int32]),
        "hmd_settle_before_drift": (st, [vp]),
        "hmd_settle_after_drift": (st, [vp, ctypes.c_float]),
        "hmd_settle_after_kick": (st, [vp, ctypes.c_float]),
        "hmd_integrator_timings": (st, [vp, ctypes.c_uint32, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p,
                                        ctypes.c_int32] + [ctypes.POINTER(ctypes.c_uint64)] * 4),
        "hmd_timings_write": (st, [ctypes.c_char_p, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_char_p)]
                              + [ctypes.POINTER(ctypes.c_uint64)] * 4),
        "hmd_replicas_destroy": (None, [vp]),
        "hmd_replicas_step": (st, [vp, ctypes.c_uint32]),
        "hmd_replicas_steps": (ctypes.c_uint64, [vp]),
        "hmd_replicas_plan": (st, [vp, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.POINTER(ctypes.c_uint32)]),
        "hmd_lan_ou_step": (st, [vp, ctypes.c_uint64, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_uint64]),
        "hmd_forcefield_create": (st, [vp, vp, ctypes.POINTER(ForceFieldDesc), pvp]),
        "hmd_forcefield_destroy": (None, [vp]),
        "hmd_forcefield_compute": (st, [vp, ctypes.c_int32]),
        "hmd_forcefield_slots": (st, [vp, ctypes.c_char_p, ctypes.c_int32]),
        "hmd_spme_scalars": (st, [vp, _c_dp]),
        "hmd_spme_grids": (st, [vp, _c_fp, _c_fp]),
        "hmd_spme_forces": (st, [vp, _c_fp]),
        "hmd_nbl_request_rebuild": (st, [vp]),
        "hmd_nbl_pair_snapshot": (st, [vp, _c_u32p, ctypes.c_uint64, _c_u64p]),
        "hmd_nbl_pair_snapshot_noalloc": (st, [vp, _c_u32p, ctypes.c_uint64, _c_u64p]),
        "hmd_nbl_count_pairs": (st, [vp, ctypes.c_double, _c_u64p]),
        "hmd_nbl_stats": (st, [vp, _c_u32p]),
        "hmd_nbl_set_capacity": (st, [vp, ctypes.c_uint32, ctypes.c_uint32]),
        "hmd_nbl_pre_step": (st, [vp]),
        "hmd_pair_force": (st, [vp, ctypes.c_int32, ctypes.POINTER(SlotOutput)]),
        "hmd_spme_compute": (st, [vp, ctypes.c_int32, ctypes.POINTER(SlotOutput)]),
        "hmd_forcefield_check": (st, [vp]),
        "hmd_ctx_create_on_stream": (st, [ctypes.c_int, vp, pvp]),
        "hmd_integrator_create": (st, [vp, vp, vp, ctypes.POINTER(IntegratorDesc), pvp]),
        "hmd_integrator_destroy": (None, [vp]),
        "hmd_integrator_step": (st, [vp, ctypes.c_uint32]),
        "hmd_integrator_capture": (st, [vp]),
        "hmd_integrator_replay": (st, [vp, ctypes.c_uint32]),
        "hmd_integrator_steps": (ctypes.c_uint64, [vp]),
        "hmd_integrator_set_forcefield": (st, [vp, vp]),
        "hmd_integrator_thermostat_state": (st, [vp, _c_dp]),
        "hmd_integrator_set_barostat": (st, [vp, ctypes.POINTER(BarostatDesc)]),
        "hmd_forcefield_add_bonded": (st, [vp, ctypes.POINTER(BondedDesc)]),
        "hmd_forcefield_set_fragment_exclusion_scales": (st, [vp, _c_fp, ctypes.c_uint32]),
        "hmd_bonded_compute": (st, [vp, ctypes.c_int32, ctypes.POINTER(SlotOutput)]),
        "hmd_integrator_barostat_state": (st, [vp, _c_dp]),
        "hmd_integrator_checkpoint": (st, [vp]),
        "hmd_integrator_rollback": (st, [vp]),
        "hmd_integrator_run_judged": (st, [vp, vp, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_double, _c_dp, ctypes.c_uint32,
                                           _c_u32p, _c_u64p]),
        "hmd_integrator_profile": (st, [vp, ctypes.c_uint32, _c_dp, ctypes.c_int32,
                                        _c_i32p, ctypes.c_char_p, ctypes.c_int32]),
        "hmd_integrator_minimize": (st, [vp, ctypes.c_uint32, ctypes.c_double]),
        "hmd_generate_velocities": (st, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.c_uint64, _c_dp, _c_fp]),
        "hmd_chacha_block": (None, [_c_u32p, ctypes.c_uint32, _c_u32p]),
        "hmd_mlp_param_count": (ctypes.c_uint32, []),
        "hmd_mlp_create": (st, [vp, ctypes.POINTER(MlpDesc), pvp]),
        "hmd_mlp_destroy": (None, [vp]),
        "hmd_mlp_set_params": (st, [vp, _c_fp]),
        "hmd_mlp_get_params": (st, [vp, _c_fp]),
        "hmd_mlp_set_force_scale": (st, [vp, ctypes.c_double]),
        "hmd_mlp_descriptors": (st, [vp, vp, _c_fp]),
        "hmd_mlp_forward": (st, [vp, vp, _c_fp]),
        "hmd_mlp_loss_grad": (st, [vp, vp, _c_fp, _c_dp, _c_fp]),
        "hmd_mlp_adam_step": (st, [vp, _c_fp]),
        "hmd_mlp_attach": (st, [vp, vp]),
        "hmd_mlp_dataset_reserve": (st, [vp, ctypes.c_uint32, ctypes.c_uint32]),
        "hmd_mlp_record_frame": (st, [vp, vp, vp]),
        "hmd_mlp_dataset_stats": (st, [vp, _c_dp]),
        "hmd_mlp_train": (st, [vp, vp, ctypes.c_uint32, ctypes.c_uint32, _c_dp]),
        "hmd_mlp_judge": (st, [vp, vp, _c_dp]),
        "hmd_traj_open": (st, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint32, pvp]),
        "hmd_traj_write_frame": (st, [vp, ctypes.c_uint64, ctypes.c_double, _c_fp, ctypes.c_uint32,
                                      _c_u32p, _c_fp, _c_fp, _c_fp, _c_fp, _c_fp, _c_fp,
                                      _c_i32p, _c_i32p, _c_i32p]),
        "hmd_traj_flush": (st, [vp]),
        "hmd_traj_close": (None, [vp]),
        "hmd_log_open": (st, [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_char_p),
                              _c_u32p, ctypes.c_uint32, pvp]),
        "hmd_log_write_row": (st, [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _c_dp]),
        "hmd_log_flush": (st, [vp]),
        "hmd_log_close": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


class HmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def check(code):
    if code != HMD_OK:
        raise HmdError(code, lib().hmd_last_error().decode())
