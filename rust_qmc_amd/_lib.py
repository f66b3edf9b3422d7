"""ctypes binding of include/qmc.h (rust_qmc_amd/libqmc_hip.so).

The library must have been built (rust_qmc_amd/build.py or
__graft_entry__.build()); there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmc_hip.so")

QMC_OK = 0
QMC_ERR_INVALID = -1
QMC_ERR_UNSUPPORTED = -2
QMC_ERR_HIP = -3
QMC_ERR_SINGULAR = -4
QMC_ERR_NOMEM = -5

QMC_FP64 = 0
QMC_MIXED = 1

QMC_FLAG_J2_GRAD_QUIRK = 1 << 0
QMC_FLAG_J2_LAP_QUIRK = 1 << 1
QMC_FLAG_HF_NN_REFSIGN = 1 << 2
QMC_FLAG_SINGULAR_IDENTITY = 1 << 3
QMC_FLAGS_REFERENCE = 0xF

QMC_AO_GTO, QMC_AO_STO, QMC_AO_SLATER1S = 0, 1, 2
QMC_JAS_NONE, QMC_JAS_PAIR, QMC_JAS_PADE = 0, 1, 2
QMC_BASIS_631G, QMC_BASIS_STO3G = 0, 1
QMC_MOVE_SINGLE, QMC_MOVE_ALL, QMC_MOVE_DRIFT = 0, 1, 2
QMC_BATCH_FUSED, QMC_BATCH_STREAMS, QMC_BATCH_SEQUENTIAL = 0, 1, 2
QMC_BATCH_ORDER_COST, QMC_BATCH_ORDER_GIVEN = 0, 1

EXPORTED = [
    "qmc_system_create", "qmc_system_destroy", "qmc_h2_molecule_mo_create", "qmc_sto_slater_det_create",
    "qmc_lithium_seeded_create", "qmc_methane_gto_create", "qmc_benzene_gto_create", "qmc_set_precision",
    "qmc_set_engine", "qmc_engine_name",
    "qmc_set_flags", "qmc_system_info", "qmc_get_mo", "qmc_nuclear_repulsion", "qmc_num_params",
    "qmc_get_params", "qmc_set_params", "qmc_set_jastrow_series", "qmc_get_nuclei", "qmc_set_nuclei",
    "qmc_eval_log_psi", "qmc_grad_log_psi", "qmc_lap_over_psi", "qmc_local_energy", "qmc_evaluate",
    "qmc_ensemble_create", "qmc_ensemble_destroy", "qmc_ensemble_get_positions", "qmc_ensemble_set_positions",
    "qmc_ensemble_get_state", "qmc_ensemble_refresh", "qmc_ensemble_sweeps", "qmc_ensemble_set_timing",
    "qmc_ensemble_timing", "qmc_vmc_run", "qmc_derivatives", "qmc_log_derivs", "qmc_dEL_dp", "qmc_hf_force",
    "qmc_nuc_grad_log", "qmc_derivatives_noalloc", "qmc_sr_solve", "qmc_ensemble_measure", "qmc_ensemble_moments", "qmc_sr_optimize",
    "qmc_sample_forces", "qmc_geometry_optimize",
    "qmc_ddvmc_run", "qmc_isdmc_run", "qmc_isdmc_block_error", "qmc_system_from_json", "qmc_system_to_json",
    "qmc_batch_create", "qmc_batch_destroy", "qmc_batch_fusable", "qmc_batch_set_config", "qmc_batch_get_config",
    "qmc_batch_sweeps", "qmc_batch_timing", "qmc_batch_tune",
    "qmc_autocorrelation_time", "qmc_blocking_error", "qmc_last_error", "qmc_version",
]

_d = C.c_double
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_u64 = C.c_uint64
_p = C.c_void_p
_pp = C.POINTER(C.c_void_p)


class QmcAO(C.Structure):
    _fields_ = [("kind", _i32), ("center", _i32), ("l", _i32), ("m", _i32), ("nprim", _i32), ("prim0", _i32)]


class QmcSystemDesc(C.Structure):
    _fields_ = [
        ("n_elec", _i32), ("spins", _p), ("n_nuc", _i32), ("nuc_pos", _p), ("nuc_charge", _p),
        ("n_ao", _i32), ("aos", _p), ("prim_exp", _p), ("prim_coef", _p), ("prim_pow", _p),
        ("n_mo", _i32), ("mo_up", _p), ("mo_dn", _p), ("jastrow_kind", _i32), ("jas_F", _d),
        ("b_ee", _d), ("b_en", _d), ("a_anti", _d), ("a_para", _d), ("k_ee", _i32), ("k_en", _i32),
        ("c_anti", _p), ("c_para", _p), ("n_species", _i32), ("species", _p), ("d_en", _p),
        ("init_nuc", _p), ("init_sigma", _d), ("flags", _u32), ("mo_rule", _i32),
    ]


class QmcSweepParams(C.Structure):
    _fields_ = [("move", _i32), ("n_steps", _i32), ("step_size", _d), ("refresh_interval", _i32),
                ("steps_per_launch", _i32), ("max_drift", _d)]


class QmcMcmcParams(C.Structure):
    _fields_ = [("n_walkers", _i64), ("n_steps", _i64), ("initial_step_size", _d), ("max_step_size", _d),
                ("min_step_size", _d), ("target_acceptance", _d), ("adaptation_interval", _i64),
                ("n_burnin", _i64), ("move", _i32), ("refresh_interval", _i32)]


class QmcMoments(C.Structure):
    _fields_ = [("n_samples", _i64), ("P", _i32), ("M", _i32), ("energy", _d), ("variance", _d),
                ("o_mean", _p), ("S", _p), ("f", _p), ("de_mean", _p), ("f_hf", _p), ("f_zv", _p),
                ("f_var", _p)]


class QmcSrParams(C.Structure):
    _fields_ = [("n_samples", _i64), ("n_walkers", _i64), ("n_equilibrate", _i64), ("max_iterations", _i64),
                ("learning_rate", _d), ("sr_epsilon", _d), ("tolerance", _d), ("step_size", _d),
                ("decorrelation", _i32), ("move", _i32), ("include_nuclear", _i32), ("refresh_interval", _i32)]


class QmcSrResults(C.Structure):
    _fields_ = [("final_energy", _d), ("final_variance", _d), ("iterations", _i32)]


class QmcForceParams(C.Structure):
    _fields_ = [("n_samples", _i64), ("n_walkers", _i64), ("n_equilibrate", _i64), ("mc_step_size", _d),
                ("zero_variance", _i32), ("decorrelation", _i32), ("move", _i32), ("refresh_interval", _i32)]


class QmcBatchConfig(C.Structure):
    _fields_ = [("mode", _i32), ("blocks_per_cu", _i32), ("steps_per_launch", _i32), ("order", _i32),
                ("refresh_interval", _i32)]


class QmcDdvmcParams(C.Structure):
    _fields_ = [("n_walkers", _i64), ("n_steps", _i64), ("n_burnin", _i64), ("time_step", _d), ("max_drift", _d),
                ("target_acceptance", _d), ("adaptation_interval", _i64), ("refresh_interval", _i32)]


class QmcDdvmcResults(C.Structure):
    _fields_ = [("energy", _d), ("error", _d), ("autocorrelation_time", _d), ("acceptance_rate", _d),
                ("final_time_step", _d)]


class QmcGeomoptParams(C.Structure):
    _fields_ = [("step_size", _d), ("max_iterations", _i64), ("force_threshold", _d), ("max_displacement", _d)]


class QmcGeomoptResults(C.Structure):
    _fields_ = [("final_energy", _d), ("final_max_force", _d), ("iterations", _i64), ("converged", _i32)]


class QmcIsdmcParams(C.Structure):
    _fields_ = [("n_walkers", _i64), ("n_steps", _i64), ("n_burnin", _i64), ("time_step", _d), ("max_drift", _d),
                ("e_ref_damping", _d), ("n_vmc_equilibrate", _i64), ("n_vmc_force_samples", _i64),
                ("capacity", _i64), ("refresh_interval", _i32), ("matched_branching", _i32)]


class QmcIsdmcResults(C.Structure):
    _fields_ = [("energy", _d), ("error", _d), ("acceptance_rate", _d), ("avg_population", _d),
                ("collapses", _i64), ("dropped", _i64), ("capacity", _i64), ("step_ms", _d)]


class QmcVmcResults(C.Structure):
    _fields_ = [("energy", _d), ("error", _d), ("autocorrelation_time", _d), ("acceptance", _d),
                ("final_step_size", _d)]


_SIGS = {
    "qmc_system_create": (_i32, [_p, _i32, _pp]),
    "qmc_system_destroy": (_i32, [_p]),
    "qmc_h2_molecule_mo_create": (_i32, [_d, _p, _p, _d, _d, _i32, _pp]),
    "qmc_sto_slater_det_create": (_i32, [_i32, _p, _p, _p, _p, _p, _d, _d, _u32, _i32, _pp]),
    "qmc_lithium_seeded_create": (_i32, [_u64, _u32, _i32, _pp]),
    "qmc_methane_gto_create": (_i32, [_d, _d, _d, _i32, _u64, _i32, _pp]),
    "qmc_benzene_gto_create": (_i32, [_d, _d, _d, _d, _i32, _pp]),
    "qmc_set_precision": (_i32, [_p, _i32]),
    "qmc_set_engine": (_i32, [_p, _i32]),
    "qmc_engine_name": (_i32, [_p, C.c_char_p, _i32]),
    "qmc_set_flags": (_i32, [_p, _u32]),
    "qmc_system_info": (_i32, [_p, _p, _p, _p, _p]),
    "qmc_get_mo": (_i32, [_p, _i32, _p]),
    "qmc_nuclear_repulsion": (_i32, [_p, _p]),
    "qmc_num_params": (_i32, [_p, _p]),
    "qmc_get_params": (_i32, [_p, _p]),
    "qmc_set_params": (_i32, [_p, _p, _i32]),
    "qmc_set_jastrow_series": (_i32, [_p, _i32, _i32]),
    "qmc_get_nuclei": (_i32, [_p, _p, _p]),
    "qmc_set_nuclei": (_i32, [_p, _p, _i32]),
    "qmc_eval_log_psi": (_i32, [_p, _i64, _p, _p, _p, _p]),
    "qmc_grad_log_psi": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_lap_over_psi": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_local_energy": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_evaluate": (_i32, [_p, _i64, _p, _p, _p, _p, _p, _p, _p]),
    "qmc_ensemble_create": (_i32, [_p, _i64, _u64, _pp]),
    "qmc_ensemble_destroy": (_i32, [_p]),
    "qmc_ensemble_get_positions": (_i32, [_p, _p]),
    "qmc_ensemble_set_positions": (_i32, [_p, _p]),
    "qmc_ensemble_get_state": (_i32, [_p, _p, _p, _p]),
    "qmc_ensemble_refresh": (_i32, [_p]),
    "qmc_ensemble_sweeps": (_i32, [_p, _p, _p, _p, _p, _p]),
    "qmc_ensemble_set_timing": (_i32, [_p, _i32]),
    "qmc_ensemble_timing": (_i32, [_p, _p, _p]),
    "qmc_vmc_run": (_i32, [_p, _p, _u64, _p, _p]),
    "qmc_derivatives": (_i32, [_p, _i64, _p, _p, _p, _p, _p, _p, _p]),
    "qmc_log_derivs": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_dEL_dp": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_hf_force": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_nuc_grad_log": (_i32, [_p, _i64, _p, _p, _p]),
    "qmc_derivatives_noalloc": (_i32, [_p, _i64, _p, _p, _p, _p, _p, _p, _p]),
    "qmc_sr_solve": (_i32, [_i32, _p, _p, _d, _p]),
    "qmc_ensemble_measure": (_i32, [_p, _i32]),
    "qmc_ensemble_moments": (_i32, [_p, _p]),
    "qmc_sr_optimize": (_i32, [_p, _p, _u64, _p, _p, _p, _p]),
    "qmc_sample_forces": (_i32, [_p, _p, _u64, _p, _p, _p, _p]),
    "qmc_geometry_optimize": (_i32, [_p, _p, _p, _u64, _p, _p, _p, _p]),
    "qmc_ddvmc_run": (_i32, [_p, _p, _u64, _p, _p]),
    "qmc_isdmc_run": (_i32, [_p, _p, _u64, _p, _p, _p, _p, _p, _p]),
    "qmc_isdmc_block_error": (_d, [_p, _i64]),
    "qmc_system_from_json": (_i32, [C.c_char_p, _i64, _i32, _pp]),
    "qmc_system_to_json": (_i32, [_p, _p, _i64, _p]),
    "qmc_batch_create": (_i32, [_p, _i32, _pp]),
    "qmc_batch_destroy": (_i32, [_p]),
    "qmc_batch_fusable": (_i32, [_p, _p]),
    "qmc_batch_set_config": (_i32, [_p, _p]),
    "qmc_batch_get_config": (_i32, [_p, _p]),
    "qmc_batch_sweeps": (_i32, [_p, _i32, _p, _i32, _p, _p, _p]),
    "qmc_batch_timing": (_i32, [_p, _p, _p]),
    "qmc_batch_tune": (_i32, [_p, _i32, _p, _i32, _p, _p]),
    "qmc_autocorrelation_time": (_d, [_p, _i64]),
    "qmc_blocking_error": (_d, [_p, _i64, _d]),
    "qmc_last_error": (C.c_char_p, []),
    "qmc_version": (C.c_char_p, []),
}

_lib = None


class QmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qmc error {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with rust_qmc_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != QMC_OK:
        raise QmcError(rc, lib().qmc_last_error().decode())
    return rc
