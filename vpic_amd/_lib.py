"""ctypes binding of libvpichip.so (the C-ABI of include/vpic_hip.h).

The library is the only compute path: when it is missing or fails to load
this module raises, there is no fallback.
"""
import ctypes
import os

from . import types as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpichip.so")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_pg = ctypes.POINTER(T.Grid)
_pmp = ctypes.POINTER(T.FieldParams)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf64 = ctypes.POINTER(ctypes.c_double)
_pdi = ctypes.POINTER(T.DumpInfo)

# name -> argtypes (all return int status)
_SIGS = {
    "vpic_hip_init": [ctypes.c_int],
    "vpic_hip_advance_p": [_vp, _i32, _f32, _f32, _vp, _i32, _pi32, _pi32,
                           _vp, _vp, _pg, _vp, _vp],
    "vpic_hip_advance_p_async": [_vp, _i32, _f32, _f32, _vp, _i32, _vp, _vp,
                                 _vp, _pg, _vp, _vp],
    "vpic_hip_advance_p_noalloc": [_vp, _i32, _f32, _f32, _vp, _i32, _vp,
                                   _vp, _vp, _pg, _vp],
    "vpic_hip_advance_p_keys": [_vp, _i32, _f32, _f32, _vp, _i32, _vp, _vp,
                                _vp, _pg, _vp, _vp, _vp],
    "vpic_hip_advance_p_species": [_i32, _vp, _vp, _vp, _pg, _vp],
    "vpic_hip_device_synchronize": [],
    "vpic_hip_push_tile": [_i32, _i32, _i32],
    "vpic_hip_push_tile_window": [_i32, _i32, _i32, _i32],
    "vpic_hip_push_get_tile": [_vp],
    "vpic_hip_push_direct_runs": [_i32],
    "vpic_hip_push_group": [_i32],
    "vpic_hip_push_bin_mode": [_i32],
    "vpic_hip_move_p": [_vp, _vp, _i32, _vp, _pg, _f32, _vp, _vp],
    "vpic_hip_sort_p": [_vp, _vp, _i32, _vp, _pg, ctypes.c_int, _vp],
    "vpic_hip_sort_p_keys": [_vp, _vp, _i32, _vp, _vp, _pg, ctypes.c_int,
                             _vp],
    "vpic_hip_sort_keys_stale": [_vp],
    "vpic_hip_takizuka_abe": [_vp, _vp, _i32, _f32, _f32, _vp, _vp, _i32,
                              _f32, _f32, ctypes.c_double, _i32,
                              ctypes.c_uint64, ctypes.c_uint64, _pg, _vp],
    "vpic_hip_energy_p": [_vp, _i32, _f32, _f32, _vp, _pg, _pf64, _vp],
    "vpic_hip_accumulate_rho_p": [_vp, _vp, _i32, _f32, _pg, _vp],
    "vpic_hip_center_p": [_vp, _i32, _f32, _f32, _vp, _pg, _i32, _vp],
    "vpic_hip_load_interpolator_array": [_vp, _vp, _pg, _vp],
    "vpic_hip_clear_accumulator_array": [_vp, _pg, _vp],
    "vpic_hip_reduce_accumulator_array": [_vp, _pg, _vp],
    "vpic_hip_unload_accumulator_array": [_vp, _vp, _pg, _vp],
    "vpic_hip_clear_jf": [_vp, _pg, _vp],
    "vpic_hip_clear_jf_unload_accumulator_array": [_vp, _vp, _pg, _vp],
    "vpic_hip_clear_rhof": [_vp, _pg, _vp],
    "vpic_hip_synchronize_jf": [_vp, _pg, _vp],
    "vpic_hip_synchronize_rho": [_vp, _pg, _vp],
    "vpic_hip_advance_b": [_vp, _pg, _f32, _vp],
    "vpic_hip_vacuum_advance_e": [_vp, _pg, _pmp, _f32, _vp],
    "vpic_hip_energy_f": [_pf64, _vp, _pg, _pmp, _vp],
    "vpic_hip_compute_div_e_err": [_vp, _pg, _pmp, _vp],
    "vpic_hip_compute_rms_div_e_err": [_pf64, _vp, _pg, _vp],
    "vpic_hip_clean_div_e": [_vp, _pg, _pmp, _vp],
    "vpic_hip_compute_div_b_err": [_vp, _pg, _vp],
    "vpic_hip_compute_rms_div_b_err": [_pf64, _vp, _pg, _vp],
    "vpic_hip_clean_div_b": [_vp, _pg, _vp],
    "vpic_hip_compute_rhob": [_vp, _pg, _pmp, _vp],
    "vpic_hip_compute_curl_b": [_vp, _pg, _pmp, _vp],
    "vpic_hip_synchronize_tang_e_norm_b": [_pf64, _vp, _pg, _vp],
    "vpic_hip_clear_hydro_array": [_vp, _pg, _vp],
    "vpic_hip_accumulate_hydro_p": [_vp, _vp, _i32, _f32, _f32, _vp, _pg, _vp,
                                    _vp],
    "vpic_hip_synchronize_hydro_array": [_vp, _pg, _vp],
    "vpic_hip_boundary_p": [_vp, _pi32, _vp, _i32, _vp, _f32, _pg, _vp,
                            _pi32, _vp],
    "vpic_hip_accumulate_rhob": [_vp, _vp, _i32, _f32, _pg, _vp],
    "vpic_hip_inject_p": [_vp, _pi32, _i32, _vp, _pi32, _i32, _vp, _i32, _vp,
                          _f32, _pg, _pi32, _vp],
    "vpic_hip_write_energies": [ctypes.c_char_p, ctypes.c_int, _pdi, _f32,
                                _vp, _vp, _vp, _i32],
    "vpic_hip_write_fields": [ctypes.c_char_p, _vp, _pg, _pdi],
    "vpic_hip_write_hydro": [ctypes.c_char_p, _vp, _pg, _pdi, _i32, _f32],
    "vpic_hip_write_particles": [ctypes.c_char_p, _vp, _i32, _i32, _pg, _pdi,
                                 _i32, _f32],
    "vpic_hip_write_grid": [ctypes.c_char_p, _pg, _pdi, _vp, _vp, _vp],
    "vpic_hip_global_header": [ctypes.c_char_p, _pg,
                               ctypes.POINTER(T.GlobalInfo),
                               ctypes.POINTER(T.DumpParams), _i32],
    "vpic_hip_field_dump": [_vp, _pg, _pdi, ctypes.POINTER(T.DumpParams),
                            ctypes.c_int64],
    "vpic_hip_hydro_dump": [_vp, _pg, _pdi, _i32, _f32,
                            ctypes.POINTER(T.DumpParams), ctypes.c_int64],
    "vpic_hip_grid_periodic": [_pg, _f64, _f64, _f64, _f64, _f64, _f64, _i32,
                               _i32, _i32, _f32, _f32, _f32],
    "vpic_hip_field_params_vacuum": [_pmp, _f32],
    "vpic_hip_grid_neighbor_table": [_pg, _vp],
    "vpic_hip_copy": [_vp, _vp, _i64, _vp],
    "vpic_hip_load_uniform_plasma": [_vp, _pg, _i32, _f32, _f32, _f32, _f32,
                                     _f32, _f32, _f32, ctypes.c_uint64, _vp],
    "vpic_hip_adjoint_bfield_energy": [_vp, _i32, _f32, _f32, _vp, _pg, _i32,
                                       _i32, _i32, _pf64, _vp, _vp, _vp,
                                       _vp],
    "vpic_hip_ensemble_create": [_vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "vpic_hip_ensemble_destroy": [_vp],
    "vpic_hip_ensemble_box": [_vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "vpic_hip_ensemble_prepare": [_vp],
    "vpic_hip_ensemble_step": [_vp, _i32, _i32],
    "vpic_hip_ensemble_shape": [_vp, _i32, _i32, _i32, _i32, _i32],
    "vpic_hip_ensemble_get_shape": [_vp, _vp],
    "vpic_hip_ensemble_tune": [_vp, _i32, _pf64],
    "vpic_hip_ensemble_sort_interval": [_vp, _i32],
    "vpic_hip_ensemble_tile": [_vp, _i32, _i32, _i32],
    "vpic_hip_ensemble_get_tile": [_vp, _vp],
    "vpic_hip_ensemble_tile_window": [_vp, _i32, _i32, _i32, _i32],
    "vpic_hip_ensemble_get_tile_window": [_vp, _vp],
    "vpic_hip_push_timing_enable": [ctypes.c_int],
    "vpic_hip_push_timing": [_pi64, _pf64, _pi64],
}

EXPORTS = sorted(_SIGS) + ["vpic_hip_last_error", "vpic_hip_version"]


class VpicError(RuntimeError):
    pass


_lib = None


def load():
    """Load libvpichip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VpicError("libvpichip.so is not built (%s); run "
                        "__graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.vpic_hip_last_error.restype = ctypes.c_char_p
    lib.vpic_hip_last_error.argtypes = []
    lib.vpic_hip_version.restype = ctypes.c_char_p
    lib.vpic_hip_version.argtypes = []
    _lib = lib
    return lib


def call(name, *args):
    """Call a C-ABI entry point; raise VpicError on a negative status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.vpic_hip_last_error().decode(errors="replace")
        raise VpicError("%s failed (%d): %s" % (name, rc, msg))
    return rc
