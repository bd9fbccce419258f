"""ctypes binding of libumt_amd.so (declarations in include/umt_sweep.h and
include/umt_rz.h).

The library is built in-tree (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C umt_amd/csrc``). Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libumt_amd.so")

_i = C.c_int
_i64 = C.c_int64
_u64 = C.c_uint64
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_vp = C.c_void_p

# name -> (restype, argtypes); include/umt_sweep.h and include/umt_rz.h
SIGNATURES = {
    "umt_last_error": (C.c_char_p, []),
    "gpu_sweepucbxyz": (None, [_pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pi, _pd, _pd, _pi, _pi,
                               _pi, _pi, _pd, _pd, _pi, _pd, _pi, _pd, _pi, _pi, _pd, _pd, _pd,
                               _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pi, _pi, _pi, _pd, _pi,
                               _pi, _pi]),
    "gpu_sweepucbxyz_noalloc": (None, [_pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pi, _pd, _pd,
                                       _pi, _pi, _pi, _pi, _pd, _pd, _pi, _pd, _pi, _pd, _pi, _pi,
                                       _pd, _pd, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pi,
                                       _pi, _pi, _pd, _pi, _pi, _pi]),
    "gpu_sweepucbxyz_noalloc_max_bytes": (C.c_size_t, []),
    "gpu_streamsynchronize": (None, [_pi]),
    "gpu_devicesynchronize": (None, []),
    "gpu_sweepucbxyz_status": (_i, []),
    "umt_problem_tiled": (_i, [_i, _i, _i, _i, _i, _i, C.POINTER(_vp)]),
    "umt_problem_destroy": (None, [_vp]),
    "umt_problem_from_hex_mesh": (_i, [_i, _pd, _i, _pi, _i, _i, _i, C.POINTER(_vp)]),
    "umt_problem_dims": (_i, [_vp, _pi64]),
    "umt_problem_mesh": (_i, [_vp, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd]),
    "umt_problem_geometry": (_i, [_vp, _pd, _pd, _pd, _pd]),
    "umt_problem_quadrature": (_i, [_vp, _pd, _pd]),
    "umt_problem_schedule": (_i, [_vp, _pi, _pi, _pi, _pi]),
    "umt_problem_cycles": (_i, [_vp, _pi, _pi, _pi]),
    "umt_draw_cross_sections": (_i, [_u64, _i, _i, _d, _d, _d, _d, _d, _d, _pd, _pd, _pd]),
    "umt_ctx_create": (_i, [_i, _i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd,
                            _pd, _pd, _pi, _pi, _pi, C.POINTER(_vp)]),
    "umt_ctx_create_from_problem": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "umt_ctx_destroy": (None, [_vp]),
    "umt_ctx_set_sigt": (_i, [_vp, _pd]),
    "umt_ctx_set_source": (_i, [_vp, _pd]),
    "umt_ctx_set_scattering": (_i, [_vp, _pd, _pd]),
    "umt_ctx_sweep": (_i, [_vp]),
    "umt_ctx_get_phi": (_i, [_vp, _pd]),
    "umt_ctx_get_psi": (_i, [_vp, _pd]),
    "umt_ctx_set_phi": (_i, [_vp, _pd]),
    "umt_ctx_source_iteration": (_i, [_vp, _d, _i, _pi, _pd, _pd]),
    "umt_ctx_time_sweeps": (_i, [_vp, _i, _i, _pd, _pd, _pi]),
    "umt_ctx_dims": (_i, [_vp, _pi64]),
    "umt_ctx_sweep_n": (_i, [_vp, _i]),
    "umt_ctx_incoming_faces": (_i, [_vp, _pd]),
    "umt_ctx_level_stats": (_i, [_vp, C.POINTER(C.c_int64)]),
    "umt_device_memory_info": (_i, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "umt_hbm_copy_probe": (_i, [_i64, _i, _pd]),
    "umt_ctx_set_groups_per_lane": (_i, [_vp, _i]),
    "umt_ctx_get_groups_per_lane": (_i, [_vp, _pi]),
    "umt_ctx_autotune": (_i, [_vp, _i, _pi, _pd]),
    "umt_ctx_autotune_layouts": (_i, [_vp, _i, _pi, _pd]),
    "umt_problem_gta": (_i, [_vp, _pi, _pd, _pd, _pi, _pi, _pi]),
    "umt_ctx_enable_gta": (_i, [_vp, _i, _pd, _pd, _pi, _pi, _pi]),
    "umt_ctx_set_acceleration": (_i, [_vp, _i]),
    "umt_ctx_set_graphs": (_i, [_vp, _i]),
    "umt_ctx_set_angles_per_block": (_i, [_vp, _i]),
    "umt_ctx_get_angles_per_block": (_i, [_vp, _pi, _pi]),
    "umt_ctx_copy_phi": (_i, [_vp, _vp]),
    "umt_ctx_set_cycles": (_i, [_vp, _pi, _pi]),
    "umt_ctx_set_stall_window": (_i, [_vp, _i]),
    "umt_ctx_set_tau": (_i, [_vp, _d]),
    "umt_ctx_set_psi_prev": (_i, [_vp, _pd]),
    "umt_ctx_get_psi_prev": (_i, [_vp, _pd]),
    "umt_ctx_save_psi": (_i, [_vp]),
    "umt_ctx_gta_stats": (_i, [_vp, _pi64, _pi64]),
    "umt_ctx_grey_sweep": (_i, [_vp, _pd, _pd, _pd]),
    "umt_ctx_autotune_grey": (_i, [_vp, _i, _pi, _pd]),
    "umt_ctx_autotune_grey_layouts": (_i, [_vp, _i, _pi, _pd]),
    "umt_ctx_grey_solve": (_i, [_vp, _pd, _pd, _d, _pd, _pi64]),
    "umt_ctx_set_gta_solver": (_i, [_vp, _i]),
    "umt_ctx_set_grey_controls": (_i, [_vp, _d, _i]),
    "umt_ctx_get_grey_controls": (_i, [_vp, _pd, _pi, _pi]),
    "umt_ctx_set_gta_options": (_i, [_vp, _i]),
    "umt_ctx_set_incident_flux": (_i, [_vp, _pd]),
    "umt_ctx_get_boundary_flux": (_i, [_vp, _pd]),
    "umt_ctx_grey_transfer": (_i, [_vp, _pd, _pd]),
    "umt_batch_create": (_i, [_i, C.POINTER(_vp), _pi, _i, C.POINTER(_vp)]),
    "umt_batch_destroy": (None, [_vp]),
    "umt_batch_dims": (_i, [_vp, _pi64]),
    "umt_batch_set_cross_sections": (_i, [_vp, _i, _pd, _pd, _pd]),
    "umt_batch_reset": (_i, [_vp]),
    "umt_batch_sweep": (_i, [_vp]),
    "umt_batch_get_phi": (_i, [_vp, _i, _pd]),
    "umt_batch_source_iteration": (_i, [_vp, _d, _i, _pi, _pd, _pi]),
    "umt_batch_set_graphs": (_i, [_vp, _i]),
    "umt_batch_set_layout": (_i, [_vp, _i, _i]),
    "umt_batch_time_iterations": (_i, [_vp, _i, _i, _pd]),
    "umt_batch_autotune": (_i, [_vp, _i, _pi, _pi, _pd]),
    "umt_batch_set_persistent": (_i, [_vp, _i, _i]),
    "umt_batch_get_persistent": (_i, [_vp, _pi, _pi]),
    "umt_learner_num_params": (_i, []),
    "umt_learner_create": (_i, [_u64, C.POINTER(_vp)]),
    "umt_learner_destroy": (None, [_vp]),
    "umt_learner_get_params": (_i, [_vp, _pf]),
    "umt_learner_set_params": (_i, [_vp, _pf]),
    "umt_learner_forward": (_i, [_vp, _pf, _i64, _pf]),
    "umt_learner_gradient": (_i, [_vp, _pf, _pf, _pf, _i64, _pf, _pd]),
    "umt_learner_step": (_i, [_vp, _pf, _pf, _pf, _i64, _d]),
    "umt_learner_add_samples": (_i, [_vp, _pf, _pf, _pf, _i64]),
    "umt_learner_num_samples": (_i, [_vp, _pi64]),
    "umt_learner_get_samples": (_i, [_vp, _pf, _pf, _pf]),
    "umt_learner_clear_samples": (_i, [_vp]),
    "umt_learner_train": (_i, [_vp, _i, _i, _d, _u64, _pd]),
    "umt_ctx_set_zone_neighbors": (_i, [_vp, _pi]),
    "umt_ctx_set_learner": (_i, [_vp, _vp, _i, _i]),
    "umt_ctx_learn_collect": (_i, [_vp, _i, _i]),
    "umt_ctx_learn_commit": (_i, [_vp, _vp, _d, _pi]),
    "umt_ctx_learn_discard": (_i, [_vp]),
    "umt_ctx_learn_features": (_i, [_vp, _pd, _pd, _pd, _pf]),
    "umt_ctx_learn_stats": (_i, [_vp, _pi64, _pi]),
    # RZ (include/umt_rz.h)
    "umt_rz_problem_structured": (_i, [_i, _i, _d, _d, _d, _d, _d, _u64, _i, _i, C.POINTER(_vp)]),
    "umt_rz_problem_from_quad_mesh": (_i, [_i, _pd, _i, _pi, _i, _i, C.POINTER(_vp)]),
    "umt_rz_problem_destroy": (None, [_vp]),
    "umt_rz_problem_dims": (_i, [_vp, _pi64]),
    "umt_rz_problem_mesh": (_i, [_vp, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd]),
    "umt_rz_problem_points": (_i, [_vp, _pd, _pi]),
    "umt_rz_problem_geometry": (_i, [_vp, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "umt_rz_problem_quadrature": (_i, [_vp, _pd, _pd, _pi, _pi, _pd, _pd, _pd]),
    "umt_rz_problem_schedule": (_i, [_vp, _pi, _pi, _pi, _pi]),
    "umt_rz_ctx_create": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "umt_rz_ctx_destroy": (None, [_vp]),
    "umt_rz_ctx_dims": (_i, [_vp, _pi64]),
    "umt_rz_ctx_set_sigt": (_i, [_vp, _pd]),
    "umt_rz_ctx_set_source": (_i, [_vp, _pd]),
    "umt_rz_ctx_set_scattering": (_i, [_vp, _pd, _pd]),
    "umt_rz_ctx_sweep": (_i, [_vp]),
    "umt_rz_ctx_get_phi": (_i, [_vp, _pd]),
    "umt_rz_ctx_set_phi": (_i, [_vp, _pd]),
    "umt_rz_ctx_get_psi": (_i, [_vp, _pd]),
    "umt_rz_ctx_source_iteration": (_i, [_vp, _d, _i, _pi, _pd, _pd]),
    "umt_rz_ctx_set_groups_per_lane": (_i, [_vp, _i]),
    "umt_rz_ctx_time_sweeps": (_i, [_vp, _i, _i, _pd, _pd]),
    "umt_rz_ctx_set_graphs": (_i, [_vp, _i]),
    "umt_rz_ctx_set_tau": (_i, [_vp, _d]),
    "umt_rz_ctx_autotune": (_i, [_vp, _i, _pi, _pd]),
    "umt_rz_ctx_set_tiles": (_i, [_vp, _i]),
    "umt_rz_ctx_set_psi_prev": (_i, [_vp, _pd]),
    "umt_rz_ctx_get_psi_prev": (_i, [_vp, _pd]),
    "umt_rz_ctx_save_psi": (_i, [_vp]),
}

_lib = None


def load():
    """Load libumt_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (make -C umt_amd/csrc). "
            "The sweep has no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class UmtError(RuntimeError):
    """A nonzero status from the C-ABI, with the library's message."""

    def __init__(self, fn, code):
        msg = load().umt_last_error().decode(errors="replace")
        super().__init__(f"{fn} returned {code}: {msg}")
        self.code = code


def check(fn, code):
    if code != 0:
        raise UmtError(fn, code)
    return code
