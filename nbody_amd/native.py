"""ctypes bindings of the two native libraries.

libnbody_hip.so  -- gfx950 kernels behind the C ABI of include/nbody_hip.h
libnbody.so      -- host layer (engines, solvers, bodies) behind include/nbody_host.h

Both are built in-tree by ``python -m nbody_amd.build``.  Loading fails loudly
when they are missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB_HIP = PKG / "libnbody_hip.so"
LIB_HOST = PKG / "libnbody.so"

_hip = None
_host = None

vp = C.c_void_p
sz = C.c_size_t
dbl = C.c_double
i32 = C.c_int
u64 = C.c_uint64
cstr = C.c_char_p
PD = C.POINTER(C.c_double)
PPD = C.POINTER(C.c_void_p)

# name: (restype, [argtypes])
HIP_API = {
    "nbh_last_error": (cstr, []),
    "nbh_version": (i32, []),
    "nbh_device_count": (i32, [C.POINTER(C.c_int)]),
    "nbh_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "nbh_ctx_destroy": (i32, [vp]),
    "nbh_ctx_synchronize": (i32, [vp]),
    "nbh_ctx_info": (i32, [vp, C.c_char_p, sz, C.POINTER(C.c_int)]),
    "nbh_malloc": (i32, [vp, sz, C.POINTER(vp)]),
    "nbh_free": (i32, [vp, vp]),
    "nbh_memcpy_h2d": (i32, [vp, vp, vp, sz]),
    "nbh_memcpy_d2h": (i32, [vp, vp, vp, sz]),
    "nbh_memcpy_d2d": (i32, [vp, vp, vp, sz]),
    "nbh_fill": (i32, [vp, vp, dbl, sz]),
    "nbh_fmaddn": (i32, [vp, vp, vp, PPD, PD, sz, sz]),
    "nbh_fmaddn_corr": (i32, [vp, vp, vp, PPD, PD, sz, sz]),
    "nbh_maxabs": (i32, [vp, vp, sz, PD]),
    "nbh_snapshot_create": (i32, [vp, sz, i32, C.POINTER(vp)]),
    "nbh_snapshot_destroy": (i32, [vp]),
    "nbh_snapshot_begin": (i32, [vp, vp, i32, vp]),
    "nbh_snapshot_wait": (i32, [vp, vp, i32, C.POINTER(vp)]),
    "nbh_clamp": (i32, [vp, vp, sz, dbl]),
    "nbh_fcompute_direct": (i32, [vp, vp, vp, vp, sz, i32, i32]),
    "nbh_fcompute_direct_noalloc": (i32, [vp, vp, vp, vp, sz, i32, i32]),
    "nbh_direct_sample": (i32, [vp, vp, vp, sz, vp, sz, vp]),
    "nbh_ah_create": (i32, [vp, sz, C.POINTER(vp)]),
    "nbh_ah_destroy": (i32, [vp]),
    "nbh_ah_fcompute_full": (i32, [vp, vp, vp, vp, vp, dbl, dbl]),
    "nbh_ah_fcompute_sparse": (i32, [vp, vp, vp, vp, vp]),
    "nbh_ah_lists": (i32, [vp, vp, C.POINTER(sz), vp, vp, vp]),
    "nbh_tree_create": (i32, [vp, sz, C.POINTER(vp)]),
    "nbh_tree_destroy": (i32, [vp]),
    "nbh_tree_build": (i32, [vp, vp, vp, vp, dbl]),
    "nbh_tree_refit": (i32, [vp, vp, vp, dbl]),
    "nbh_tree_build_radix": (i32, [vp, vp, vp, vp, dbl]),
    "nbh_tree_set_radix_lean": (i32, [vp, vp, i32]),
    "nbh_tree_get_children": (i32, [vp, vp, vp, vp]),
    "nbh_bh_fcompute": (i32, [vp, vp, vp, vp, vp, i32, i32]),
    "nbh_bh_fcompute_group": (i32, [vp, vp, vp, vp]),
    "nbh_bh_fcompute_group_quad": (i32, [vp, vp, vp, vp, C.c_float]),
    "nbh_bh_fcompute_group_local": (i32, [vp, vp, vp, vp, C.c_float]),
    "nbh_bh_fcompute_group_local2": (i32, [vp, vp, vp, vp, C.c_float]),
    "nbh_bh_fcompute_group_local2_mlp": (i32, [vp, vp, vp, vp, vp, C.c_float]),
    "nbh_bh_super_samples": (i32, [vp, vp, C.c_float, vp, sz, vp, vp, vp, vp, sz, C.POINTER(sz)]),
    "nbh_tree_set_list_cap": (i32, [vp, vp, C.c_int]),
    "nbh_tree_bytes": (i32, [vp, vp, C.POINTER(sz)]),
    "nbh_tree_get": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "nbh_bh_counters": (i32, [vp, vp, C.POINTER(u64), C.POINTER(u64)]),
    "nbh_bh_walk_counters": (i32, [vp, vp, C.POINTER(u64)]),
    "nbh_mlp_create": (i32, [vp, u64, C.POINTER(vp)]),
    "nbh_mlp_destroy": (i32, [vp]),
    "nbh_mlp_param_count": (i32, [C.POINTER(C.c_int)]),
    "nbh_mlp_set_params": (i32, [vp, vp, vp]),
    "nbh_mlp_get_params": (i32, [vp, vp, vp]),
    "nbh_mlp_forward": (i32, [vp, vp, vp, sz, vp]),
    "nbh_mlp_gradient": (i32, [vp, vp, vp, vp, vp, sz, vp, C.POINTER(C.c_float)]),
    "nbh_mlp_train": (i32, [vp, vp, vp, vp, vp, sz, i32, sz, C.c_float, vp]),
    "nbh_bh_band_samples": (i32, [vp, vp, vp, sz, C.c_float, C.c_float, vp, vp, vp, vp, sz, C.POINTER(sz)]),
    "nbh_bh_fcompute_group_mlp": (i32, [vp, vp, vp, vp, vp, C.c_float, C.c_float]),
    "nbh_stats": (i32, [vp, vp, vp, sz, i32, PD]),
    "nbh_timing_enable": (i32, [vp, i32]),
    "nbh_timing_reset": (i32, [vp]),
    "nbh_timing_read": (i32, [vp, cstr, PD, C.POINTER(u64)]),
}

HOST_API = {
    "nbs_log_count": (sz, []),
    "nbs_log_last": (cstr, []),
    "nbs_log_quiet": (None, [i32]),
    "nbs_data_create": (vp, []),
    "nbs_data_destroy": (None, [vp]),
    "nbs_data_load": (i32, [vp, cstr, cstr]),
    "nbs_data_save": (i32, [vp, cstr]),
    "nbs_data_make_universe": (None, [vp, sz, dbl, dbl, dbl]),
    "nbs_data_generate": (i32, [vp, cstr, sz, u64]),
    "nbs_data_set": (None, [vp, sz, PD, PD, PD, PD]),
    "nbs_data_count": (sz, [vp]),
    "nbs_data_get": (None, [vp, PD, PD, PD, PD]),
    "nbs_data_time": (dbl, [vp]),
    "nbs_data_step": (sz, [vp]),
    "nbs_data_set_time": (None, [vp, dbl]),
    "nbs_data_set_step": (None, [vp, sz]),
    "nbs_data_advise_time": (None, [vp, dbl]),
    "nbs_data_box_size": (sz, [vp]),
    "nbs_data_set_check_list": (None, [vp, cstr]),
    "nbs_data_set_quiet": (None, [vp, i32]),
    "nbs_data_print_statistics": (None, [vp, vp]),
    "nbs_data_statistics": (None, [vp, vp, PD]),
    "nbs_data_errors": (None, [vp, PD]),
    "nbs_data_is_equal": (i32, [vp, vp, dbl]),
    "nbs_engine_create": (vp, [cstr]),
    "nbs_engine_destroy": (None, [vp]),
    "nbs_engine_type_name": (cstr, [vp]),
    "nbs_engine_init": (i32, [vp, vp]),
    "nbs_engine_get_data": (None, [vp, vp]),
    "nbs_engine_problem_size": (sz, [vp]),
    "nbs_engine_get_y": (vp, [vp]),
    "nbs_engine_advise_time": (None, [vp, dbl]),
    "nbs_engine_get_time": (dbl, [vp]),
    "nbs_engine_set_time": (None, [vp, dbl]),
    "nbs_engine_get_step": (sz, [vp]),
    "nbs_engine_set_step": (None, [vp, sz]),
    "nbs_engine_fcompute": (None, [vp, dbl, vp, vp]),
    "nbs_engine_clamp": (None, [vp, vp, dbl]),
    "nbs_engine_create_buffer": (vp, [vp, sz]),
    "nbs_engine_free_buffer": (None, [vp, vp]),
    "nbs_engine_read_buffer": (None, [vp, vp, vp]),
    "nbs_engine_write_buffer": (None, [vp, vp, vp]),
    "nbs_engine_copy_buffer": (None, [vp, vp, vp]),
    "nbs_engine_fill_buffer": (None, [vp, vp, dbl]),
    "nbs_engine_fmadd_inplace": (None, [vp, vp, vp, dbl]),
    "nbs_engine_fmadd": (None, [vp, vp, vp, vp, dbl]),
    "nbs_engine_fmaddn_inplace": (None, [vp, vp, PPD, sz, PD, sz]),
    "nbs_engine_fmaddn_corr": (None, [vp, vp, vp, PPD, sz, PD, sz]),
    "nbs_engine_fmaddn": (None, [vp, vp, vp, PPD, sz, PD, sz]),
    "nbs_engine_fmaxabs": (None, [vp, vp, PD]),
    "nbs_engine_compute_count": (sz, [vp]),
    "nbs_engine_print_info": (None, [vp]),
    "nbs_engine_has_error": (i32, [vp]),
    "nbs_engine_hip_context": (vp, [vp]),
    "nbs_engine_hip_tree": (vp, [vp]),
    "nbs_engine_hip_neighbours": (vp, [vp]),
    "nbs_memory_device_ptr": (vp, [vp]),
    "nbs_memory_size": (sz, [vp]),
    "nbs_memory_fake_create": (vp, [sz]),
    "nbs_memory_fake_destroy": (None, [vp]),
    "nbs_solver_create": (vp, [cstr]),
    "nbs_solver_destroy": (None, [vp]),
    "nbs_solver_type_name": (cstr, [vp]),
    "nbs_solver_set_engine": (None, [vp, vp]),
    "nbs_solver_set_time_step": (None, [vp, dbl, dbl]),
    "nbs_solver_run": (i32, [vp, vp, vp, dbl, dbl, dbl]),
    "nbs_parse_arguments": (sz, [i32, C.POINTER(C.c_char_p), C.c_char_p, sz]),
    "nbs_stream_create": (vp, []),
    "nbs_stream_destroy": (None, [vp]),
    "nbs_stream_open": (i32, [vp, cstr, C.c_int64, vp]),
    "nbs_stream_write": (i32, [vp, vp]),
    "nbs_stream_close": (None, [vp]),
    "nbs_stream_reader_create": (vp, []),
    "nbs_stream_reader_destroy": (None, [vp]),
    "nbs_stream_reader_load": (i32, [vp, cstr]),
    "nbs_stream_reader_close": (None, [vp]),
    "nbs_stream_reader_frame_count": (sz, [vp]),
    "nbs_stream_reader_steps_count": (sz, [vp]),
    "nbs_stream_reader_max_time": (dbl, [vp]),
    "nbs_stream_reader_seek": (i32, [vp, sz]),
    "nbs_stream_reader_current_frame": (sz, [vp]),
    "nbs_stream_reader_body_count": (sz, [vp]),
    "nbs_stream_reader_coord_size": (sz, [vp]),
    "nbs_stream_reader_box_size": (dbl, [vp]),
    "nbs_stream_reader_last_file_n": (sz, [vp]),
    "nbs_stream_reader_read": (i32, [vp, vp]),
    "nbs_data_get_color": (None, [vp, C.POINTER(C.c_float)]),
    "nbs_solver_advise": (None, [vp, dbl]),
    "nbs_solver_print_info": (None, [vp]),
    "nbs_solver_rk_configure": (i32, [vp, sz, sz, sz, dbl]),
    "nbs_solver_rk_table": (i32, [vp, C.POINTER(sz), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "nbs_solver_rk_coeffs": (i32, [vp, PD, PD, PD, PD]),
    "nbs_solver_rk_substeps": (sz, [vp]),
    "nbs_solver_create_rk_butcher": (vp, [cstr]),
    "nbs_solver_create_adams": (vp, [sz, i32]),
    "nbs_data_resize": (None, [vp, sz]),
    "nbs_extrapolator_create": (vp, [cstr, vp, sz, C.POINTER(sz), sz]),
    "nbs_extrapolator_destroy": (None, [vp]),
    "nbs_extrapolator_update_table": (None, [vp, sz, vp]),
    "nbs_extrapolator_estimate_error": (dbl, [vp, sz]),
    "nbs_extrapolator_extrapolate": (None, [vp, sz, vp]),
}


def _bind(lib, api):
    for name, (res, args) in api.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip_lib():
    """libnbody_hip.so (kernels + C ABI)."""
    global _hip
    if _hip is None:
        if not LIB_HIP.exists():
            raise RuntimeError(f"{LIB_HIP} is not built: run `python -m nbody_amd.build`")
        _hip = _bind(C.CDLL(str(LIB_HIP), mode=C.RTLD_GLOBAL), HIP_API)
    return _hip


def host_lib():
    """libnbody.so (engines, solvers, bodies)."""
    global _host
    if _host is None:
        hip_lib()
        if not LIB_HOST.exists():
            raise RuntimeError(f"{LIB_HOST} is not built: run `python -m nbody_amd.build`")
        _host = _bind(C.CDLL(str(LIB_HOST), mode=C.RTLD_GLOBAL), HOST_API)
    return _host


def device_count() -> int:
    n = C.c_int(0)
    hip_lib().nbh_device_count(C.byref(n))
    return n.value


def check(status: int, what: str = "nbh call") -> None:
    if status != 0:
        msg = hip_lib().nbh_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({status}): {msg}")


def header_functions(header: Path) -> list[str]:
    """Function names declared in one of include/*.h."""
    import re
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nb[hs]_[a-z0-9_]+)\s*\(", text)))
