"""CPU checks of the C ABI: both libraries load without a GPU, export every
function their headers declare, and refuse invalid input before touching a
device."""
import ctypes as C

import pytest

from nbody_amd import native, nbody_create_engine


@pytest.mark.parametrize("header,loader", [("nbody_hip.h", native.hip_lib), ("nbody_host.h", native.host_lib)])
def test_header_symbols_exported(header, loader):
    names = native.header_functions(native.INCLUDE / header)
    assert len(names) > 10
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_headers():
    for header, api in (("nbody_hip.h", native.HIP_API), ("nbody_host.h", native.HOST_API)):
        names = set(native.header_functions(native.INCLUDE / header))
        assert names == set(api), (header, names ^ set(api))


def test_no_gpu_paths_fail_cleanly():
    lib = native.hip_lib()
    if native.device_count() > 0:
        pytest.skip("a GPU is visible")
    ctx = C.c_void_p()
    assert lib.nbh_ctx_create(0, C.byref(ctx)) != 0
    assert b"device" in lib.nbh_last_error() or len(lib.nbh_last_error()) > 0
    assert nbody_create_engine({"engine": "hip"}) is None


def test_null_context_rejected():
    lib = native.hip_lib()
    r = C.c_double()
    assert lib.nbh_maxabs(None, None, 0, C.byref(r)) == 1
    assert lib.nbh_fill(None, None, 0.0, 1) == 1
    assert b"null context" in lib.nbh_last_error()


def test_factory_validates_before_device():
    assert nbody_create_engine({"engine": "nope"}) is None
    assert nbody_create_engine({"engine": "hip", "force_precision": "fp8"}) is None
    assert nbody_create_engine({"engine": "hip", "device": "0,1"}) is None
    assert nbody_create_engine({"engine": "hip", "block_size": "-3"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "tree_layout": "tree"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "traverse_type": "zigzag"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "tree_build_rate": "-1"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": "0"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "force_precision": "fp32_corr"}) is None   # direct only
    assert nbody_create_engine({"engine": "hip_block", "force_precision": "fp32"}) is None     # FP64 only
    assert nbody_create_engine({"engine": "hip_ah", "force_precision": "fp32"}) is None
    assert nbody_create_engine({"engine": "hip_ah", "full_recompute_rate": 0}) is None         # reference: x % 0
    assert nbody_create_engine({"engine": "hip_ah", "max_dist": "-1"}) is None
    assert nbody_create_engine({"engine": "hip_ah", "min_force": "abc"}) is None
    grp = {"engine": "hip_bh", "force_precision": "fp32", "bh_walk": "group"}
    assert nbody_create_engine({**grp, "bh_far_field": "tree"}) is None
    assert nbody_create_engine({**grp, "bh_far_field": "local"}) is None                     # needs quadrupole
    assert nbody_create_engine({**grp, "bh_multipole": "quadrupole", "bh_far_field": "local",
                                "bh_quad_near": "0.5"}) is None                              # cell inside the group
