"""The C-ABI library loads on a host without a GPU and exports every entry
point declared in include/alb_abi.h; struct layouts agree with the ctypes mirror."""
import ctypes as C
import os
import re

import pytest

from pyalbsim_amd import abi, runtime

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "alb_abi.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(alb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    assert _declared() == sorted(runtime.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = runtime.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_layout_and_version():
    lib = runtime.load_library()
    assert lib.alb_abi_version() == abi.ALB_ABI_VERSION
    assert lib.alb_scene_size() == C.sizeof(abi.Scene)
    assert C.sizeof(abi.Rng) == 32
    assert C.sizeof(abi.Records) == 144 and C.sizeof(abi.Queries) == 144 and C.sizeof(abi.Map) == 160
    assert C.sizeof(abi.SceneIndex) == 32
    assert C.sizeof(abi.Layer) == 160 and abi.Layer.z_lut.offset == 88 and abi.Layer.phase_n.offset == 128
    assert abi.Layer.a_tab.offset == 136 and abi.Layer.tab_n.offset == 152
    import oracle

    assert oracle.layer_layout() == (C.sizeof(abi.Layer), abi.Layer.a_tab.offset, abi.Layer.tab_n.offset,
                                     abi.Layer.phase_n.offset)


def test_last_error_is_set_by_argument_checks():
    lib = runtime.load_library()
    rc = lib.alb_trace(None, None, 0, 0, 1, None, None, None, None)
    assert rc == abi.ALB_ERR_ARG
    assert b"null" in lib.alb_last_error()
    rc = lib.alb_map_build(None, 0, 2, None, None, 0, None)
    assert rc == abi.ALB_ERR_ARG
    rc = lib.alb_trace_batched(None, 1, None, None, 0, None, None, None, 256, 4, None, None)
    assert rc == abi.ALB_ERR_ARG and b"null" in lib.alb_last_error()
    rc = lib.alb_map_build_batched(None, 0, 2, None, None, None, 0, None)
    assert rc == abi.ALB_ERR_ARG
    rc = lib.alb_knn_batched(None, None, 1, 100, None, None, 0, 1, None, None, None, None, None, None)
    assert rc == abi.ALB_ERR_ARG
    rc = lib.alb_contrib_batched(None, 1, 100, None, 1, None, None, 0, 1, None, None, None, None, None, 0, None, 0, 0,
                                 None)
    assert rc == abi.ALB_ERR_ARG
    rc = lib.alb_radius_gather(None, None, 0, None, 0, 1, None, 0.05, None, 10, 0, 0, None, None, None, None)
    assert rc == abi.ALB_ERR_ARG and b"null" in lib.alb_last_error()


def test_product_path_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("host has a GPU")
    from pyalbsim_amd.model import build_scene
    from tests.scenes import single_shot

    with pytest.raises(runtime.AlbError):
        runtime.DeviceScene(build_scene(single_shot()))


def test_k_limit_matches_the_header_and_is_checked_before_device_work():
    m = re.search(r"#define ALB_MAX_K (\d+)", open(HEADER).read())
    assert m and int(m.group(1)) == abi.MAX_K
    from pyalbsim_amd.config.simulation import SimulationConfig
    from pyalbsim_amd.execution.batched import simulate_batched
    from pyalbsim_amd.execution.linear import simulate
    from pyalbsim_amd.model import build_scene

    for k in (abi.MAX_K + 1, 1000):
        model = build_scene(SimulationConfig(photon_mapping_k=k))
        with pytest.raises(runtime.AlbError, match=rf"outside \[1, {abi.MAX_K}\]"):
            simulate(model, 1000, 250, (1, 2))
        with pytest.raises(runtime.AlbError, match=rf"outside \[1, {abi.MAX_K}\]"):
            simulate_batched([model], 1000, 250, (1, 2))
    runtime.check_k([build_scene(SimulationConfig(photon_mapping_k=abi.MAX_K))])
