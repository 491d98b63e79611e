"""The C-ABI library loads, exports every symbol include/fvc.h declares, and
its host-only entry points agree with the oracle (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from frustumvolumecaching_amd import abi, lib, synthetic as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fvc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    declared = _declared()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(lib.exported_symbols()) == declared


def test_struct_sizes_match_header_layout():
    assert C.sizeof(abi.SceneInfo) == 4 * 4 + 12 * 4 + 3 * 4
    assert C.sizeof(abi.StepsizeInfo) == 12 * 4
    assert C.sizeof(abi.CameraInfo) == 3 * 4 + 16 + 2 * 48
    assert C.sizeof(abi.RaymarchInfo) == C.sizeof(abi.CameraInfo) + C.sizeof(abi.StepsizeInfo) + 12
    assert C.sizeof(abi.RenderStats) == 40


@pytest.mark.parametrize("make", [S.small_plain, S.unbounded_1080p])
def test_stepsize_setup_bit_identical_to_oracle(oracle_mod, make):
    wl = make(width=32, height=32, stepsize_setup=lib.stepsize_setup)
    ref = abi.StepsizeInfo()
    C.memmove(C.byref(ref), C.byref(wl.step), C.sizeof(ref))
    oracle_mod.stepsize_setup(ref, wl.scene)
    a = np.frombuffer(bytes(wl.step), np.uint32)
    b = np.frombuffer(bytes(ref), np.uint32)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("log2_t", [19, 21])
def test_param_counts_agree(oracle_mod, log2_t):
    net = S.net_desc(log2_hashmap_size=log2_t)
    assert lib.net_param_counts(net) == S.param_counts(net) == oracle_mod.net_param_counts(net)


def test_unsupported_network_is_rejected():
    net = S.net_desc()
    net.head_width = 128
    with pytest.raises(lib.FvcError) as e:
        lib.net_param_counts(net)
    assert e.value.code == lib.E_UNSUPPORTED


def test_model_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    wl = S.small_plain(width=16, height=16, stepsize_setup=lib.stepsize_setup)
    with pytest.raises(lib.FvcError) as e:
        lib.NerfNetwork.from_workload(wl)
    assert e.value.code == lib.E_NO_DEVICE
