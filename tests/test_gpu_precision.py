"""BF16, two-term BF16 (BF16X3) and FP8-E4M3 MLP variants against the
oracle's emulation of the same scheme (same operand rounding / splitting
and per-layer scales, FP32 accumulation in a different order), through the
C ABI.

Bars (measured error x ~5-10 headroom):
* network outputs: BF16 and BF16X3 max 2e-3, 99.9th percentile 2e-4; FP8
  max 5e-2, 99.9th percentile 1e-2 (absolute, outputs of order 1);
* images within 2/255 of the same-mode oracle, full and cached renders;
* PSNR vs the FP64 oracle: FP16 above 60 dB; BF16X3 within 0.5 dB of FP16
  (the precision target); BF16 above 55 dB and FP8 above 45 dB, their
  measured floors with margin: the 0.5 dB target is not met by them (8 and
  4 mantissa bits against FP16's 11; the xfail test states it).
"""
import copy

import numpy as np
import pytest
import torch

from frustumvolumecaching_amd import abi, lib, synthetic as S

from .gpu_util import dev, empty, host_f16, to_8bit

pytestmark = pytest.mark.gpu

RES = 96
PRECS = {"bf16": (abi.PREC_BF16, 2e-3, 2e-4), "bf16x3": (abi.PREC_BF16X3, 2e-3, 2e-4),
         "fp8": (abi.PREC_FP8, 5e-2, 1e-2)}


@pytest.fixture(scope="module")
def wl():
    return S.small_plain(stepsize_setup=lib.stepsize_setup)


@pytest.fixture(scope="module")
def omodel(oracle_mod, wl):
    return oracle_mod.Model.from_workload(wl)


def _model(wl, prec):
    net = copy.copy(wl.net)
    net.precision = prec
    return lib.NerfNetwork(wl.scene, net, wl.occupancy, wl.params)


def _mode(oracle_mod, name):
    return {"bf16": oracle_mod.EMU_BF16, "bf16x3": oracle_mod.EMU_BF16X3, "fp8": oracle_mod.EMU_FP8}[name]


def _close(got, ref, tmax, tq):
    err = np.abs(np.asarray(got, np.float64) - ref)
    scale = max(1.0, np.abs(ref).max())
    assert err.max() <= tmax * scale, err.max()
    assert np.quantile(err, 0.999) <= tq * scale, np.quantile(err, 0.999)


@pytest.mark.parametrize("name", ["bf16", "bf16x3", "fp8"])
def test_network_outputs_match_oracle(wl, omodel, oracle_mod, name):
    prec, tmax, tq = PRECS[name]
    g = _model(wl, prec)
    rng = np.random.default_rng(5)
    p = rng.random((8192, 3)).astype(np.float32)
    d = rng.standard_normal((8192, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d01 = ((d + 1) / 2).astype(np.float32)
    mode = _mode(oracle_mod, name)
    out = empty((len(p), 16), torch.int16)
    g.inferenceLatents(dev(p), len(p), out)
    out4 = empty((len(p), 4), torch.int16)
    g.inference(dev(p), dev(d01), len(p), out4)
    hin = (rng.standard_normal((4096, 16)) * 0.05).astype(np.float16)
    outh = empty((4096, 4), torch.int16)
    g.inferenceHead(dev(hin.view(np.int16)), dev(d01[:4096]), 4096, outh)
    torch.cuda.synchronize()
    _close(host_f16(out)[:, :9], omodel.latents(p, mode)[:, :9], tmax, tq)
    _close(host_f16(out4), omodel.inference(p, d01, mode), tmax, tq)
    _close(host_f16(outh), omodel.head(hin.astype(np.float32), d01[:4096], mode), tmax, tq)
    g.close()


def _render(r, rm, cached=False):
    n = rm.cam_info.res_x * rm.cam_info.res_y
    img = torch.zeros(n * 4, dtype=torch.uint8, device="cuda")
    res = torch.zeros(n * 4, dtype=torch.float32, device="cuda")
    (r.render_cached if cached else r.render_full)(rm, img)
    r.last_result(res, n * 4)
    torch.cuda.synchronize()
    return res.cpu().numpy().reshape(rm.cam_info.res_y, rm.cam_info.res_x, 4)


@pytest.mark.parametrize("name", ["bf16", "bf16x3", "fp8"])
def test_renders_match_oracle(wl, omodel, oracle_mod, name):
    prec = PRECS[name][0]
    mode = _mode(oracle_mod, name)
    g = _model(wl, prec)
    r = lib.Renderer(g)
    rm = wl.rm(width=RES, height=RES)
    full = _render(r, rm)
    ref, _ = omodel.render_full(rm, mode)
    assert np.abs(to_8bit(full) - to_8bit(ref)).max() <= 2
    # cache filled in this precision, reprojected after a 2 degree yaw
    r.initCacheBlockwise(rm)
    rm2 = wl.rm(S.yaw(wl.cam2world, 2.0), RES, RES)
    cached = _render(r, rm2, cached=True)
    oc = omodel.cache_fill(rm, mode)
    ref_c, _, _ = oc.render(rm2)
    assert np.abs(to_8bit(cached) - to_8bit(ref_c)).max() <= 2
    r.close()
    g.close()


@pytest.fixture(scope="module")
def psnr(wl, omodel, oracle_mod):
    rm = wl.rm(width=RES, height=RES)
    ref64, _ = omodel.render_full(rm, oracle_mod.F64)
    out = {}
    for name, prec in [("fp16", abi.PREC_FP16), ("bf16", abi.PREC_BF16), ("bf16x3", abi.PREC_BF16X3),
                       ("fp8", abi.PREC_FP8)]:
        g = _model(wl, prec)
        r = lib.Renderer(g)
        out[name] = oracle_mod.psnr_8bit(to_8bit(_render(r, rm)), to_8bit(ref64))
        r.close()
        g.close()
    print("PSNR vs FP64 oracle (dB):", {k: round(v, 2) for k, v in out.items()})
    return out


def test_psnr_vs_fp64_per_precision(psnr):
    assert psnr["fp16"] > 60
    assert psnr["bf16x3"] >= psnr["fp16"] - 0.5
    assert psnr["bf16"] > 55 and psnr["fp8"] > 45


@pytest.mark.xfail(strict=True, reason="plain BF16 (8 mantissa bits) and FP8-E4M3 (4) MLP operands lose more than "
                                       "0.5 dB against FP16 (11); the two-term BF16X3 split meets the target")
def test_bf16_and_fp8_within_half_db_of_fp16(psnr):
    assert psnr["bf16"] >= psnr["fp16"] - 0.5 and psnr["fp8"] >= psnr["fp16"] - 0.5
