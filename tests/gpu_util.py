"""Helpers for the GPU parity tests: device buffers (torch), fp16 unpacking,
ulp distances."""
import numpy as np
import torch


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def empty(shape, dtype):
    return torch.empty(shape, dtype=dtype, device="cuda")


def host_f16(t):
    """uint16/int16 device tensor of fp16 bits -> float64 numpy."""
    return t.cpu().numpy().view(np.uint16).view(np.float16).astype(np.float64)


def f16_ulps(a, b, floor=2.0 ** -8):
    """Distance in fp16 ulps of max(|b|, floor).  The floor keeps outputs that
    cancel to nearly zero from being measured in ulps of a tiny value: a
    1-ulp change of an fp16 hidden activation moves such an output by far
    more than its own ulp."""
    b = np.asarray(b, np.float64)
    e = np.floor(np.log2(np.maximum(np.abs(b), floor)))
    ulp = 2.0 ** (e - 10)
    return np.abs(np.asarray(a, np.float64) - b) / ulp


def to_8bit(rgbt):
    rgb = rgbt[..., :3] + rgbt[..., 3:4]
    return np.clip((rgb * np.float32(256.0)).astype(np.int32), 0, 255)
