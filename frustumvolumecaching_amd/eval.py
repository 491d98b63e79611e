"""Evaluation protocol of the reference (eval.cu:230-509) over libfvc.so.

For every camera pose of a path:

* render without the cache (one warm-up, then ``iters`` timed renders) ->
  ``timings.json["no_cache"]``, ``render_time_new_ms``;
* fill the cache at the pose itself and render from it; the cached image
  must stay within 35 dB of the full render (the reference's sanity gate,
  eval.cu:315-326);
* rotations: with the focus point f = the camera position projected onto
  the line through the origin along the view direction and c = pos - f,
  for 5, 10 and 15 degrees the vector c rotated about the camera's up axis,
  then about c at 6 angles 2 pi k / 6; the cache is filled at
  lookAt(f + v, f, up) and the ORIGINAL pose is rendered from it;
* translations: the cache filled at pos + s * view_dir for s in
  {-0.25, -0.1, -dt/2, dt/2, 0.1, 0.25} (dt = the stepsize), the original
  pose rendered from it.

Per pose a directory with ``metrics.json`` ({"rot": [...], "trans": [...]},
entries {image, n_bricks_set, n_bricks_reserved, overflow,
n_cache_samples, n_resamples}) and ``timings.json`` ({"no_cache": {...},
"rot": [...], "trans": [...], "render_time_new_ms"}, entries {image,
cache_init_ms, render_time_ms, timings}) as the reference writes them; each
metrics entry also carries ``psnr_vs_full_db`` (8-bit PSNR of the cached
render against the full render of the same pose).

Poses: a NeRF ``transforms.json`` (``frames[].transform_matrix``, the
reference's camera path input) or, without one, a synthetic path.

    python -m frustumvolumecaching_amd.eval --workload small --poses 2 --out eval_out
"""
import argparse
import json
import math
import os
import struct
import time
import zlib

import numpy as np
import torch

from . import lib
from . import synthetic as S

VIEW_DEGREES = (5, 10, 15)
N_CIRCLE = 6


def look_at(eye, center, up):
    """lookAt of eval.cu:73-85: columns right, up, view direction, eye."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    view = center - eye
    view /= np.linalg.norm(view)
    right = np.cross(up, view)
    right /= np.linalg.norm(right)
    up_new = np.cross(view, right)
    return np.stack([right, up_new, view, eye], axis=1).astype(np.float32)


def rotate(v, angle, axis):
    """glm::rotate of a vector about a unit axis (Rodrigues)."""
    v, k = np.asarray(v, np.float64), np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    c, s = math.cos(angle), math.sin(angle)
    return v * c + np.cross(k, v) * s + k * np.dot(k, v) * (1.0 - c)


def sub_poses(cam2world, stepsize):
    """(name, kind, cam2world) of the rotation and translation sub-poses."""
    m = np.asarray(cam2world, np.float64).reshape(3, 4)
    view = m[:, 2] / np.linalg.norm(m[:, 2])
    up = m[:, 1] / np.linalg.norm(m[:, 1])
    pos = m[:, 3]
    focus = pos + np.dot(-pos, view) * view
    cam_vec = pos - focus
    out = []
    for deg in VIEW_DEGREES:
        start = rotate(cam_vec, math.radians(deg), up)
        for k in range(N_CIRCLE):
            theta = k / N_CIRCLE * 2.0 * math.pi
            new_pos = rotate(start, theta, cam_vec)
            name = f"{deg}/{int(round(theta * 180.0 / math.pi))}"
            out.append((name, "rot", look_at(focus + new_pos, focus, up)))
    for step in (-0.25, -0.1, -stepsize * 0.5, stepsize * 0.5, 0.1, 0.25):
        t = m.astype(np.float32).copy()
        t[:, 3] = (pos + step * view).astype(np.float32)
        out.append((f"{step:f}", "trans", t))
    return out


def write_png(path, rgba):
    """8-bit RGBA PNG (zlib, no filter)."""
    h, w = rgba.shape[:2]
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
                chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def psnr_8bit(a, b):
    d = (a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)) / 255.0
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0.0 else -10.0 * math.log10(mse)


def load_poses(path):
    with open(path) as f:
        tj = json.load(f)
    out = []
    for i, fr in enumerate(tj["frames"]):
        m = np.asarray(fr["transform_matrix"], np.float32)[:3, :4]
        name = os.path.splitext(os.path.basename(fr.get("file_path", f"pose_{i:03d}")))[0]
        out.append((name, m))
    return out


def _timings(r):
    return [{"name": k, "time_ms": v["ms"], "n_executions": v["launches"]} for k, v in r.profile().items()]


def evaluate(model, wl, poses, out_dir, iters=3, write_images=False, width=None, height=None, log=print):
    """Run the protocol; returns {pose name: (metrics, timings)}."""
    r = lib.Renderer(model)
    results = {}
    W = width or wl.width
    H = height or wl.height
    for name, cam in poses:
        d = os.path.join(out_dir, name)
        os.makedirs(d, exist_ok=True)
        rm = wl.rm(cam, W, H)
        full = torch.zeros(H, W, 4, dtype=torch.uint8, device="cuda")
        img = torch.zeros_like(full)
        # render without the cache
        r.render_full(rm, full)
        r.profile_enable(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(iters):
            r.render_full(rm, full)
        torch.cuda.synchronize()
        t_new = (time.perf_counter() - t) * 1e3 / iters
        timings = {"no_cache": {"image": os.path.join(d, "no_cache"), "cache_init_ms": 0.0,
                                "render_time_ms": t_new, "timings": _timings(r)}}
        r.profile_enable(False)
        full_h = full.cpu().numpy()
        # cache at the pose itself: the sanity gate
        r.initCacheBlockwise(rm)
        r.render_cached(rm, img)
        sanity = psnr_8bit(img.cpu().numpy(), full_h)
        if sanity < 35.0:
            raise RuntimeError(f"sanity check failed (cache render) at {name}: PSNR {sanity:.2f} dB")
        if write_images:
            write_png(os.path.join(d, "no_cache.png"), full_h)
            write_png(os.path.join(d, "cache.png"), img.cpu().numpy())
        metrics = {"rot": [], "trans": []}
        timings.update({"rot": [], "trans": []})
        for sub, kind, c2w in sub_poses(cam, rm.stepsize_info.stepsize):
            sub_rm = wl.rm(c2w, W, H)
            torch.cuda.synchronize()
            t = time.perf_counter()
            init = r.initCacheBlockwise(sub_rm)
            torch.cuda.synchronize()
            t_init = (time.perf_counter() - t) * 1e3
            r.profile_enable(True)
            t = time.perf_counter()
            for _ in range(iters):
                st = r.render_cached(rm, img)
            torch.cuda.synchronize()
            t_r = (time.perf_counter() - t) * 1e3 / iters
            det = _timings(r)
            r.profile_enable(False)
            cached_h = img.cpu().numpy()
            metrics[kind].append({"image": sub, "n_bricks_set": int(init.n_bricks_set),
                                  "n_bricks_reserved": int(init.n_bricks_reserved),
                                  "overflow": bool(init.n_bricks_set >= init.n_bricks_reserved),
                                  "n_cache_samples": int(st.n_cache_hits), "n_resamples": int(st.n_resamples),
                                  "psnr_vs_full_db": round(psnr_8bit(cached_h, full_h), 3)})
            timings[kind].append({"image": ("r_" if kind == "rot" else "t_") + sub, "cache_init_ms": t_init,
                                  "render_time_ms": t_r, "timings": det})
            if write_images:
                sd = os.path.join(d, kind, *sub.split("/")[:-1]) if kind == "rot" else os.path.join(d, kind)
                os.makedirs(sd, exist_ok=True)
                write_png(os.path.join(sd, sub.split("/")[-1] + ".png"), cached_h)
        timings["render_time_new_ms"] = t_new
        with open(os.path.join(d, "metrics.json"), "w") as f:
            json.dump(metrics, f)
        with open(os.path.join(d, "timings.json"), "w") as f:
            json.dump(timings, f)
        psnrs = [e["psnr_vs_full_db"] for k in ("rot", "trans") for e in metrics[k]]
        log(f"[eval] {name}: full {t_new:.2f} ms, cached {np.mean([e['render_time_ms'] for e in timings['rot']]):.2f} "
            f"ms (rot mean), PSNR vs full min {min(psnrs):.2f} dB")
        results[name] = (metrics, timings)
    r.close()
    return results


def synthetic_poses(wl, kind, n):
    """An orbit around the small scene, or the 1080p path's frames."""
    if kind == "small":
        return [(f"pose_{k:03d}", S.look_at(S.eye_on_sphere(4.0, 30.0 + 360.0 * k / max(n, 1), 20.0), (0, 0, 0)))
                for k in range(n)]
    path = S.camera_path_1080p(wl, max(n, 1) * 10)
    return [(f"pose_{k:03d}", path[10 * k]) for k in range(n)]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=["small", "unbounded"], default="small")
    ap.add_argument("--nerfacc", help="nerfacc export directory (instead of a synthetic workload's model)")
    ap.add_argument("--transforms", help="NeRF transforms.json with the camera path")
    ap.add_argument("--poses", type=int, default=2, help="synthetic path length (without --transforms)")
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--write-images", action="store_true")
    ap.add_argument("--out", default="eval_out")
    a = ap.parse_args()
    wl = S.small_plain() if a.workload == "small" else S.unbounded_1080p()
    model = lib.NerfNetwork.from_nerfacc(a.nerfacc) if a.nerfacc else lib.NerfNetwork.from_workload(wl)
    if a.transforms:
        poses = load_poses(a.transforms)
    else:
        poses = synthetic_poses(wl, a.workload, a.poses)
    evaluate(model, wl, poses, a.out, a.iters, a.write_images, a.width, a.height)
    model.close()


if __name__ == "__main__":
    main()
