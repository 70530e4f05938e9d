"""Launch-shape autotuner for the HIP backend.

Picks, by timing candidates on the device, the shapes the kernels leave
open:
  * the guided filter's strip width and rows per strip, for an image size,
    radius and channel count: radii up to 96 run one-wave strips of 128 or
    256 columns, larger radii workgroup strips of 256, 512 or 1024 threads
    (the middle width - 2r columns of a strip are its output columns);
  * the persistent solve's chunk size (pixels a wave takes from the device
    queue at once), on a sample of the image's pixels.
Changing a shape does not change what is computed: every candidate runs the
same per-pixel arithmetic.  Results are cached per device in a JSON file
($UNBLENDING_AUTOTUNE_CACHE, default ~/.cache/unblending_amd/autotune.json);
`tuned_params` returns ready-to-use hip_params.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import hip_params

FILTER_WAVE_MAX_RADIUS = 96
FILTER_WAVE_BLOCKS = (128, 256)
FILTER_BLOCKS = (256, 512, 1024)
FILTER_ROWS = (0, 32, 64, 128, 256, 512)
CHUNKS = (32, 64, 128, 256, 1024)


def _torch():
    import torch
    return torch


def _cache_path():
    return os.environ.get("UNBLENDING_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "unblending_amd", "autotune.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    cache = _load_cache()
    cache[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the result is still returned


def _device_name():
    torch = _torch()
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _time_ms(fn, repeats):
    torch = _torch()
    fn()  # warm-up (code objects, caches)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(repeats):
        start.record()
        fn()
        stop.record()
        stop.synchronize()
        best = min(best, start.elapsed_time(stop))
    return best


def balanced_rows(width, height, radius, block, cus, rounds=(1, 2, 3, 4)):
    """Rows per strip that make a one-wave-strip launch k whole rounds of
    tiles on `cus` compute units (a pass-A tile is one workgroup of up to
    eight waves, which a CU holds one of): a launch of a few tiles more than
    a round leaves most CUs idle for its last round."""
    out = []
    tw = block - 2 * radius
    if tw < 1 or cus < 1:
        return out
    strips = -(-width // tw)
    for k in rounds:
        chunks = max(1, (k * cus) // strips)
        rows = max(32, -(-height // chunks))
        if rows not in out:
            out.append(rows)
    return out


def filter_candidates(radius, width=0, height=0, cus=0):
    out = []
    if 0 < radius <= FILTER_WAVE_MAX_RADIUS:
        for block in FILTER_WAVE_BLOCKS:
            if block - 2 * radius < 64:
                continue
            rows_set = list(FILTER_ROWS)
            for rows in balanced_rows(width, height, radius, block, cus):
                if rows not in rows_set:
                    rows_set.append(rows)
            for rows in rows_set:
                out.append((block, rows))
        return out
    for block in FILTER_BLOCKS:
        if block - 2 * radius < 64:
            continue
        for rows in FILTER_ROWS:
            out.append((block, rows))
    if not out:
        out = [(1024, rows) for rows in FILTER_ROWS]
    return out


def tune_filter(width, height, radius, channels, eps=1e-4, repeats=3, use_cache=True, log=None):
    """Best (block, rows, ms) for the guided filter of `channels` planes of a
    width x height image with this radius (synthetic inputs of that shape)."""
    torch = _torch()
    from . import device
    key = "%s|filter2|%dx%d|r%d|c%d" % (_device_name(), width, height, radius, channels)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return tuple(hit)
    rng = np.random.default_rng(0)
    rgb = torch.from_numpy(rng.uniform(0, 1, (height, width, 3))).cuda()
    planes = [torch.from_numpy(rng.uniform(0, 1, (height, width))).cuda() for _ in range(channels)]
    out = [torch.empty_like(p) for p in planes]
    n = (device._lib.unblending_hip_filter_workspace_size(width, height, channels) + 7) // 8
    ws = torch.empty(max(1, n), dtype=torch.float64, device="cuda")
    best = None
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    for block, rows in filter_candidates(radius, width, height, cus):
        ms = _time_ms(lambda: device.guided_filter(rgb, planes, radius, eps, out=out, ws=ws, block=block, rows=rows),
                      repeats)
        if log:
            log("filter %dx%d r=%d C=%d block %d rows %d: %.3f ms" % (width, height, radius, channels, block, rows, ms))
        if best is None or ms < best[2]:
            best = (block, rows, ms)
    _store(key, list(best))
    return best


def tune_chunk(rgb, layer_infos, sample_rows=256, repeats=1, use_cache=True, log=None):
    """Best (chunk, ms) for the first-pass solve, timed on the first
    sample_rows rows of rgb (h, w, 3) on the device."""
    torch = _torch()
    from . import device
    h, w, _ = rgb.shape
    rows = min(h, sample_rows)
    key = "%s|chunk|w%d|L%d" % (_device_name(), w, len(layer_infos))
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return tuple(hit)
    sample = rgb[:rows].contiguous()
    L = len(layer_infos)
    out = torch.empty((L, rows, w, 4), dtype=torch.float64, device=rgb.device)
    ws = device.workspace(w, rows, L, False, device=rgb.device)
    best = None
    for chunk in CHUNKS:
        p = hip_params(chunk_size=chunk)
        ms = _time_ms(lambda: device.unmix(sample, layer_infos, out=out, ws=ws, params=p), repeats)
        if log:
            log("chunk %d: %.3f ms" % (chunk, ms))
        if best is None or ms < best[1]:
            best = (chunk, ms)
    _store(key, list(best))
    return best


def tuned_params(rgb, layer_infos, radius, eps, channels=None, log=None, use_cache=True, **kw):
    """hip_params with the tuned filter shape and chunk size for this image."""
    h, w, _ = rgb.shape
    C = channels if channels is not None else len(layer_infos)
    block, rows, _ = tune_filter(w, h, radius, C, eps, log=log, use_cache=use_cache)
    chunk, _ = tune_chunk(rgb, layer_infos, log=log, use_cache=use_cache)
    return hip_params(radius=radius, eps=eps, chunk_size=chunk, filter_block=block, filter_rows=rows, **kw)
