"""GPU: the command-line tool (unblending_amd/unblending-cli, the reference's
unblending-cli/main.cpp:13-96 without Qt) end to end on a small PPM: both
passes with an opaque, smoothed background, layer and recomposite files, and
the layers equal to the Python API's result on the same 8-bit input."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import layer_infos

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "unblending_amd", "unblending-cli")
NAMES = ["Normal", "Multiply", "Screen", "Overlay", "Darken", "Lighten", "ColorDodge", "ColorBurn", "HardLight",
         "SoftLight", "Difference", "Exclusion", "LinearDodge"]


def read_pam(path):
    data = open(path, "rb").read()
    head, body = data.split(b"ENDHDR\n", 1)
    fields = dict(l.split(b" ", 1) for l in head.split(b"\n")[1:] if b" " in l)
    w, h, d = int(fields[b"WIDTH"]), int(fields[b"HEIGHT"]), int(fields[b"DEPTH"])
    return np.frombuffer(body, dtype=np.uint8).reshape(h, w, d)


def read_pnm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    w, h = (int(v) for v in parts[1].split())
    arr = np.frombuffer(parts[3], dtype=np.uint8)
    return arr.reshape(h, w, -1)


def to8(v):
    return np.round(np.clip(v, 0.0, 1.0) * 255.0).astype(np.uint8)


def test_cli_end_to_end(tmp_path):
    import unblending_amd as U
    from unblending_amd import synth
    assert os.path.exists(CLI), "unblending-cli is not built"
    wl = synth.config(1, (48, 40))
    rgb8 = to8(wl.rgb)
    img = tmp_path / "in.ppm"
    img.write_bytes(b"P6\n48 40\n255\n" + rgb8.tobytes())
    var = 0.01  # c1's covariance: 0.01 I, written as one number
    entries = [{"color_model": {"color_variance": var, "primary_color": [float(v) for v in l.mean]},
                "comp_op": "source-over", "mode": NAMES[l.mode]} for l in wl.layers]
    (tmp_path / "layers.json").write_text(json.dumps(entries))
    out = tmp_path / "out"
    print("case cli 48x40 c1 layers", flush=True)
    r = subprocess.run([CLI, "-v", "-e", "-o", str(out), str(img), str(tmp_path / "layers.json")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    L = len(wl.layers)
    for i, l in enumerate(wl.layers):
        for name in ("layer_%d_%s.pam" % (i, NAMES[l.mode]), "non-smoothed-layer_%d_%s.pam" % (i, NAMES[l.mode]),
                     "layer-alpha_%d.pgm" % i, "non-smoothed-layer-alpha_%d.pgm" % i):
            assert (out / name).exists(), name
    for name in ("input.ppm", "recomposited.ppm", "non-smoothed-recomposited.ppm", "layer_infos.json"):
        assert (out / name).exists(), name
    assert np.array_equal(read_pnm(out / "input.ppm"), rgb8)

    # the same passes through the Python API on the same input; the precision
    # the tool derives from the variance by the cofactor inverse
    x = rgb8.astype(np.float64) / 255.0
    p = (var * var) * (1.0 / (var * (var * var)))
    infos = layer_infos([type(l)(l.mode, l.mean, np.eye(3) * p, l.comp) for l in wl.layers])
    first = U.compute_color_unmixing(x, infos, True)
    second = U.perform_matte_refinement(x, first, infos, True, True)
    for i, l in enumerate(wl.layers):
        assert np.array_equal(read_pam(out / ("layer_%d_%s.pam" % (i, NAMES[l.mode]))), to8(second[i]))
        assert np.array_equal(read_pam(out / ("non-smoothed-layer_%d_%s.pam" % (i, NAMES[l.mode]))), to8(first[i]))
    rec = read_pnm(out / "recomposited.ppm").astype(float)
    err = np.abs(rec - rgb8.astype(float)).mean()
    print("  recomposite mean |error| %.3f levels" % err, flush=True)
    assert err < 2.0
    assert json.load(open(out / "layer_infos.json"))[1]["mode"] == NAMES[wl.layers[1].mode]
