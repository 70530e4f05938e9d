"""The LayerInfos reader and writer of include/unblending_hip.hpp (the
reference's import_layer_infos / export_layer_infos, unblending_io.cpp:69-170,
LayerInfos.md), compiled with g++ into a small driver (no device work).

Inputs: the reference's example LayerInfos files (tests/golden/layer_infos),
plus written cases for a full covariance, the plus comp-op and bad input.
Expected values: Python's json module and numpy's inverse, independently."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "layer_infos")
NAMES = ["Normal", "Multiply", "Screen", "Overlay", "Darken", "Lighten", "ColorDodge", "ColorBurn", "HardLight",
         "SoftLight", "Difference", "Exclusion", "LinearDodge"]


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tool") / "layer_infos_tool")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "layer_infos_tool.cpp"), "-o", exe], check=True)
    return exe


def ref_dump(v):
    """JSON in the reference writer's layout: sorted keys, ', ' and ': '
    separators, numbers as %.17g."""
    if isinstance(v, dict):
        return "{" + ", ".join(json.dumps(k) + ": " + ref_dump(v[k]) for k in sorted(v)) + "}"
    if isinstance(v, list):
        return "[" + ", ".join(ref_dump(x) for x in v) + "]"
    if isinstance(v, str):
        return json.dumps(v)
    return "%.17g" % v


def parse(tool, path):
    out = subprocess.run([tool, "parse", path], capture_output=True, text=True, check=True).stdout
    rows = []
    for line in out.splitlines():
        v = line.split()
        rows.append({"mode": int(v[0]), "comp": tuple(int(x) for x in v[1:4]), "mu": np.array([float(x) for x in v[4:7]]),
                     "precision": np.array([float(x) for x in v[7:16]]).reshape(3, 3),
                     "sigma": np.array([float(x) for x in v[16:25]]).reshape(3, 3)})
    return rows


def expected(entry):
    m = entry["color_model"]
    var = m["color_variance"]
    sigma = np.eye(3) * var if isinstance(var, (int, float)) else np.array(var, dtype=float).reshape(3, 3).T
    return {"mode": NAMES.index(entry["mode"]),
            "comp": (1, 1, 1) if entry["comp_op"] == "source-over" else (2, 1, 1),
            "mu": np.array(m["primary_color"], dtype=float), "sigma": sigma}


@pytest.mark.parametrize("name", ["magic.json", "electron.json", "cezanne.json"])
def test_reference_examples_parse(tool, name):
    path = os.path.join(GOLDEN, name)
    ref = [expected(e) for e in json.load(open(path))]
    got = parse(tool, path)
    assert len(got) == len(ref) > 0
    for g, r in zip(got, ref):
        assert g["mode"] == r["mode"] and g["comp"] == r["comp"]
        assert np.array_equal(g["mu"], r["mu"])
        np.testing.assert_allclose(g["precision"], np.linalg.inv(r["sigma"]), rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", ["magic.json", "electron.json", "cezanne.json"])
def test_export_round_trip_in_the_reference_format(tool, tmp_path, name):
    """The writer's layout is the reference writer's (sorted keys, ', ' and
    ': ' separators, %.17g numbers, a uniform variance as one number); a
    written file reads back to the same layers."""
    path = os.path.join(GOLDEN, name)
    subprocess.run([tool, "export", path, str(tmp_path)], check=True)
    text = open(tmp_path / "layer_infos.json").read()
    src = json.load(open(path))
    back = json.loads(text)
    assert [e["mode"] for e in back] == [e["mode"] for e in src]
    assert [e["comp_op"] for e in back] == [e["comp_op"] for e in src]
    for a, b in zip(back, src):
        assert a["color_model"]["primary_color"] == b["color_model"]["primary_color"]
        assert isinstance(a["color_model"]["color_variance"], float)
        assert a["color_model"]["color_variance"] == pytest.approx(b["color_model"]["color_variance"], rel=1e-14)
    # byte layout: that of the reference's own files (which ref_dump reproduces)
    assert ref_dump(src) == open(path).read().strip()
    assert text == ref_dump(back)
    again = parse(tool, str(tmp_path / "layer_infos.json"))
    first = parse(tool, path)
    for a, b in zip(again, first):
        np.testing.assert_allclose(a["precision"], b["precision"], rtol=1e-13)


def test_full_covariance_and_plus(tool, tmp_path):
    """Nine numbers are a covariance read column by column
    (unblending_io.cpp:41-52); anything but "source-over" is plus."""
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 3))
    sigma = a @ a.T + 0.1 * np.eye(3)
    sigma[0, 1] += 1e-3  # not symmetric: the reading order matters
    entry = {"color_model": {"color_variance": list(sigma.T.ravel()), "primary_color": [0.25, 0.5, 0.75]},
             "comp_op": "plus", "mode": "SoftLight"}
    path = tmp_path / "in.json"
    path.write_text(json.dumps([entry]))
    (g,) = parse(tool, str(path))
    assert g["mode"] == 9 and g["comp"] == (2, 1, 1)
    np.testing.assert_allclose(g["precision"], np.linalg.inv(sigma), rtol=1e-12)
    subprocess.run([tool, "export", str(path), str(tmp_path)], check=True)
    back = json.load(open(tmp_path / "layer_infos.json"))
    assert back[0]["comp_op"] == "plus" and len(back[0]["color_model"]["color_variance"]) == 9
    # the writer's element order: row-major of the covariance (unblending_io.cpp:14-24)
    np.testing.assert_allclose(np.array(back[0]["color_model"]["color_variance"]).reshape(3, 3), sigma, rtol=1e-12)


def test_names_and_bad_input(tool, tmp_path):
    out = subprocess.run([tool, "names"], capture_output=True, text=True, check=True).stdout.split()
    assert out[0::2] == NAMES and [int(v) for v in out[1::2]] == list(range(13))
    bad = tmp_path / "bad.json"
    for text in ('[{"color_model": {"color_variance": 0.1, "primary_color": [0, 0, 0]}, "mode": "Glow"}]',
                 '[{"color_model": {"color_variance": [1, 2], "primary_color": [0, 0, 0]}, "mode": "Normal"}]',
                 '[{"color_model": {"color_variance": 0.1, "primary_color": [0, 0, 0]}, "mode": "Normal"}',
                 '{"mode": "Normal"}'):
        bad.write_text(text)
        r = subprocess.run([tool, "parse", str(bad)], capture_output=True, text=True)
        assert r.returncode == 3 and "error" in r.stderr
