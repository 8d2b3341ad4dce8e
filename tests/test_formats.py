"""The reference's file formats: .env options, the <scene>.bin cache and the
benchmark result file (hippie_amd/formats.py)."""
import os
import struct
import warnings

import numpy as np
import pytest

from hippie_amd import formats as F

ENV_TEXT = """\
Application {
\t#mode benchmark
}

Benchmark {
\toutput sweep-a   # trailing comment
\timages yes
}
Resolution { width 640 height 480 }
Scene {
\tfilename ../data/scene.obj
\tlight 0.48 0.08 0.17
}
Bvh
{
\tlayout bin
\tct 3.5
\tsbvhAlpha 0.00001
\tplocRadius 25
\tadaptiveLeafSize OFF
\tmethod ploc
\tmethod lbvh
}
"""


def _env(text):
    e = F.Environment()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ok = e.read_env_text(text)
    return e, ok


def test_env_blocks_values_and_defaults():
    e, ok = _env(ENV_TEXT)
    assert ok
    assert e.get("Application.mode") == "interactive"          # commented out: default
    assert e.get("Benchmark.output") == "sweep-a"
    assert e.get("Benchmark.images") is True
    assert (e.get("Resolution.width"), e.get("Resolution.height")) == (640, 480)
    assert e.get("Scene.filename") == "../data/scene.obj"
    assert e.get("Scene.light") == tuple(float(np.float32(x)) for x in (0.48, 0.08, 0.17))
    assert e.get("Bvh.ct") == 3.5
    assert e.get("Bvh.plocRadius") == 25
    assert e.get("Bvh.adaptiveLeafSize") is False
    assert e.get("Bvh.method") == "ploc"                       # first value wins
    assert e.get_values("Bvh.method") == ["ploc", "lbvh"]
    assert e.get("Bvh.ci") == 2.0                              # default
    assert e.get("Scene.environment") is None                  # no default, no value
    assert e.get("No.such") is None


def test_env_canonical_text_of_values():
    # filterValue stores QString::number(value): six significant digits.
    assert F.filter_value("0.00001", F.OPT_FLOAT) == "1e-05"
    assert F.filter_value("3.14159265", F.OPT_FLOAT) == "3.14159"
    assert F.filter_value("100000000", F.OPT_FLOAT) == "1e+08"
    assert F.filter_value("1e39", F.OPT_FLOAT) is None          # outside float32
    assert F.filter_value("12", F.OPT_INT) == "12"
    assert F.filter_value("-7", F.OPT_INT) == "-7"
    assert F.filter_value("1.5", F.OPT_INT) is None
    assert F.filter_value("3000000000", F.OPT_INT) is None
    for w in ("true", "yes", "on", "1", "TRUE", "YES", "ON"):
        assert F.filter_value(w, F.OPT_BOOL) == "1"
    for w in ("false", "no", "off", "0", "FALSE", "NO", "OFF"):
        assert F.filter_value(w, F.OPT_BOOL) == "0"
    assert F.filter_value("True", F.OPT_BOOL) is None
    assert F.filter_value("1 2 3", F.OPT_VECTOR) == "1 2 3"
    assert F.filter_value("1 2", F.OPT_VECTOR) is None
    assert F.filter_value("", F.OPT_STRING) is None


def test_env_errors_stop_reading():
    e, ok = _env("Bvh {\n method ploc\n bogus 1\n plocRadius 9\n}\n")
    assert not ok and e.get("Bvh.method") == "ploc" and e.get("Bvh.plocRadius") == 1
    e, ok = _env("Bvh {\n plocRadius many\n}\n")
    assert not ok and e.get("Bvh.plocRadius") == 1
    e, ok = _env("Bvh {\n plocRadius\n 5\n}\n")           # the value must be on the key's line
    assert not ok
    e, ok = _env("}\n")
    assert not ok
    e, ok = _env("Camera {\n position 1 2\n}\n")
    assert not ok
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert not F.Environment().read_env_file("/nonexistent/x.env")


def test_env_float_values_default_is_truncated():
    # getFloatValues returns the default converted with toInt (Environment.cpp:230).
    e = F.Environment()
    assert e.get_values("Bvh.ct") == [3.0]
    assert e.get_values("Renderer.aoRadius") == [0.0]      # "0.01" does not parse as an int
    e, ok = _env("Renderer {\n aoRadius 0.5\n aoRadius 0.25\n}\n")
    assert ok and e.get_values("Renderer.aoRadius") == [0.5, 0.25]


def test_env_to_build_params():
    pytest.importorskip("torch")
    from hippie_amd import core
    e, ok = _env(ENV_TEXT)
    assert ok
    p = core.BuildParams.from_env(e)
    assert p.method == core.METHODS["ploc"]
    assert p.ploc_radius == 25
    assert p.adaptive_leaf_size == 0
    assert p.ct == 3.5 and p.ci == 2.0
    assert p.morton_bits == 60
    assert p.layout == core.LAYOUTS["bin"]
    with pytest.raises(ValueError):
        core.BuildParams.from_env({"Bvh.layout": "hex"})


@pytest.mark.skipif(not os.path.exists("/root/reference/bin/default.env"), reason="reference tree not present")
def test_env_reads_the_reference_default_env():
    e = F.Environment()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert e.read_env_file("/root/reference/bin/default.env")
    assert e.get("Bvh.method") == "hlbvh"
    assert e.get("Bvh.plocRadius") == 100
    assert e.get("Bvh.atrIterations") == 25
    assert e.get("Renderer.rayType") == "path"
    assert e.get("Camera.fieldOfView") == 45.0


def _scene(rng, nv=7, nt=5, nm=2):
    v = rng.random((nv, 3), dtype=np.float32) * 4 - 1
    t = rng.integers(0, nv, (nt, 3), dtype=np.int32)
    m = np.zeros(nm, F.MATERIAL_DTYPE)
    m["tex_index"] = [-1, 0][:nm]
    m["shininess"] = 12.5
    m["diffuse"] = rng.random((nm, 3), dtype=np.float32)
    return F.BinScene(v, t, normals=rng.random((nv, 3), dtype=np.float32),
                      tex_coords=rng.random((nv, 2), dtype=np.float32),
                      mat_indices=rng.integers(0, nm, nt, dtype=np.int32), materials=m,
                      texture_filenames=["tex/wood.png", "mörtel.jpg", "", None])


def test_bin_scene_bytes_match_the_exporter_layout(tmp_path):
    rng = np.random.default_rng(3)
    s = _scene(rng)
    path = str(tmp_path / "a.obj.bin")
    F.write_bin_scene(path, s)
    raw = open(path, "rb").read()
    # The layout written by BinExporter, assembled independently here.
    expect = struct.pack("<4i", 7, 5, 2, 4) + s.bounding_box.tobytes()
    expect += s.vertices.tobytes() + s.normals.tobytes() + s.tex_coords.tobytes()
    expect += s.triangles.tobytes() + s.mat_indices.tobytes()
    for m in s.materials:
        expect += struct.pack("<if3f3f", m["tex_index"], m["shininess"], *m["diffuse"], *m["specular"])
    for name in ("tex/wood.png", "mörtel.jpg", ""):
        b = name.encode("utf-16-be")
        expect += struct.pack(">I", len(b)) + b
    expect += b"\xff\xff\xff\xff"
    assert raw == expect
    r = F.read_bin_scene(path)
    for k in ("vertices", "triangles", "normals", "tex_coords", "mat_indices", "bounding_box"):
        assert np.array_equal(getattr(r, k), getattr(s, k)), k
    assert r.materials.tobytes() == s.materials.tobytes()
    assert r.texture_filenames == ["tex/wood.png", "mörtel.jpg", "", None]


def test_bin_scene_missing_and_truncated(tmp_path):
    assert F.read_bin_scene(str(tmp_path / "none.bin")) is None
    rng = np.random.default_rng(4)
    path = str(tmp_path / "b.bin")
    F.write_bin_scene(path, _scene(rng))
    raw = open(path, "rb").read()
    for cut in (10, 40, 60, len(raw) - 3):
        open(path, "wb").write(raw[:cut])
        with pytest.raises(ValueError):
            F.read_bin_scene(path)


OBJ = """# two quads and a triangle
v 0 0 0
v 2 0 0
v 2 1 0
v 0 1 0
v 0 0 4
vt 0 0
vn 0 0 1
f 1/1/1 2/1/1 3/1/1 4/1/1
f -5//1 -1//1 -2//1
f 2 3 5
"""


def test_load_scene_writes_and_reads_the_cache(tmp_path):
    obj = tmp_path / "s.obj"
    obj.write_text(OBJ)
    s = F.load_scene(str(obj))
    assert s.triangles.tolist() == [[0, 1, 2], [0, 2, 3], [0, 4, 3], [1, 2, 4]]
    assert os.path.exists(str(obj) + ".bin")
    obj.write_text("v 0 0 0\n")                       # the cache, not the file, is read now
    s2 = F.load_scene(str(obj))
    assert np.array_equal(s2.vertices, s.vertices) and np.array_equal(s2.triangles, s.triangles)
    st = F.load_static_scene(str(obj))
    # (v - mn) / largest extent, box (0,0,0)-(2,1,4): every coordinate divided by 4.
    assert np.array_equal(st.vertices, (s.vertices * np.float32(0.25)).astype(np.float32))


def test_normalisation_rounds_each_operation():
    rng = np.random.default_rng(5)
    v = (rng.random((1000, 3)) * 3 - 7).astype(np.float32)
    box = F.vertex_box(v)
    out = F.normalize_scene_vertices(v, box)
    diag = box[3:] - box[:3]
    scale = np.float32(1) / diag.max()
    for i in range(0, 1000, 97):
        for k in range(3):
            d = np.float32(v[i, k] - box[k])
            assert out[i, k] == np.float32(d * scale)
    assert out.min() == 0.0 and out.max() <= 1.0


def test_benchmark_log_layout_and_reader(tmp_path):
    stats = {"RAY TYPE": "primary", "BVH CONSTRUCTION TIME KERNELS": 0.00123456789, "CT": 3.0, "CI": 2.0,
             "BVH COST": 1823.4412, "AVG. LEAF SIZE": 3.25, "LEAF SIZE HISTOGRAM": [0, 5, 7],
             "NUMBER OF RAYS": 1048576, "NUMBER OF PRIMARY RAYS": 1048576, "RT PERFORMANCE": 1544.46,
             "RT PERFORMANCE PRIMARY": 1544.46}
    text = F.benchmark_log_text(stats)
    head = ("RAY TYPE\nprimary\n\nBVH CONSTRUCTION TIME KERNELS\n0.00123457\n\n"
            "BVH CONSTRUCTION TIME ABSOLUTE\n3.40282e+38\n\nCT\n3\n\nCI\n2\n\nBVH COST\n1823.44\n\n"
            "AVG. LEAF SIZE\n3.25\n\nLEAF SIZE HISTOGRAM\n0\n5\n7\n\nNUMBER OF RAYS\n1048576\n\n")
    assert text.startswith(head)
    assert text.endswith("TIME TO IMAGE ABSOLUTE\n3.40282e+38\n")
    assert "RT PERFORMANCE SHADOW\n-3.40282e+38\n\n" in text
    path = F.benchmark_log_path(str(tmp_path), "conference-bin-lbvh-split0-anneal0-sort0-cam0")
    F.write_benchmark_log(path, stats)
    assert path.endswith("conference-bin-lbvh-split0-anneal0-sort0-cam0/"
                         "test_conference-bin-lbvh-split0-anneal0-sort0-cam0.log")
    assert F.parse_stat(path, "BVH COST") == pytest.approx(1823.44)
    assert F.parse_stat(path, "RT PERFORMANCE PRIMARY") == pytest.approx(1544.46)
    assert F.parse_stat(path, "NUMBER OF PRIMARY RAYS") == 1048576
    assert F.parse_stat(path, "NO SUCH STAT") is None
