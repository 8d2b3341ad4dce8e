"""hippie's on-disk and wire formats, so its configuration files, scene caches
and benchmark sweeps work unchanged against this library.

* ``Environment`` -- the ``.env`` block format and its typed option registry
  (src/hippie/environment/Environment.cpp:10-465, AppEnvironment.cpp:10-91).
* ``read_bin_scene`` / ``write_bin_scene`` -- the ``<scene>.bin`` cache written
  beside an imported scene (src/hippie/util/BinImporter.cpp:16-86,
  BinExporter.cpp:15-62), and ``load_static_scene`` -- the loader that reads the
  cache or imports the scene, then normalises it (rt/scene/SceneLoader.cpp:20-153).
* ``write_benchmark_log`` / ``parse_stat`` -- the static benchmark's result file
  (src/hippie/benchmark/Benchmark.cpp:73-127), in the layout the reference's
  sweep scripts read (test/Comparison/table-csv.py:24-34).

All of this is host-side I/O; nothing here touches the device.
"""
import math
import os
import re
import struct
import warnings

import numpy as np

OPT_INT, OPT_FLOAT, OPT_BOOL, OPT_VECTOR, OPT_STRING = range(5)

# AppEnvironment::registerOptions (AppEnvironment.cpp:10-91): name, default, type.
APP_OPTIONS = [
    ("Application.mode", "interactive", OPT_STRING),
    ("Application.stats", "1", OPT_BOOL),
    ("Benchmark.output", "default", OPT_STRING),
    ("Benchmark.images", "0", OPT_BOOL),
    ("Resolution.width", "1024", OPT_INT),
    ("Resolution.height", "768", OPT_INT),
    ("Scene.filename", None, OPT_STRING),
    ("Scene.filefilter", None, OPT_STRING),
    ("Scene.environment", None, OPT_STRING),
    ("Scene.light", "5.0 -2.0 1.0", OPT_VECTOR),
    ("Renderer.numberOfPrimarySamples", "1", OPT_INT),
    ("Renderer.numberOfAOSamples", "8", OPT_INT),
    ("Renderer.numberOfShadowSamples", "2", OPT_INT),
    ("Renderer.aoRadius", "0.01", OPT_FLOAT),
    ("Renderer.shadowRadius", "0.001", OPT_FLOAT),
    ("Renderer.rayType", "primary", OPT_STRING),
    ("Renderer.nodeSizeThreshold", "5000", OPT_INT),
    ("Renderer.thermalThreshold", "250", OPT_INT),
    ("Renderer.recursionDepth", "1", OPT_INT),
    ("Renderer.keyValue", "0.6", OPT_FLOAT),
    ("Renderer.whitePoint", "2.0", OPT_FLOAT),
    ("Renderer.russianRoulette", "0", OPT_BOOL),
    ("Renderer.sortShadowRays", "0", OPT_BOOL),
    ("Renderer.sortAORays", "0", OPT_BOOL),
    ("Renderer.sortPathRays", "0", OPT_BOOL),
    ("Renderer.shadowRayLength", "0.0", OPT_FLOAT),
    ("Renderer.aoRayLength", "0.25", OPT_FLOAT),
    ("Renderer.pathRayLength", "0.25", OPT_FLOAT),
    ("Renderer.shadowMortonCodeBits", "32", OPT_INT),
    ("Renderer.aoMortonCodeBits", "32", OPT_INT),
    ("Renderer.pathMortonCodeBits", "32", OPT_INT),
    ("Renderer.shadowMortonCodeMethod", "aila", OPT_STRING),
    ("Renderer.aoMortonCodeMethod", "aila", OPT_STRING),
    ("Renderer.pathMortonCodeMethod", "aila", OPT_STRING),
    ("Bvh.layout", "bin", OPT_STRING),
    ("Bvh.ct", "3.0", OPT_FLOAT),
    ("Bvh.ci", "2.0", OPT_FLOAT),
    ("Bvh.maxLeafSize", "8", OPT_INT),
    ("Bvh.adaptiveLeafSize", "1", OPT_BOOL),
    ("Bvh.presplitting", "0", OPT_BOOL),
    ("Bvh.lbvhMortonCodeBits", "60", OPT_INT),
    ("Bvh.hlbvhMortonCodeBits", "60", OPT_INT),
    ("Bvh.hlbvhMortonCodeSAHBits", "15", OPT_INT),
    ("Bvh.trIterations", "2", OPT_INT),
    ("Bvh.trMortonCodeBits", "60", OPT_INT),
    ("Bvh.atrIterations", "2", OPT_INT),
    ("Bvh.atrMortonCodeBits", "60", OPT_INT),
    ("Bvh.insertionMortonCodeBits", "60", OPT_INT),
    ("Bvh.insertionMod", "1", OPT_INT),
    ("Bvh.insertionSbvh", "0", OPT_BOOL),
    ("Bvh.plocRadius", "1", OPT_INT),
    ("Bvh.plocMaxRadius", "512", OPT_INT),
    ("Bvh.plocMortonCodeBits", "60", OPT_INT),
    ("Bvh.plocAdaptive", "0", OPT_BOOL),
    ("Bvh.sbvhAlpha", "0.00001", OPT_FLOAT),
    ("Bvh.presplitterBeta", "0.1", OPT_FLOAT),
    ("Bvh.method", "lbvh", OPT_STRING),
    ("Bvh.update", "refit", OPT_STRING),
    ("Camera.position", "0.0 0.0 0.0", OPT_VECTOR),
    ("Camera.direction", "0.0 0.0 -1.0", OPT_VECTOR),
    ("Camera.wheelAngle", "0.0", OPT_FLOAT),
    ("Camera.nearPlane", "0.001", OPT_FLOAT),
    ("Camera.farPlane", "3.0", OPT_FLOAT),
    ("Camera.fieldOfView", "45.0", OPT_FLOAT),
    ("Camera.step", "0.025", OPT_FLOAT),
    ("Animation.frameRate", "30.0", OPT_FLOAT),
    ("Animation.loop", "1", OPT_BOOL),
    ("Animation.pause", "0", OPT_BOOL),
    ("Screenshots.directory", "screenshots", OPT_STRING),
]

_TRUE = ("true", "yes", "on", "1", "TRUE", "YES", "ON")
_FALSE = ("false", "no", "off", "0", "FALSE", "NO", "OFF")
_WS = re.compile(r"[ \n\t]")
_INT_RE = re.compile(r"^\s*[+-]?\d+\s*$")


def _number(x):
    """Decimal text of a float as Qt writes it (QString::number: %.6g)."""
    return "%.6g" % x


def _to_float32(text):
    """A float option's value, or None (QString::toFloat fails: not a number,
    or a finite number outside the float32 range)."""
    try:
        v = float(text)
    except ValueError:
        return None
    # Round to nearest float32 overflows from half an ulp above the largest float.
    if math.isfinite(v) and abs(v) >= 3.4028235677973366e38:
        return None
    return float(np.float32(v))


def filter_value(text, opt_type):
    """Environment::filterValue (Environment.cpp:10-48): the canonical text of
    a value, or None if it does not parse as the option's type."""
    if opt_type == OPT_INT:
        if not _INT_RE.match(text):
            return None
        v = int(text)
        if not -2**31 <= v < 2**31:
            return None
        return str(v)
    if opt_type == OPT_FLOAT:
        v = _to_float32(text)
        return None if v is None else _number(v)
    if opt_type == OPT_BOOL:
        if text in _TRUE:
            return "1"
        if text in _FALSE:
            return "0"
        return None
    if opt_type == OPT_VECTOR:
        parts = _WS.split(text)
        if len(parts) != 3:
            return None
        vals = [_to_float32(p) for p in parts]
        if any(v is None for v in vals):
            return None
        return " ".join(_number(v) for v in vals)
    return text if text else None


class Option:
    def __init__(self, name, default, opt_type):
        self.name = name
        self.type = opt_type
        self.default = None if default is None else filter_value(default, opt_type)
        if default is not None and self.default is None:
            raise ValueError(f"invalid default value for option '{name}'")
        self.values = []


class Environment:
    """Typed options read from ``.env`` files (Environment.h / Environment.cpp).

    A file is a sequence of blocks ``Name { key value ... }``; ``#`` starts a
    comment; words are separated by spaces and tabs; a vector takes three
    words.  Each occurrence of a key appends a value (the getters return the
    first; the ``*_values`` getters all of them).  An unknown key, a value that
    does not parse, or an unpaired ``}`` stops reading: ``read_env_file``
    warns and returns False, keeping the values read so far, as the reference
    does.
    """

    def __init__(self, options=APP_OPTIONS):
        self.options = {}
        for name, default, t in options:
            self.options[name] = Option(name, default, t)

    def register_option(self, name, default, opt_type):
        self.options[name] = Option(name, default, opt_type)

    def read_env_file(self, filename):
        """Environment::readEnvFile (Environment.cpp:309-465)."""
        if not os.path.exists(filename):
            warnings.warn(f"file '{filename}' does not exist")
            return False
        with open(filename, "r", encoding="utf-8", errors="replace") as f:
            return self.read_env_text(f.read(), filename)

    def read_env_text(self, text, filename="<text>"):
        block_name, last, block = "", "", False
        for line_number, line in enumerate(text.splitlines(), 1):
            c = line.find("#")
            if c != -1:
                line = line[:c]
            words = [w for w in _WS.split(line) if w]
            i = 0
            while i < len(words):
                w = words[i]
                if w == "{":
                    block_name = last
                    block = True
                elif w == "}":
                    if not block_name:
                        warnings.warn(f"unpaired }} in '{filename}' (line {line_number})")
                        return False
                    block = False
                    block_name = ""
                elif block:
                    name = block_name + "." + w
                    opt = self.options.get(name)
                    if opt is None:
                        warnings.warn(f"unknown option '{name}' in '{filename}' (line {line_number})")
                        return False
                    if opt.type == OPT_VECTOR:
                        value = (filter_value(" ".join(words[i + 1:i + 4]), OPT_VECTOR)
                                 if i + 3 < len(words) else None)
                        i += 3
                    else:
                        value = filter_value(words[i + 1], opt.type) if i + 1 < len(words) else None
                        i += 1
                    if value is None:
                        warnings.warn(f"mismatch in variable {name} in '{filename}' (line {line_number})")
                        return False
                    opt.values.append(value)
                    w = words[i] if i < len(words) else w
                last = w
                i += 1
        return True

    # -- getters (Environment.cpp:105-298) -------------------------------------

    def _texts(self, name):
        opt = self.options.get(name)
        if opt is None:
            return None, None
        if opt.values and opt.values[0]:
            return opt, list(opt.values)
        if opt.default:
            return opt, None
        return opt, []

    @staticmethod
    def _convert(text, t):
        if t == OPT_INT:
            return int(text)
        if t == OPT_FLOAT:
            return float(np.float32(float(text)))
        if t == OPT_BOOL:
            return bool(int(text))
        if t == OPT_VECTOR:
            return tuple(float(np.float32(float(x))) for x in _WS.split(text))
        return text

    def get(self, name):
        """The first value of an option, else its default, else None."""
        opt, texts = self._texts(name)
        if opt is None:
            return None
        if texts:
            return self._convert(texts[0], opt.type)
        if texts is None:
            return self._convert(opt.default, opt.type)
        return None

    def get_values(self, name):
        """Every value of an option, else [default], else None.  A float
        option's default comes back truncated to an integer, as the
        reference's getFloatValues converts it with toInt (Environment.cpp:230)."""
        opt, texts = self._texts(name)
        if opt is None:
            return None
        if texts:
            return [self._convert(x, opt.type) for x in texts]
        if texts is None:
            if opt.type == OPT_FLOAT:
                d = opt.default
                return [float(int(d))] if _INT_RE.match(d) else [0.0]
            return [self._convert(opt.default, opt.type)]
        return None

    def as_dict(self, prefix=""):
        """{name: value} of every option with a value or a default."""
        out = {}
        for name in self.options:
            if name.startswith(prefix):
                v = self.get(name)
                if v is not None:
                    out[name] = v
        return out


# ---------------------------------------------------------------------------
# <scene>.bin cache (BinImporter.cpp:16-86, BinExporter.cpp:15-62)
# ---------------------------------------------------------------------------

# struct Material { int texIndex; float shininess; Vec3f diffuse; Vec3f specular; } (rt/scene/Material.h:14-19)
MATERIAL_DTYPE = np.dtype([("tex_index", "<i4"), ("shininess", "<f4"), ("diffuse", "<f4", (3,)),
                           ("specular", "<f4", (3,))])
assert MATERIAL_DTYPE.itemsize == 32


class BinScene:
    """The arrays of a hippie scene cache: positions, normals and texture
    coordinates per vertex; vertex and material indices per triangle;
    materials; texture file names; the bounding box (mn xyz, mx xyz)."""

    def __init__(self, vertices, triangles, normals=None, tex_coords=None, mat_indices=None, materials=None,
                 texture_filenames=(), bounding_box=None):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        nv, nt = self.vertices.shape[0], self.triangles.shape[0]
        self.normals = (np.zeros((nv, 3), np.float32) if normals is None
                        else np.ascontiguousarray(normals, np.float32).reshape(nv, 3))
        self.tex_coords = (np.zeros((nv, 2), np.float32) if tex_coords is None
                           else np.ascontiguousarray(tex_coords, np.float32).reshape(nv, 2))
        self.mat_indices = (np.zeros(nt, np.int32) if mat_indices is None
                            else np.ascontiguousarray(mat_indices, np.int32).reshape(nt))
        self.materials = (np.zeros(0, MATERIAL_DTYPE) if materials is None
                          else np.ascontiguousarray(materials, MATERIAL_DTYPE).reshape(-1))
        self.texture_filenames = list(texture_filenames)
        if bounding_box is None:
            bounding_box = vertex_box(self.vertices)
        self.bounding_box = np.asarray(bounding_box, np.float32).reshape(6)

    @property
    def n_vertices(self):
        return self.vertices.shape[0]

    @property
    def n_triangles(self):
        return self.triangles.shape[0]


def vertex_box(vertices):
    """AABB of the vertices (mn xyz, mx xyz); the reset box (+max, -max) when empty."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    if v.shape[0] == 0:
        big = np.finfo(np.float32).max
        return np.array([big, big, big, -big, -big, -big], np.float32)
    return np.concatenate([v.min(axis=0), v.max(axis=0)]).astype(np.float32)


def _qstring_bytes(s):
    """QDataStream << QString (big-endian byte count, then UTF-16BE); a null
    string (None) is the count 0xFFFFFFFF."""
    if s is None:
        return struct.pack(">I", 0xFFFFFFFF)
    data = s.encode("utf-16-be")
    return struct.pack(">I", len(data)) + data


def _read_qstring(buf, pos):
    if pos + 4 > len(buf):
        raise ValueError("truncated texture file name")
    (n,) = struct.unpack_from(">I", buf, pos)
    pos += 4
    if n == 0xFFFFFFFF:
        return None, pos
    if n % 2 or pos + n > len(buf):
        raise ValueError("malformed texture file name")
    return buf[pos:pos + n].decode("utf-16-be"), pos + n


def write_bin_scene(filename, scene):
    """BinExporter::exportScene: header (vertices, triangles, materials,
    textures as int32), AABB (6 float32), vertices, normals, texture
    coordinates, vertex indices, material indices, materials, then the texture
    file names as Qt data-stream strings.  Native little-endian, as the
    reference writes its structs."""
    with open(filename, "wb") as f:
        f.write(struct.pack("<4i", scene.n_vertices, scene.n_triangles, scene.materials.shape[0],
                            len(scene.texture_filenames)))
        f.write(scene.bounding_box.astype("<f4").tobytes())
        for a in (scene.vertices, scene.normals, scene.tex_coords, scene.triangles, scene.mat_indices,
                  scene.materials):
            f.write(np.ascontiguousarray(a).tobytes())
        for name in scene.texture_filenames:
            f.write(_qstring_bytes(name))
    return True


def read_bin_scene(filename):
    """BinImporter::importScene; None when the file does not exist (the
    reference then imports the scene itself).  Unlike the reference, a file
    shorter than its header promises raises ValueError instead of leaving
    the buffers partly undefined."""
    if not os.path.exists(filename):
        return None
    with open(filename, "rb") as f:
        buf = f.read()
    if len(buf) < 40:
        raise ValueError(f"'{filename}': {len(buf)} bytes is shorter than the 40-byte header")
    nv, nt, nm, ntex = struct.unpack_from("<4i", buf, 0)
    if min(nv, nt, nm, ntex) < 0:
        raise ValueError(f"'{filename}': negative counts in the header")
    box = np.frombuffer(buf, "<f4", 6, 16).copy()
    pos = 40
    sizes = [(nv * 3, "<f4"), (nv * 3, "<f4"), (nv * 2, "<f4"), (nt * 3, "<i4"), (nt, "<i4")]
    arrays = []
    for count, dt in sizes:
        nbytes = 4 * count
        if pos + nbytes > len(buf):
            raise ValueError(f"'{filename}' is truncated")
        arrays.append(np.frombuffer(buf, dt, count, pos).copy())
        pos += nbytes
    if pos + 32 * nm > len(buf):
        raise ValueError(f"'{filename}' is truncated")
    materials = np.frombuffer(buf, MATERIAL_DTYPE, nm, pos).copy()
    pos += 32 * nm
    names = []
    for _ in range(ntex):
        s, pos = _read_qstring(buf, pos)
        names.append(s)
    v, n, tc, tri, mi = arrays
    return BinScene(v.reshape(nv, 3), tri.reshape(nt, 3), n.reshape(nv, 3), tc.reshape(nv, 2), mi, materials, names,
                    box)


def import_obj(filename):
    """A Wavefront OBJ reader for the importer step of ``load_scene``: vertex
    positions, texture coordinates and normals, and polygon faces (``v``,
    ``v/t``, ``v//n``, ``v/t/n``, negative indices) split into triangle fans.
    The reference imports through Assimp (rt/scene/AssimpImporter.cpp), which
    is not part of this library: vertices keep the file's order and faces
    share them, materials are not read (every triangle has material 0)."""
    pos, tex, nor = [], [], []
    faces = []
    with open(filename, "r", encoding="utf-8", errors="replace") as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            tag = parts[0]
            if tag == "v" and len(parts) >= 4:
                pos.append([float(x) for x in parts[1:4]])
            elif tag == "vt" and len(parts) >= 3:
                tex.append([float(x) for x in parts[1:3]])
            elif tag == "vn" and len(parts) >= 4:
                nor.append([float(x) for x in parts[1:4]])
            elif tag == "f" and len(parts) >= 4:
                idx = []
                for p in parts[1:]:
                    k = int(p.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(pos) + k)
                for j in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[j], idx[j + 1]])
    v = np.asarray(pos, np.float32).reshape(-1, 3)
    t = np.asarray(faces, np.int32).reshape(-1, 3)
    if t.size and (t.min() < 0 or t.max() >= v.shape[0]):
        raise ValueError(f"'{filename}': a face refers to a vertex that does not exist")
    return BinScene(v, t)


def load_scene(filename, importer=import_obj, cache=True):
    """SceneLoader::loadScene (SceneLoader.cpp:20-88): the ``filename.bin``
    cache if it exists, else the importer's result, which is then written to
    the cache."""
    if cache:
        s = read_bin_scene(filename + ".bin")
        if s is not None:
            return s
    s = importer(filename)
    if cache:
        write_bin_scene(filename + ".bin", s)
    return s


def normalize_scene_vertices(vertices, box):
    """SceneLoader::loadStaticScene's normalisation (SceneLoader.cpp:134-141)
    with the scene file's box: v = (v - mn) * (1 / largest extent), each
    operation rounded to float32 as the reference's Vec3f arithmetic does."""
    v = np.asarray(vertices, np.float32)
    box = np.asarray(box, np.float32)
    mn, mx = box[:3], box[3:]
    diag = (mx - mn).astype(np.float32)
    scale = np.float32(1.0) / np.float32(max(max(diag[0], diag[1]), diag[2]))
    return ((v - mn).astype(np.float32) * scale).astype(np.float32)


def load_static_scene(filename, importer=import_obj, cache=True):
    """SceneLoader::loadStaticScene: load (through the cache) and normalise
    with the file's bounding box.  Returns the BinScene with normalised
    vertices."""
    s = load_scene(filename, importer, cache)
    s.vertices = normalize_scene_vertices(s.vertices, s.bounding_box)
    return s


# ---------------------------------------------------------------------------
# Benchmark result file (Benchmark.cpp:73-127)
# ---------------------------------------------------------------------------

# Keys in the order Benchmark::testStatic writes them.
BENCHMARK_KEYS = [
    "RAY TYPE", "BVH CONSTRUCTION TIME KERNELS", "BVH CONSTRUCTION TIME ABSOLUTE", "CT", "CI", "BVH COST",
    "AVG. LEAF SIZE", "LEAF SIZE HISTOGRAM", "NUMBER OF RAYS", "NUMBER OF PRIMARY RAYS", "NUMBER OF SHADOW RAYS",
    "NUMBER OF AO RAYS", "NUMBER OF PATH RAYS", "RT PERFORMANCE", "RT PERFORMANCE PRIMARY", "RT PERFORMANCE SHADOW",
    "RT PERFORMANCE AO", "RT PERFORMANCE PATH", "RENDER TIME KERNELS", "RENDER TIME ABSOLUTE",
    "TIME TO IMAGE KERNELS", "TIME TO IMAGE ABSOLUTE",
]

MAX_FLOAT = float(np.finfo(np.float32).max)


def _text(v):
    """A value as QTextStream writes it: floats with 6 significant digits
    (SmartNotation, %g), integers and strings as they are."""
    if isinstance(v, (bool, np.bool_)):
        return str(int(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return _number(float(np.float32(v)))
    return str(v)


def benchmark_log_text(stats):
    """The text of a static benchmark result file.  ``stats`` maps the keys of
    BENCHMARK_KEYS to values (``LEAF SIZE HISTOGRAM``: a sequence of counts);
    a missing performance figure is written as the reference's initial value
    (-MAX_FLOAT for RT PERFORMANCE*, MAX_FLOAT for times and cost, 0 for ray
    counts)."""
    out = []
    for i, key in enumerate(BENCHMARK_KEYS):
        out.append(key + "\n")
        if key == "LEAF SIZE HISTOGRAM":
            for h in stats.get(key, ()):
                out.append(_text(int(h)) + "\n")
            out.append("\n")
            continue
        if key in stats:
            v = stats[key]
        elif key.startswith("RT PERFORMANCE"):
            v = -MAX_FLOAT
        elif key.startswith("NUMBER OF"):
            v = 0
        elif key == "RAY TYPE":
            v = "primary"
        else:
            v = MAX_FLOAT
        out.append(_text(v) + ("\n\n" if i + 1 < len(BENCHMARK_KEYS) else "\n"))
    return "".join(out)


def write_benchmark_log(filename, stats):
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "w") as f:
        f.write(benchmark_log_text(stats))


def benchmark_log_path(root, test_name):
    """Where the sweep scripts look for a run's log: <root>/<name>/test_<name>.log
    (test/Comparison/table-csv.py:73-74)."""
    return os.path.join(root, test_name, "test_" + test_name + ".log")


def parse_stat(filename, stat):
    """The value on the line after the first line containing ``stat``, as a
    float (the sweep scripts' reader); None when the stat is absent."""
    with open(filename) as f:
        lines = f.read().split("\n")
    for i, line in enumerate(lines):
        if stat in line:
            return float(lines[i + 1])
    return None
