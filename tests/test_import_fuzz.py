"""Malformed inputs for the C++ glTF importer (include/kittlespt/model_importer.hpp)
and the C-ABI PNG decoder (kpt_png_decode_rgb8): seeded byte flips, byte
runs and truncations of valid .glb, .gltf and PNG files.  Every input must
either import / decode or be refused with an error — never crash, hang or
touch memory it does not own.  The importer runs in an application built
with AddressSanitizer and UndefinedBehaviorSanitizer (g++), so an
out-of-bounds read or write, or undefined arithmetic, ends it with a report
instead of a clean exit; under the sanitized CPU suite the decoder's library
is itself instrumented.  No reference-side assets exist; the files are
the synthetic ones of tests/test_gltf.py.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from kittlespt_amd import abi, build
import test_gltf as tg

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "gltf_import.cpp")
APP = os.path.join(HERE, "cpp", "_build", "gltf_import_san")
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _app():
    deps = [SRC, abi.LIB_PATH] + [os.path.join(build.INCLUDE, "kittlespt", f) for f in
                                  ("model_importer.hpp", "pod_types.hpp")] + [os.path.join(build.INCLUDE, "kpt_abi.h")]
    if not os.path.exists(APP) or os.path.getmtime(APP) < max(os.path.getmtime(d) for d in deps):
        build.compile_cpp_app([SRC], APP, extra=SAN)
    return APP


def _mutants(data, n, seed):
    """n seeded variants of data: byte flips, random byte runs, truncations."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        b = bytearray(data)
        kind = k % 3
        if kind == 0:
            for _ in range(int(rng.integers(1, 8))):
                i = int(rng.integers(0, len(b)))
                b[i] ^= 1 << int(rng.integers(0, 8))
        elif kind == 1:
            i = int(rng.integers(0, len(b)))
            m = int(rng.integers(1, 16))
            b[i:i + m] = rng.integers(0, 256, min(m, len(b) - i), dtype=np.uint8).tobytes()
        else:
            b = b[:int(rng.integers(0, len(b)))]
        out.append(bytes(b))
    return out


# The sanitized CPU suite preloads clang's ASan runtime into its processes;
# the application carries g++'s own, so it runs in the ordinary suite only.
needs_plain_env = pytest.mark.skipif(os.environ.get("KPT_UNDER_SANITIZERS") == "1",
                                     reason="the importer application is sanitized by its own build")


def _run_all(paths):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    codes = {}
    for p in paths:
        r = subprocess.run([_app(), str(p), str(p) + ".dump"], capture_output=True, text=True, timeout=60, env=env)
        assert r.returncode in (0, 3), f"{p}: exit {r.returncode}\n{r.stdout[-2000:]}{r.stderr[-4000:]}"
        codes[r.returncode] = codes.get(r.returncode, 0) + 1
    return codes


@needs_plain_env
def test_importer_survives_mutated_glb(tmp_path):
    path, _, _ = tg._write(tmp_path, True)
    paths = []
    for k, m in enumerate(_mutants(path.read_bytes(), 90, 21)):
        p = tmp_path / f"m{k}.glb"
        p.write_bytes(m)
        paths.append(p)
    codes = _run_all(paths)
    assert codes.get(3, 0) > 0          # most mutants are refused, with a reason


@needs_plain_env
def test_importer_survives_mutated_gltf_json(tmp_path):
    path, _ = tg._rich_files(tmp_path)
    text = path.read_bytes()
    paths = []
    for k, m in enumerate(_mutants(text, 90, 22)):
        p = tmp_path / f"j{k}.gltf"
        p.write_bytes(m)
        paths.append(p)
    # structural edits the byte mutants rarely make: wrong types and huge or
    # negative numbers in index, count, offset and stride fields
    doc = json.loads(text)
    edits = [("accessors", 0, "count", 1e30), ("accessors", 1, "byteOffset", -4), ("bufferViews", 0, "byteStride", 3),
             ("bufferViews", 1, "byteLength", 1e19), ("accessors", 3, "bufferView", "x"),
             ("accessors", 4, "componentType", 5124), ("bufferViews", 2, "buffer", 7)]
    for k, (arr, i, key, val) in enumerate(edits):
        d = json.loads(text)
        d[arr][i][key] = val
        p = tmp_path / f"e{k}.gltf"
        p.write_text(json.dumps(d))
        paths.append(p)
    d = json.loads(text)
    d["nodes"][0]["mesh"] = 1e300
    (tmp_path / "huge_index.gltf").write_text(json.dumps(d))
    paths.append(tmp_path / "huge_index.gltf")
    (tmp_path / "deep.gltf").write_text("[" * 100000)
    paths.append(tmp_path / "deep.gltf")
    codes = _run_all(paths)
    assert codes.get(3, 0) >= len(edits)
    assert doc["asset"]["version"] == "2.0"


def test_png_decoder_survives_mutated_images():
    L = abi.lib()
    rng = np.random.default_rng(23)
    base = [tg._encode_png(rng.integers(0, 256, (9, 11, c), dtype=np.uint8), t, [0, 1, 2, 3, 4])
            for t, c in ((2, 3), (6, 4), (0, 1), (4, 2))]
    pal = rng.integers(0, 256, (5, 3), dtype=np.uint8)
    base.append(tg._encode_png(rng.integers(0, 5, (7, 6, 1), dtype=np.uint8), 3, [1, 4], palette=pal))
    seen = set()
    for k, png in enumerate(base):
        for m in _mutants(png, 80, 30 + k):
            buf = np.frombuffer(m, np.uint8) if m else np.zeros(1, np.uint8)
            w, h = ctypes.c_int32(), ctypes.c_int32()
            rc = L.kpt_png_decode_rgb8(buf.ctypes.data, len(m), ctypes.byref(w), ctypes.byref(h), None, 0)
            seen.add(rc)
            if rc != abi.KPT_OK:
                continue
            assert 0 < w.value and 0 < h.value and w.value * h.value <= (1 << 28)
            if w.value * h.value > 1 << 20:     # a mutated header asking for a huge image: refuse the buffer
                rc2 = L.kpt_png_decode_rgb8(buf.ctypes.data, len(m), ctypes.byref(w), ctypes.byref(h),
                                            buf.ctypes.data, 0)
                assert rc2 == abi.KPT_ERR_INVALID_ARG
                continue
            out = np.zeros(w.value * h.value * 3, np.uint8)
            rc = L.kpt_png_decode_rgb8(buf.ctypes.data, len(m), ctypes.byref(w), ctypes.byref(h), out.ctypes.data,
                                       out.size)
            seen.add(rc)
    assert abi.KPT_OK in seen and abi.KPT_ERR_INVALID_ARG in seen
