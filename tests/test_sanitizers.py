"""The CPU suite under AddressSanitizer and UndefinedBehaviorSanitizer.

The host code of libkpt_hip.so (the C ABI, the host BVH builders, image
decoding, the scene and context plumbing in the host pass of the .hip
sources) and the oracle are rebuilt with -fsanitize=address,undefined by
clang (the compiler hipcc drives for the host pass; device code is built
as usual) into kittlespt_amd/_lib/san/ and oracle/_build/san/.  The rest of
the CPU suite then runs in a child process against those builds, with the
ASan runtime preloaded into Python.  Any report ends the child (ASan halts
on the first error, UBSan does not recover), which fails this test.  The
builds are redone only when a source is newer.  Runs by hand too:
`python tests/test_sanitizers.py`.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kittlespt_amd import build  # noqa: E402

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_SAN = os.path.join(ROOT, "kittlespt_amd", "_lib", "san", "libkpt_hip.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SRCS = [os.path.join(ORACLE_DIR, f) for f in ("kpt_oracle.cpp", "kpt_oracle_post.cpp", "kpt_oracle_cache.cpp")]
ORACLE_SAN = os.path.join(ORACLE_DIR, "_build", "san", "libkpt_oracle.so")
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer", "-shared-libsan"]


def _stale(out, sources):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(s) > t for s in sources)


def _headers(*dirs):
    out = []
    for d in dirs:
        out += glob.glob(os.path.join(d, "**", "*.h"), recursive=True)
    return out


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed:\n{r.stdout[-4000:]}{r.stderr[-4000:]}")


def build_libraries():
    """Sanitized builds of the product library (host pass) and the oracle."""
    hip_deps = build.SOURCES + _headers(build.CSRC, build.INCLUDE)
    if _stale(HIP_SAN, hip_deps):
        os.makedirs(os.path.dirname(HIP_SAN), exist_ok=True)
        flags = [f for f in build.HIPCC_FLAGS if f != "-O3"] + ["-O1", "-g"]
        for f in SAN:                      # host pass only
            flags += ["-Xarch_host", f]
        _run([build.hipcc()] + flags + ["-o", HIP_SAN] + build.SOURCES + ["-lpthread", "-lz"])
    oracle_deps = ORACLE_SRCS + _headers(ORACLE_DIR, build.INCLUDE, os.path.join(build.CSRC, "common"))
    if _stale(ORACLE_SAN, oracle_deps):
        os.makedirs(os.path.dirname(ORACLE_SAN), exist_ok=True)
        # the oracle's own IEEE flags (oracle/Makefile), built by clang
        _run([os.path.join(ROCM, "llvm", "bin", "clang++"), "-std=c++17", "-O1", "-g", "-fPIC", "-shared",
              "-ffp-contract=off", "-fno-fast-math", "-fno-strict-aliasing"] + SAN +
             ["-o", ORACLE_SAN] + ORACLE_SRCS + ["-lpthread"])


def asan_runtime():
    found = sorted(glob.glob(os.path.join(ROCM, "llvm", "lib", "clang", "*", "lib", "linux",
                                          "libclang_rt.asan-x86_64.so")))
    if not found:
        raise RuntimeError("clang's ASan runtime (libclang_rt.asan-x86_64.so) not found")
    return found[-1]


def run_suite():
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan_runtime(), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", KPT_LIB_PATH=HIP_SAN,
               KPT_ORACLE_LIB=ORACLE_SAN, KPT_UNDER_SANITIZERS="1")
    return subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-x", "-q",
                           "-m", "not gpu", "-p", "no:cacheprovider"],
                          cwd=ROOT, env=env, capture_output=True, text=True)


@pytest.mark.skipif(os.environ.get("KPT_UNDER_SANITIZERS") == "1", reason="this is the sanitized run itself")
def test_cpu_suite_under_asan_and_ubsan():
    build_libraries()
    r = run_suite()
    tail = (r.stdout + r.stderr)[-6000:]
    print(tail)
    assert r.returncode == 0, tail
    assert "runtime error" not in r.stdout + r.stderr
    assert "AddressSanitizer" not in r.stdout + r.stderr


if __name__ == "__main__":
    build_libraries()
    res = run_suite()
    sys.stdout.write(res.stdout[-6000:] + res.stderr[-6000:])
    sys.exit(res.returncode)
