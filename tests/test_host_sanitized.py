"""CPU: the library's host code under AddressSanitizer and UBSan.

bsofi_engine.hip (the planner and executor), bsofi_abi.cpp (the C and
Fortran entry points, pinned staging, the handle API) and bsofi_serve.cpp
(the serving loop) are compiled for the host only, with
-fsanitize=address,undefined, and linked against tests/host/fake_hip.cpp, a
stand-in HIP runtime whose device memory is host memory and whose kernels do
nothing.  tests/host/host_harness.cpp then drives: plans of many shapes
(n = 1 .. 1024, batches, every phase set, both BSOF schedules, one and two
lanes, graphs), the no-allocation entry point, argument errors, the
host-array paths and the device-workspace variants, the handle API with a
double finalize, and the serving loop's admission under a binding budget,
release, an oversized request and an injected device-allocation failure.
Any memory error, undefined behaviour or leak aborts the harness."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
CLANG = "/opt/rocm/llvm/bin/clang"
NM = shutil.which("nm") or "/opt/rocm/llvm/bin/llvm-nm"

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANGXX)),
                                reason="needs the ROCm compilers")


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    return r.stdout


def test_host_code_under_sanitizers(tmp_path):
    src = os.path.join(ROOT, "bsofi_amd", "csrc")
    # host-only compile: no device code, sanitizers on the host side only
    flags = ["-x", "hip", "--offload-host-only", "-O1", "-g", "-std=c++17", "-fPIC", "-fno-omit-frame-pointer",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-I" + os.path.join(ROOT, "include"), "-I" + src]
    units = [os.path.join(src, f) for f in ("bsofi_engine.hip", "bsofi_abi.cpp", "bsofi_serve.cpp")]
    units += [os.path.join(ROOT, "tests", "host", f) for f in ("fake_hip.cpp", "host_harness.cpp")]
    procs, objs = [], []
    for u in units:
        o = str(tmp_path / (os.path.basename(u) + ".o"))
        objs.append(o)
        procs.append(subprocess.Popen([HIPCC, *flags, "-c", u, "-o", o], cwd=tmp_path, stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    for u, p in zip(units, procs):
        out, _ = p.communicate()
        assert p.returncode == 0, f"compiling {u}:\n{out[-4000:]}"
    # the host-only objects name the (absent) device binaries: define them
    syms = set()
    for ln in _run([NM, "-u", *objs[:3]], tmp_path).splitlines():
        parts = ln.split()
        if parts and parts[-1].startswith("__hip_fatbin"):
            syms.add(parts[-1])
    (tmp_path / "fatbin.c").write_text("".join(f"char {s}[64];\n" for s in sorted(syms)))
    _run([CLANG, "-c", "fatbin.c", "-o", "fatbin.o"], tmp_path)
    _run([CLANGXX, "-fsanitize=address,undefined", "-g", *objs, "fatbin.o", "-o", "harness", "-lpthread"], tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(tmp_path / "harness")], cwd=tmp_path, capture_output=True, text=True, env=env,
                       timeout=900)
    print(r.stdout)
    assert r.returncode == 0, f"harness failed:\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    assert "host harness: all checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    shutil.rmtree(tmp_path, ignore_errors=True)
