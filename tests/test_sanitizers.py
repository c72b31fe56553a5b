"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

  * The oracle (oracle/*.cpp) built with -fsanitize=address,undefined runs
    the oracle test files in a child process that preloads the sanitizer
    runtimes (PALZ_ORACLE_LIB selects that build).
  * The OBJ loader (palzilla_amd/csrc/scene_io.hip: palz_mesh_load_obj and
    its copy/destroy calls), which parses files it is handed, built with the
    sanitizers on its host code (-Xarch_host), loads the reference's table
    mesh and a set of malformed files made by mutating it: every load either
    succeeds or returns an error, and no sanitizer report appears.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "palzilla_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _runtimes():
    libs = [subprocess.run(["gcc", f"-print-file-name={n}"], capture_output=True, text=True).stdout.strip()
            for n in ("libasan.so", "libubsan.so")]
    if not all(os.path.isabs(p) and os.path.exists(p) for p in libs):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    return libs


def _newer(target, sources):
    return os.path.exists(target) and all(os.path.getmtime(target) >= os.path.getmtime(s) for s in sources)


def _clean(out: str) -> bool:
    return "ERROR: AddressSanitizer" not in out and "runtime error:" not in out


def test_oracle_under_asan_and_ubsan():
    libs = _runtimes()
    sys.path.insert(0, ROOT)
    from oracle import build as ob
    srcs = [os.path.join(ROOT, "oracle", s) for s in ob.SOURCES] + [os.path.join(ROOT, "oracle", "oracle.hpp")]
    lib = ob.SANITIZED if _newer(ob.SANITIZED, srcs) else ob.build_sanitized()
    syms = subprocess.run(["nm", "-D", lib], capture_output=True, text=True).stdout
    assert "__asan_init" in syms and "__ubsan_handle" in syms
    env = dict(os.environ, LD_PRELOAD=" ".join(libs), PALZ_ORACLE_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    files = ["test_oracle_lbvh.py", "test_oracle_traversal.py", "test_oracle_occlusion.py",
             "test_oracle_interval.py", "test_oracle_caustics.py", "test_oracle_pathcuts.py",
             "test_oracle_primary.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        *[os.path.join(HERE, f) for f in files]], capture_output=True, text=True, env=env,
                       cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert _clean(out), out[-4000:]


def _mutants(text: str, rng: np.random.Generator, n: int):
    lines = text.splitlines()
    faces = [i for i, l in enumerate(lines) if l.startswith("f ")]
    verts = [i for i, l in enumerate(lines) if l.startswith("v ")]
    out = []
    for k in range(n):
        ls = list(lines)
        kind = k % 6
        if kind == 0:  # truncated mid-line
            cut = int(rng.integers(1, len(text)))
            out.append(text[:cut])
            continue
        i = int(rng.choice(faces))
        if kind == 1:  # a corner index past the end, zero or far negative
            ls[i] = "f 1 2 " + str(int(rng.choice([0, 10 ** 6, -10 ** 6, 2 ** 40])))
        elif kind == 2:  # malformed corners
            ls[i] = rng.choice(["f 1/", "f //", "f 1//", "f a b c", "f 1/2/3/4 5 6", "f", "f 1 2"])
        elif kind == 3:  # numbers out of range or not numbers
            j = int(rng.choice(verts))
            ls[j] = rng.choice(["v 1e999 0 0", "v nan inf -inf", "v 1 2", "v", "v 0x1p3 - +"])
        elif kind == 4:  # normals that do not exist
            ls[i] = "f 1//99999 2//1 3//-99999"
        else:  # a very long face
            ls[i] = "f " + " ".join(str(1 + (j % 500)) for j in range(5000))
        out.append("\n".join(ls) + "\n")
    return out


def test_obj_loader_under_asan_and_ubsan(tmp_path):
    _runtimes()
    exe = tmp_path / "obj_driver"
    cmd = [HIPCC, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "--offload-arch=gfx950",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "scene_io.hip"), os.path.join(CSRC, "runtime.hip"),
           "-x", "c++", os.path.join(HERE, "sanitize", "obj_driver.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    if r.returncode != 0:
        pytest.fail(r.stderr[-3000:])
    table = os.path.join(HERE, "golden", "table.obj")
    text = open(table).read()
    rng = np.random.default_rng(11)
    paths = [table]
    for k, m in enumerate(_mutants(text, rng, 48)):
        p = tmp_path / f"m{k}.obj"
        p.write_text(m)
        paths.append(str(p))
    paths.append(str(tmp_path / "missing.obj"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe), *paths], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and _clean(out), out[-4000:]
    rows = [l.split() for l in r.stdout.splitlines()]
    assert len(rows) == len(paths)
    assert rows[0][1:] == ["0", "3068"], "the table mesh loads with its 3,068 triangles"
    assert rows[-1][1] != "0", "a missing file is an error"
    assert sum(row[1] != "0" for row in rows) >= 10, "malformed files must be rejected"
