"""AddressSanitizer + UndefinedBehaviorSanitizer over the host code (CPU only).

tests/sanitize/san_main.cpp links the library's host translation units
(chorus_api.cpp, host_ops.cpp, tecplot.cpp) and the C oracle, all built with
-fsanitize=address,undefined (no recovery: the first report fails the run),
with link-time stand-ins for the device launchers (tests/sanitize/
device_stubs.cpp).  On golden cases it runs the oracle's compresid and time
steps (checked bit for bit against the reference's fixtures), the library's
host topology (checked against the oracle's), its plot-node topology, and
chorus_init's host set-up, which ends in CHORUS_ERR_HIP without a GPU.
"""
import os
import shutil
import struct
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import REPO, golden_params, load_golden

CASES = ["ot_n8_N4_jitter_av", "ot_n5_N6_free", "visc_n6_N5_free_av", "rotor_n16_N6_free_av_g14", "loop_n16x8_N5"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]
BTYPES = {"CYCL": 1, "FREE": 2}

pytestmark = pytest.mark.skipif(os.path.exists("/dev/kfd"),
                                reason="CPU-only build: on a GPU machine chorus_init would reach the device path")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("gcc/g++ not available")
    d = tmp_path_factory.mktemp("san")
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + os.path.join(REPO, "chorus-mhd_amd", "csrc"),
           "-I" + os.path.join(REPO, "oracle"), "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    csrc = os.path.join(REPO, "chorus-mhd_amd", "csrc")
    jobs = [["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math"] + SAN +
            ["-c", os.path.join(REPO, "oracle", "chorus_oracle.c"), "-o", str(d / "oracle.o")]]
    for f in ("host_ops.cpp", "chorus_api.cpp", "tecplot.cpp"):
        jobs.append(["g++", "-std=c++17", "-O1", "-ffp-contract=off"] + SAN + inc +
                    ["-c", os.path.join(csrc, f), "-o", str(d / (f[:-4] + ".o"))])
    for f in ("san_main.cpp", "device_stubs.cpp"):
        jobs.append(["g++", "-std=c++17", "-O1"] + SAN + inc +
                    ["-c", os.path.join(REPO, "tests", "sanitize", f), "-o", str(d / (f[:-4] + ".o"))])
    with ThreadPoolExecutor(max_workers=6) as ex:
        res = list(ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), jobs))
    for c, r in zip(jobs, res):
        assert r.returncode == 0, " ".join(c) + "\n" + r.stderr[-3000:]
    objs = sorted(str(p) for p in d.glob("*.o"))
    exe = str(d / "san_main")
    r = subprocess.run(["g++"] + SAN + ["-o", exe] + objs + ["-L/opt/rocm/lib", "-lamdhip64",
                                                            "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("name", CASES)
def test_host_code_under_sanitizers(driver, name, tmp_path):
    g = load_golden(name)
    mesh, N = g["mesh"], int(g["N"])
    p = golden_params(g)
    steps, nre = int(g["steps"]), 2
    nbnd = mesh.bnd.shape[0]
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        f.write(struct.pack("<8i", mesh.ncell, mesh.nvert, N, p["artiv"], p["k_stage"], steps, nbnd, nre))
        f.write(struct.pack("<6d", p["mu"], p["eta0"], p["gam"], float(g["dt"]), mesh.dxcycl, mesh.dycycl))
        f.write(np.ascontiguousarray(mesh.xv, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(mesh.yv, dtype="<f8").tobytes())
        f.write(np.asfortranarray(mesh.cells, dtype="<i4").ravel(order="F").tobytes())
        f.write(np.asfortranarray(mesh.bnd, dtype="<i4").ravel(order="F").tobytes())
        f.write(np.array([BTYPES.get(t, 0) for t in mesh.bnd_type], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(g["q0"], dtype="<f8").tobytes())
    out = tmp_path / "out.bin"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([driver, str(case), str(out)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    raw = open(out, "rb").read()
    n = 9 * N * N * mesh.ncell
    resid = np.frombuffer(raw, "<f8", n, 0).reshape(g["resid"].shape)
    qend = np.frombuffer(raw, "<f8", n, 8 * n).reshape(g["qend"].shape)
    nface, nnode, ncellp, init_rc = struct.unpack_from("<4i", raw, 16 * n)
    assert np.array_equal(resid, g["resid"])
    assert np.array_equal(qend, g["qend"])
    assert nface == int(g["topo"][2])
    assert ncellp == mesh.ncell * nre * nre and nnode > 0
    assert init_rc == -2          # CHORUS_ERR_HIP: the host set-up ran, no device
