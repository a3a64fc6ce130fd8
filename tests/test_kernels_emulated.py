"""The kernels of sdhash_amd/csrc, unmodified, built for the host against a
CPU emulation of HIP (tests/hipemu) with AddressSanitizer, and run on small
parity cases (tests/emu_parity.py).  This checks every memory access of the
kernels and their barrier/wave structure on the CPU.  A second build with a
12-step warm-up forces the score fix-point, uncertified tiles and the
hand-off re-run, which real data reaches only rarely with the 64-step warm-up.
CPU only."""
import os
import platform
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(platform.machine() != "x86_64", reason="emulator fibers are x86-64")


def build(tmp, extra):
    out = str(tmp / "libsdbfx_emu.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-fPIC", "-shared", "-I", os.path.join(ROOT, "tests", "hipemu"),
                    "-I", os.path.join(ROOT, "include"), *extra, "-x", "c++",
                    os.path.join(ROOT, "sdhash_amd", "csrc", "sdbfx_host.cpp"),
                    os.path.join(ROOT, "sdhash_amd", "csrc", "sdbfx_serve.cpp"), "-o", out], check=True)
    return out


def run(lib):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run(["python3", os.path.join(ROOT, "tests", "emu_parity.py"), lib], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "compare ok" in r.stdout
    assert "serve ok" in r.stdout
    assert "admission ok" in r.stdout
    assert "noalloc ok" in r.stdout
    assert "index ok" in r.stdout


@pytest.mark.parametrize("variant", [[], ["-DSDBFX_WARM=12", "-DSDBFX_INSERT_REPLAY=1"]],
                         ids=["default", "warm12-replay"])
def test_emulated_kernels(tmp_path, variant):
    run(build(tmp_path, variant))
