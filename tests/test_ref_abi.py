"""The reference's own entry points (include/pfg_ref.h): a C program drives
libpfg.so through fluid_*, sem_* in the reference host's call order and
checks the run against the same run through pfg_* (tests/ref_abi/drive_ref.c).
On the CPU the program must compile and link against the library; on the
GPU it must run and agree."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "ref_abi", "drive_ref.c")
PKG = os.path.join(ROOT, "platelet_fiber_gpu_amd")


def build(d):
    exe = os.path.join(d, "drive_ref")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), SRC,
                    "-o", exe, "-L", PKG, "-lpfg", "-lm", f"-Wl,-rpath,{PKG}",
                    "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    return exe


def test_reference_driver_links(pfg):
    with tempfile.TemporaryDirectory() as d:
        exe = build(d)
        if pfg.device_count() > 0:
            return  # the GPU test runs it
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 77, r.stdout + r.stderr
        assert "REF_ABI skip" in r.stdout


def test_reference_host_declarations_link(pfg, tmp_path):
    """The reference host's own declarations of the entry points (gsl_rng
    opaque, no pfg header) compile without warnings and link against
    libpfg.so; every name resolves."""
    src = os.path.join(ROOT, "tests", "ref_abi", "ref_prototypes.c")
    exe = str(tmp_path / "ref_prototypes")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", src, "-o", exe, "-L", PKG,
                    "-lpfg", f"-Wl,-rpath,{PKG}", "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "resolved 13" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_reference_driver_matches_pfg(gpu):
    with tempfile.TemporaryDirectory() as d:
        exe = build(d)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "REF_ABI ok" in r.stdout
