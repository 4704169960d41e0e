"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

A copy of libgdock_hip.so is built with both sanitizers on its host side only
(each -fsanitize after -Xarch_host: the gfx950 code is the ordinary build),
and tests/sanitize/host_driver.cpp, built the same way, drives the host
code: PDB parsing of every fixture and of damaged text, record formatting,
ligand positioning, the host transform, pose records over the host thread
pool, the project stream, and whole GA runs of the host driver with
synthetic fitness (Hall of Fame past its capacity, early stop, final breed,
model selection).  Any report aborts the driver.  No GPU is used."""
import gzip
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    from gdock_amd import build as b
    if shutil.which(b.hipcc()) is None and not Path(b.hipcc()).exists():
        pytest.skip("hipcc is not available")
    d = tmp_path_factory.mktemp("sanitized")
    lib = b.build(verbose=False, lib=d / "libgdock_hip.so", objdir=d / "obj", extra_flags=SAN)
    exe = d / "host_driver"
    res = subprocess.run([b.hipcc(), "-O1", "-g", "-std=c++17", *SAN, str(ROOT / "tests/sanitize/host_driver.cpp"),
                          "-L", str(d), "-lgdock_hip", f"-Wl,-rpath,{d}", "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return lib, exe, d


def test_host_code_clean_under_asan_and_ubsan(sanitized):
    _, exe, d = sanitized
    data = ROOT / "tests/golden/data"
    poses = d / "1h590B_0A_poses.pdb"
    poses.write_bytes(gzip.decompress((data / "1h590B_0A_poses.pdb.gz").read_bytes()))
    files = [data / "2oob_A.pdb", data / "2oob_B.pdb", data / "2oob.pdb", data / "1h590A_0A.pdb", poses]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([str(exe), *map(str, files)], capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, (res.stdout[-2000:], res.stderr[-6000:])
    assert "host driver ok" in res.stdout
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr
