"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The scene model and the importers (luz_amd/csrc/scene.cpp, ingest.cpp) parse
files a user hands in: .luz/.luzbin, glTF/GLB and OBJ/MTL.  They are built
here with g++ -fsanitize=address,undefined together with a small driver
(tests/sanitize/ingest_driver.cpp) and run over the committed assets and over
corrupted copies of them: truncations at many offsets and seeded byte flips.
Every file must either load (and survive a save/reload round trip) or be
rejected with an error message; a sanitizer report fails the test.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
ASSETS = ROOT / "tests" / "golden" / "assets"
SOURCES = [ROOT / "luz_amd" / "csrc" / "scene.cpp", ROOT / "luz_amd" / "csrc" / "ingest.cpp",
           ROOT / "tests" / "sanitize" / "ingest_driver.cpp"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("san") / "ingest_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", str(ROOT / "include"), *map(str, SOURCES), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def _run(driver, scratch, files):
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=86",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1:exitcode=87", "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(driver), str(scratch), *map(str, files)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    return r.stdout


def test_assets_load_and_round_trip(driver, tmp_path):
    files = sorted(p for p in ASSETS.iterdir() if p.suffix in (".glb", ".obj", ".luz"))
    out = _run(driver, tmp_path, files)
    print(out)
    for p in files:
        if p.suffix in (".glb", ".obj"):
            assert f"{p}: loaded" in out and "same round trip" in out.split(f"{p}: loaded")[1].split("\n")[0]


def test_corrupted_assets_are_rejected_or_loaded_cleanly(driver, tmp_path):
    rng = np.random.default_rng(7)
    files = []
    for p in sorted(ASSETS.iterdir()):
        if p.suffix not in (".glb", ".obj", ".luz"):
            continue
        data = p.read_bytes()
        cuts = sorted(set(np.linspace(0, len(data) - 1, 24).astype(int).tolist()))
        for k, c in enumerate(cuts):
            f = tmp_path / f"{p.stem}_cut{k}{p.suffix}"
            f.write_bytes(data[:c])
            files.append(f)
        for k in range(24):
            b = bytearray(data)
            for pos in rng.integers(0, len(b), 1 + k % 6):
                b[int(pos)] = int(rng.integers(0, 256))
            f = tmp_path / f"{p.stem}_flip{k}{p.suffix}"
            f.write_bytes(bytes(b))
            files.append(f)
    out = _run(driver, tmp_path, files)
    assert out.count("\n") == len(files)
    assert "rejected" in out
