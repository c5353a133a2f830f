"""CPU checks of the C++ host driver's command line (no device work)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "q-vti_amd", "qvti_shots")


@pytest.mark.skipif(not os.path.exists(EXE), reason="host driver not built")
def test_help_and_argument_errors():
    r = subprocess.run([EXE, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage:" in r.stdout
    r = subprocess.run([EXE, "--nx", "8", "--bogus", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unknown option --bogus" in r.stderr
    r = subprocess.run([EXE, "--ny", "8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "must be positive" in r.stderr
    r = subprocess.run([EXE, "--nx", "8", "--ny", "8", "--nz", "8", "--eps", "e.bin"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "go together" in r.stderr
    base = [EXE, "--nx", "8", "--ny", "8", "--nz", "8"]
    r = subprocess.run(base + ["--rtm-type", "5"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--rtm-type must be" in r.stderr
    r = subprocess.run(base + ["--rtm-type", "1", "--imaging-condition", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "partial correlation" in r.stderr
    r = subprocess.run(base + ["--checkpoint", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--checkpoint must be positive" in r.stderr
    r = subprocess.run(base + ["--snap-every", "5", "--snap-field", "9"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--snap-field in 0..8" in r.stderr
    r = subprocess.run(base + ["--media", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--media must be" in r.stderr
