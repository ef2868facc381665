"""The generic Stockham pass kernels (used for line lengths without compiled
four-step kernels) against the oracle, selected with QBG_GENERIC_FFT=1 in a
child process so the setting cannot leak into other tests."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


def test_generic_kernels_parity():
    env = dict(os.environ, QBG_GENERIC_FFT="1")
    cmd = [sys.executable, "-u", "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
           str(ROOT / "tests" / "test_gpu_fft.py"), "-k", "fp64 and (10 or 11 or 0-16)"]
    r = subprocess.run(cmd, env=env, cwd=str(ROOT), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0
