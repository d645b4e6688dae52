"""The torch bridge (spectrecon_amd.dl) on the GPU.  Its checks run in a
child process (tests/dl_checks.py) that imports torch before the library:
torch ships its own HIP runtime, and this pytest process has loaded the
library against the system one already."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_torch_projector_and_bregem(gpu):
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "dl_checks.py")], cwd=os.path.dirname(HERE),
                       capture_output=True, text=True, timeout=240)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    for name in ("autograd", "bregem_vs_oracle", "training"):
        assert f"ok {name}" in r.stdout, out[-3000:]
