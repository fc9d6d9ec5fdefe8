"""bench.py's command-line contract on the CPU (no GPU work runs)."""

from __future__ import annotations

import subprocess
import sys

from tests.conftest import ROOT


def test_more_than_one_gpu_is_refused_with_a_message():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""
