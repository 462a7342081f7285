"""empi-hip argument handling that needs no device: a non-positive
--epochs-per-batch is rejected before the configuration is read."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "empi_amd", "empi-hip")


@pytest.mark.skipif(not os.path.exists(CLI), reason="empi-hip not built")
@pytest.mark.parametrize("value", ["0", "-3", "x"])
def test_epochs_per_batch_must_be_positive(tmp_path, value):
    cfg = tmp_path / "run.cfg"
    cfg.write_text("energyError 0.01\n")
    r = subprocess.run([CLI, f"--epochs-per-batch={value}", str(cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "--epochs-per-batch needs a positive count" in r.stderr
