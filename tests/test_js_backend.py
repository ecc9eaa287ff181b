"""The Node backend (vitrum_amd/js) over the N-API addon: the Engine contract
checks of vitrum_amd/js/test_contract.js, without a device (argument and scene
validation, error classes, matrix helpers) and on device 0."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "vitrum_amd", "napi", "vitrum_hip.node")
SCRIPT = os.path.join(ROOT, "vitrum_amd", "js", "test_contract.js")

needs_node = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built N-API addon is not available")


def run(gpu):
    env = dict(os.environ, VITRUM_HIP_GPU="1" if gpu else "0")
    r = subprocess.run(["node", SCRIPT], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@needs_node
def test_contract_without_device():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible; the device variant covers this host")
    assert "contract ok (no device)" in run(False)


@needs_node
@pytest.mark.gpu
def test_contract_on_device():
    assert "contract ok (device 0)" in run(True)
