"""The C++ drop-in (include/feec/..., libsimplefem_amd_cxx.so): host-side
classes checked by a caller written like the reference's own tests."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "simplefem_amd", "_build", "dropin_test")


def test_dropin_host_classes():
    assert os.path.exists(BIN), "build first: python -m simplefem_amd.build"
    proc = subprocess.run([BIN, "cpu"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failure(s)" in proc.stdout
