"""The frame transport's WebSocket layer (node/ws.js) on the CPU: text and
binary messages of every length class, masking and fragmentation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_ws_roundtrip():
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "ws_roundtrip.mjs")],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "WS OK" in r.stdout
