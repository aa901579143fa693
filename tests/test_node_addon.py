"""The N-API addon (node-gyp, Node's own headers) on the CPU: it builds, loads libbd.so, and its
constructor reports the library's validation errors as JS Errors (tests/node/create_errors.js)."""
import subprocess

import pytest

import node_addon

pytestmark = pytest.mark.skipif(not node_addon.toolchain(), reason="node / node-gyp not installed")


def test_addon_builds_and_reports_create_errors():
    node_addon.build()
    assert node_addon.ADDON.exists()
    r = subprocess.run(["node", str(node_addon.ROOT / "tests" / "node" / "create_errors.js")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout
