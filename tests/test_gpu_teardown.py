"""Process exit with library objects still alive: a model kept by a stored
traceback (as pytest keeps the last failed test's frames in
sys.last_traceback) after its context was closed, and a server that was
never closed with a request still queued.  The interpreter must exit cleanly
(return code 0, no allocator abort); the binding closes servers, models and
contexts in that order at the start of shutdown."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import sys
sys.path.insert(0, {pkg!r})
import numpy as np
from elpaso_amd import Context, GpuModel, Server
from elpaso_amd import meshgen as MG

ctx = Context(0)
g = GpuModel(ctx, MG.acoustic_cavity(6, 5, 4))
g.assemble()
X, rel, it, st = g.sweep(np.array([100.0, 400.0]), block=4)
assert (rel <= 1e-8).all()
def failing(model):
    assert model.n < 0, "planted failure"
try:
    failing(g)
except AssertionError:
    sys.last_type, sys.last_value, sys.last_traceback = sys.exc_info()
sv = Server(ctx, 1 << 30, max_slots=1)
for k in range(3):
    sv.submit(GpuModel(ctx, MG.acoustic_cavity(8, 6, 5)), np.array([120.0, 480.0]), block=4)
ctx.close()
del g
print("exiting with a model in a traceback and an open server", flush=True)
'''


def test_exit_with_live_objects_is_clean():
    script = SCRIPT.format(pkg=os.path.join(ROOT, "elpaso-core_amd"))
    print("[teardown] subprocess exit with live objects", flush=True)
    p = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=100)
    assert p.returncode == 0, (p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    assert "exiting with a model" in p.stdout
    assert "corrupted" not in p.stderr and "Aborted" not in p.stderr, p.stderr[-2000:]
