"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU.

The oracle (oracle/) is the checker here and only here: tests build it if
missing and compare libgpv's HIP path against it on identical inputs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


def _ensure_built():
    if not os.path.exists(os.path.join(ROOT, "oracle", "liborc.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "gpview_amd", "libgpv.so")):
        subprocess.run([sys.executable, "-m", "gpview_amd.build"], check=True, cwd=ROOT)
    native = os.path.join(ROOT, "tests", "native")
    if not all(os.path.exists(os.path.join(native, f)) for f in ("libslabcheck.so", "libdevpredicates.so")):
        subprocess.run(["make", "-s", "-C", native], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    return o


@pytest.fixture(scope="session")
def gv():
    import gpview_amd
    return gpview_amd


@pytest.fixture(scope="session")
def ctx(gv):
    c = gv.Context()
    yield c
    c.close()


def say(msg):
    """Name the case before a GPU step (flushed, so a hung box still shows it)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def soup_of(gv, verts, faces):
    return gv.flatten(verts, faces)


def compare_results(name, got, want, n2):
    """Bit-exact comparison of the four result arrays with a readable report."""
    msgs = []
    for key in ("l1", "prefix", "bidx", "l2"):
        a, b = np.asarray(got[key]), np.asarray(want[key])
        if a.shape != b.shape:
            msgs.append(f"{key}: shape {a.shape} != {b.shape}")
            continue
        bad = np.nonzero(a != b)[0]
        if len(bad):
            i = bad[0]
            extra = ""
            if key == "l2" and n2:
                n23 = n2 ** 3
                extra = f" (boundary {i // n23}, sub {i % n23})"
            msgs.append(f"{key}: {len(bad)} differ, first at {i}{extra}: got {a[i]} want {b[i]}")
    assert not msgs, f"{name}: " + "; ".join(msgs)
