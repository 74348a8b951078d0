"""The C++ Object-level drop-in (include/gpv_object.hpp) end to end.

tests/native/object_dropin is a small C++ program built against the facade
and libgpv.  It runs the sequence GPView runs per object
(src/GPView.cpp:1374-1381): ReadObject(cessna.obj) ->
CreateFlatTriangleData -> PerformVoxelization(glParam, bufferSize), which
calls SaveVoxelization.  All six .raw/.txt files must be byte-identical to
the oracle's writer for the same mesh and grid (src/Object.cpp:2934-3075).
bufferSize 1 forces the reference's capacity retry (src/Object.cpp:3088-3089,
3198-3214); -1 keeps the default capacity.
"""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT, say

pytestmark = pytest.mark.gpu

FILES = ("VoxelConfig.txt", "Level1InOut.raw", "Level1BoundaryPrefixSum.raw", "Level2InOut.raw",
         "Level1Normal.raw", "Level2Normal.raw")


@pytest.mark.parametrize("vc,vc2,buf", [(8, 4, 1), (8, 4, -1), (64, 8, 1), (64, 8, -1)])
def test_object_dropin_cessna(gv, orc, tmp_path, vc, vc2, buf):
    exe = os.path.join(ROOT, "tests", "native", "object_dropin")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: run `make -C tests/native` before the GPU tests")
    mesh = os.path.join(GOLDEN, "cessna.obj")
    got, want = tmp_path / "got", tmp_path / "want"
    got.mkdir()
    want.mkdir()
    say(f"object drop-in: cessna {vc}/{vc2} bufferSize {buf}")
    r = subprocess.run([exe, mesh, str(vc), str(vc2), str(buf), str(got), "-1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    fields = r.stdout.split()
    stats = dict(zip(fields[0::2], fields[1::2]))
    v, f, _ = gv.read_obj(mesh)
    soup = gv.flatten(v, f)
    bmin, bmax, mms = gv.mesh_bbox(v)
    g = orc.grid_setup(bmin, bmax, mms, vc, vc2)
    orc.save_raw(soup, g, -1, str(want), normals=True)
    assert int(stats["triangles"]) == len(soup)
    if buf == 1:
        assert int(stats["regrows"]) >= 1, r.stdout
    names = sorted(os.listdir(want))
    assert names == sorted(f"Obj-1{n}" for n in FILES)
    assert sorted(os.listdir(got)) == names
    for n in names:
        assert (want / n).read_bytes() == (got / n).read_bytes(), n
