"""CPU-side checks of the boundary: the C ABI loads and exports every symbol
the headers declare, host-side geometry (bbox, grid derivation, readers,
generators) matches the oracle and closed forms, and the compiled gfx950 code
keeps the exactness contract (no contraction, IEEE division, denormals)."""
import os
import re
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpv_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol(gv):
    L = gv.lib()
    declared = _declared("gpv.h") + _declared("gpv_mesh.h") + _declared("gpv_serve.h")
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), f"{name} declared but not exported"
    assert sorted(declared) == sorted(gv.EXPORTED)


def test_batch_descriptors_match_struct(gv):
    """The batch's numpy descriptor array holds, byte for byte, the
    gpv_mesh_desc structs that the ctypes mirror builds for each mesh, also
    for read-only, empty, float64 and strided soups and for device pointers."""
    import ctypes
    ro = np.arange(18, dtype=np.float32)
    ro.setflags(write=False)
    soups = [np.arange(27, dtype=np.float32).reshape(3, 9), ro, np.zeros((0, 9), np.float32),
             np.arange(27, dtype=np.float64).reshape(3, 9), np.arange(36, dtype=np.float32).reshape(2, 18)[:, :9]]
    extra = [{}, {"voxel_count2": 8, "level2": False}, {"normals": True, "initial_capacity": 1 << 40},
             {"cuda_level2": True}, {"l1_stencil": True}]
    meshes = [dict(soup=x, bmin=(0.25 * i, -1.0, 2.0), bmax=np.array([3.0, 4.0 + i, 5.0], np.float32), mms=6.5 + i,
                   voxel_count=16 + i, **e) for i, (x, e) in enumerate(zip(soups, extra))]
    for ptrs in (None, [4096 * (i + 1) for i in range(len(meshes))]):
        a, keep = gv.mesh_descs(meshes, ptrs)
        assert a.dtype.itemsize == ctypes.sizeof(gv.MeshDesc)
        for i, m in enumerate(meshes):
            d, soup = gv.mesh_desc(m, None if ptrs is None else ptrs[i])
            if ptrs is None:
                # the ctypes mirror converted its own copy; the batch keeps its soup
                assert a["tris"][i] == keep[i].__array_interface__["data"][0]
                assert np.array_equal(keep[i].reshape(-1, 9), soup)
                d.tris = int(a["tris"][i])
            assert a[i:i + 1].tobytes() == ctypes.string_at(ctypes.addressof(d), ctypes.sizeof(d)), i
    with pytest.raises(ValueError):
        gv.mesh_descs([dict(meshes[0], soup=np.zeros(10, np.float32))])


def test_status_strings(gv):
    L = gv.lib()
    for s in range(9):
        assert L.gpv_status_string(s) != b"unknown status"


def test_bbox_and_grid_match_oracle(gv, orc):
    rng = np.random.default_rng(5)
    for trial in range(50):
        v = (rng.normal(0, 1, (rng.integers(3, 200), 3)) * rng.uniform(0.01, 100)).astype(np.float32)
        b1 = gv.mesh_bbox(v)
        b2 = orc.bbox(v)
        assert np.array_equal(b1[0], b2[0]) and np.array_equal(b1[1], b2[1]) and b1[2] == b2[2]
        vc = int(rng.integers(1, 300))
        vc2 = int(rng.integers(1, 17))
        i = gv.grid_derive(b1[0], b1[1], b1[2], vc, vc2)
        g = orc.grid_setup(b1[0], b1[1], b1[2], vc, vc2)
        assert list(i.n) == list(g.n) and list(i.g) == list(g.g) and list(i.g2) == list(g.g2)


def test_cessna_closed_form(gv, orc):
    """files/cessna.obj: raw vertex extremes, pad and divisions re-derived in
    numpy from the formulas of src/Object.cpp:572-583 and :3094-3136."""
    v, f, _ = gv.read_obj(os.path.join(GOLDEN, "cessna.obj"))
    assert v.shape == (3745, 3) and f.shape == (7446, 3)
    mn, mx = v.min(0), v.max(0)
    np.testing.assert_allclose(mn, [-19.563246, -5.119924, -22.152081], atol=1e-6)
    np.testing.assert_allclose(mx, [21.633591, 5.119924, 22.152081], atol=1e-6)
    d = (mx - mn).astype(np.float32)
    sq = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
    sq = np.float32(sq + np.float32(d[2] * d[2]))
    off = np.float32(0.001 * float(np.float32(np.sqrt(np.float64(sq)))))
    bmin = (mn - off).astype(np.float32)
    bmax = (off + mx).astype(np.float32)
    b = gv.mesh_bbox(v)
    assert np.array_equal(b[0], bmin) and np.array_equal(b[1], bmax)
    ext = (bmax - bmin).astype(np.float32)
    mms = ext.max()
    assert b[2] == mms
    for vc in (8, 64, 256):
        nominal = np.float32(float(mms) / float(vc))
        n = [int(np.float32(e / nominal)) for e in ext]
        n = [max(k, 1) for k in n]
        n = [k if k % 4 == 0 else k + 4 - k % 4 for k in n]
        g = [np.float32(float(e) / k) for e, k in zip(ext, n)]
        info = gv.grid_derive(b[0], b[1], b[2], vc, 4)
        assert list(info.n) == n and [np.float32(x) for x in info.g] == g
    # z is the longest axis (44.4): nominal 5.55 -> 7.4, 1.85, 8.0 -> 8, 4, 8
    assert list(gv.grid_derive(b[0], b[1], b[2], 8, 4).n) == [8, 4, 8]


def test_obj_reader_matches_numpy_parse(gv):
    path = os.path.join(GOLDEN, "cessna.obj")
    v, f, bff = gv.read_obj(path)
    vs, fs = [], []
    for line in open(path):
        w = line.split()
        if w and w[0] == "v":
            vs.append([np.float32(x) for x in w[1:4]])
        elif w and w[0] == "f":
            fs.append([int(x.split("/")[0]) - 1 for x in w[1:4]])
    assert np.array_equal(v, np.array(vs, np.float32)) and np.array_equal(f, np.array(fs, np.int32))
    assert not bff


def test_obj_reader_reference_quirks(gv, tmp_path):
    """Last line without a newline is dropped (src/Object.cpp:419-424);
    'a/b/c' and 'a//c' face forms; unreferenced vertices still count."""
    p = tmp_path / "q.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 5 5 5\nf 1/1/1 2//2 3\nf 1 2 3")
    v, f, _ = gv.read_obj(str(p))
    assert len(v) == 4 and f.tolist() == [[0, 1, 2]]
    assert gv.mesh_bbox(v)[1][0] > 5.0


def test_off_reader(gv, tmp_path):
    p = tmp_path / "t.off"
    p.write_text("OFF\n4 2 0\n0 0 0\n1 0 0\n0 1 0\n9 9 9\n3 0 1 2\n3 0 2 1\n")
    v, f, bff = gv.read_off(str(p))
    assert len(v) == 4 and f.tolist() == [[0, 1, 2], [0, 2, 1]] and bff


def _watertight(f):
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    fw = set(map(tuple, e.tolist()))
    return len(fw) == len(e) and all((b, a) in fw for a, b in e.tolist())


def test_generators_deterministic_and_watertight(gv):
    import hashlib
    v, f = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)
    assert v.shape == (2562, 3) and f.shape == (5120, 3) and _watertight(f)
    v2, f2 = gv.icosphere(4, 0.4, (0.5, 0.5, 0.5), 0.02, seed=1)
    assert hashlib.sha256(v.tobytes()).digest() == hashlib.sha256(v2.tobytes()).digest()
    r = np.linalg.norm(v.astype(np.float64) - 0.5, axis=1)
    assert 0.4 * 0.98 - 1e-6 <= r.min() and r.max() <= 0.4 * 1.02 + 1e-6
    for kind in range(3):
        v, f = gv.mixed(kind, 4000, 100 + kind)
        assert _watertight(f), kind
        assert v.min() >= 0.0 and v.max() <= 1.0
    v, f = gv.torus_knot(3, 7, 128, 16)
    assert f.shape == (2 * 128 * 16, 3) and _watertight(f)
    ext = v.max(0) - v.min(0)
    assert abs(ext.max() - 0.9) < 1e-6


def test_stress_keeps_watertight_and_bbox(gv):
    v, f = gv.icosphere(3)
    vs, fs = gv.stress(v, f, 32, 4, 0.3, 0.1, seed=3)
    assert _watertight(fs) and len(fs) > len(f)
    assert np.array_equal(gv.mesh_bbox(v)[0], gv.mesh_bbox(vs)[0])


def test_stress_coplanar_fraction(gv):
    """f_coplanar = 0.05: about 5% of triangles end with all three vertices on
    one Level1 lattice plane (bmin + k*g along some axis), watertight, bbox kept."""
    v, f = gv.icosphere(4)
    vs, fs = gv.stress(v, f, 64, 4, p_snap=0.0, f_sliver=0.0, f_coplanar=0.05, seed=9)
    assert _watertight(fs) and np.array_equal(gv.mesh_bbox(v)[0], gv.mesh_bbox(vs)[0])
    bmin, bmax, mms = gv.mesh_bbox(vs)
    info = gv.grid_derive(bmin, bmax, mms, 64, 4).as_dict()
    t = vs[fs]                                    # [T, 3, 3]
    on = np.zeros(len(fs), bool)
    for q in range(3):
        same = (t[:, 0, q] == t[:, 1, q]) & (t[:, 1, q] == t[:, 2, q])
        k = np.rint((t[:, 0, q].astype(np.float64) - bmin[q]) / np.float64(info["g"][q]))
        plane = (bmin[q] + k * np.float64(info["g"][q])).astype(np.float32)
        on |= same & (plane == t[:, 0, q])
    frac = on.mean()
    assert 0.03 < frac < 0.08, frac


def test_invalid_arguments(gv):
    with pytest.raises(gv.GpvError):
        gv.grid_derive(np.float32([0, 0, 0]), np.float32([1, 1, 1]), 1.0, 8, 17)
    with pytest.raises(gv.GpvError):
        gv.grid_derive(np.float32([0, 0, 0]), np.float32([1, 1, 1]), 1.0, 0, 4)


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs hipcc")
def test_c_headers_compile_as_c99(tmp_path):
    """include/gpv.h, gpv_mesh.h and gpv_serve.h are plain C: a C99 caller
    written like the integration notes (stage sequence, batch, serving loop,
    no-allocation entry point) compiles with warnings as errors and links
    against libgpv (tests/native/c_abi_usage.c)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    src = os.path.join(ROOT, "tests", "native", "c_abi_usage.c")
    inc = os.path.join(ROOT, "include")
    obj = str(tmp_path / "c_abi_usage.o")
    subprocess.run([cc, "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", inc, "-c", src, "-o", obj],
                   check=True)
    libdir = os.path.join(ROOT, "gpview_amd")
    if os.path.exists(os.path.join(libdir, "libgpv.so")):
        out = str(tmp_path / "libcheck.so")
        subprocess.run([cc, "-shared", "-o", out, obj, "-L", libdir, "-lgpv", "-Wl,--no-undefined",
                        "-Wl,-rpath," + libdir], check=True)


def test_codegen_exactness():
    """Our own gfx950 ISA: f32 denormals kept, f32 division is the
    correctly rounded v_div_scale/v_div_fmas/v_div_fixup expansion, f32 sqrt
    the correctly rounded v_sqrt_f32 + neighbour-residual expansion, no packed
    FP32, and every f32 FMA-family instruction belongs to one of those
    expansions (no contracted multiply-add)."""
    from gpview_amd import build
    s = build.disassemble().splitlines()
    assert not any(re.match(r"\s+v_pk_(add|mul|fma)_f32", l) for l in s)
    assert all(l.strip().endswith(" 3") for l in s if "amdhsa_float_denorm_mode_32" in l)
    ins = [l.strip() for l in s if re.match(r"\s+[vs]_", l)]
    assert any(x.startswith("v_div_fixup_f32") for x in ins)
    fma = re.compile(r"v_(fma_f32|fmac_f32|fmamk_f32|fmaak_f32|mad_f32|mac_f32)")
    for i, x in enumerate(ins):
        if fma.match(x):
            win = " ".join(ins[max(0, i - 12):i + 12])
            ok = ("v_div_scale_f32" in win or "v_div_fmas_f32" in win or "v_rcp_iflag_f32" in win
                  or ("v_rcp_f32" in win and ("v_cvt_u32_f32" in win or "v_mul_hi" in win))
                  or "v_sqrt_f32" in " ".join(ins[max(0, i - 4):i + 1]))
            assert ok, f"unexplained {x}"
    # every hardware sqrt (1 ulp) is followed by the neighbour residual test
    for i, x in enumerate(ins):
        if x.startswith("v_sqrt_f32"):
            win = ins[i + 1:i + 12]
            assert sum(1 for y in win if y.startswith("v_fma_f32")) >= 2, "bare v_sqrt_f32"
            assert any(y.startswith("v_add_u32") and ", -1," in y for y in win), "bare v_sqrt_f32"


def test_predict_footprint_monotone(gv):
    """Admission estimate: grows with triangles and with either resolution."""
    from gpview_amd import workloads
    m = workloads.plain_small(gv)
    base = gv.predict_footprint(m)
    assert base > 0
    bigger = dict(m, soup=np.concatenate([m["soup"]] * 4))
    assert gv.predict_footprint(bigger) > base
    assert gv.predict_footprint(dict(m, voxel_count=128)) > base
    assert gv.predict_footprint(dict(m, voxel_count2=8)) > base
    with pytest.raises(gv.GpvError):
        gv.predict_footprint(dict(m, voxel_count2=17))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_server_without_device_fails_cleanly(gv):
    with pytest.raises(gv.GpvError) as e:
        gv.Server(lanes=2)
    assert e.value.status == gv.GPV_ERR_NO_DEVICE
