"""Volume files, implicit datasets, histogram, TF builders and GUI settings (CPU only).

Pinned by the data the reference holds and by closed forms:
* .cvol: files written here byte by byte with struct (the header layout of
  renderer/volume.cpp:75-84) for all three voxel types are read back; files
  written by Volume.save are parsed back with struct; the reference's
  C60Large.cvol matches the checksums numpy decoded from it
  (tests/golden/c60large_cvol.json, made by tests/golden/make_io_fixtures.py);
* implicit datasets: each equation's value at grid points where its
  expression reduces to a closed form (renderer/volume_implicit.cpp), and the
  grid spacing (box / (resolution-1), renderer/volume.cpp:375-383);
* histogram: a volume with known values (renderer/volume.cpp:287-356);
* colour space: published sRGB / CIE L*a*b* (D65) values and round trips;
* TF builders: hand-merged axes and texel-centre interpolation
  (renderer/tf_utils.cpp:12-172);
* settings: two of the reference's GUI settings files (pytests/config-files).
"""
import json
import math
import os
import struct

import numpy as np
import pytest
import torch

from diffdvr_amd import pyrenderer as pr
from diffdvr_amd import settings as st

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
C60 = "/root/reference/pytests/volumes/C60Large.cvol"


def write_cvol(path, vox_xyz, voxel_size, typ, compressed=0):
    """A .cvol file from an [x][y][z] array, written with struct (x fastest on disk)."""
    X, Y, Z = vox_xyz.shape
    dt = {0: np.uint8, 1: np.uint16, 2: np.float32}[typ]
    body = np.ascontiguousarray(np.transpose(vox_xyz, (2, 1, 0))).astype(dt).tobytes()
    with open(path, "wb") as f:
        f.write(b"cvol" + struct.pack("<3Q3di", X, Y, Z, *voxel_size, typ) + bytes([compressed]) + bytes(7))
        f.write(body)


@pytest.mark.parametrize("typ,scale", [(0, 255.0), (1, 65535.0), (2, 1.0)])
def test_cvol_read_all_types(tmp_path, typ, scale):
    rng = np.random.default_rng(10 + typ)
    X, Y, Z = 5, 7, 3
    if typ == 2:
        raw = rng.uniform(0, 1, (X, Y, Z)).astype(np.float32)
    else:
        raw = rng.integers(0, int(scale) + 1, (X, Y, Z))
    p = tmp_path / f"v{typ}.cvol"
    write_cvol(p, raw, (0.5, 0.25, 2.0), typ)
    v = pr.Volume(str(p))
    d = v.getDataCpu(0)
    assert tuple(d.shape) == (1, X, Y, Z)
    assert d.stride() == (X * Y * Z, 1, X, X * Y)  # x fastest, as the reference's blob
    want = raw.astype(np.float32) / np.float32(scale) if typ != 2 else raw
    assert np.array_equal(d[0].numpy(), want.astype(np.float32))
    ws = v.world_size
    assert (ws.x, ws.y, ws.z) == (0.5 * X, 0.25 * Y, 2.0 * Z)
    r = v.resolution
    assert (r.x, r.y, r.z) == (X, Y, Z)


def test_cvol_save_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(20)
    a = rng.uniform(0, 1, (4, 6, 5)).astype(np.float32)
    v = pr.Volume.from_numpy(a)
    v.set_world_size(2.0, 3.0, 10.0)
    p = tmp_path / "out.cvol"
    v.save(str(p))
    raw = open(p, "rb").read()
    assert raw[:4] == b"cvol" and len(raw) == 64 + 4 * a.size
    assert struct.unpack("<3Q", raw[4:28]) == (4, 6, 5)
    assert struct.unpack("<3d", raw[28:52]) == (0.5, 0.5, 2.0)  # world size / size
    assert struct.unpack("<i", raw[52:56])[0] == 2 and raw[56] == 0
    body = np.frombuffer(raw[64:], dtype=np.float32).reshape(5, 6, 4)  # [z][y][x]
    assert np.array_equal(np.transpose(body, (2, 1, 0)), a)
    w = pr.Volume(str(p))
    assert torch.equal(w.getDataCpu(0), v.getDataCpu(0))
    assert (w.world_size.x, w.world_size.y, w.world_size.z) == (2.0, 3.0, 10.0)


def test_cvol_errors(tmp_path):
    p = tmp_path / "bad.cvol"
    open(p, "wb").write(b"nope" + bytes(60))
    with pytest.raises(RuntimeError, match="Illegal magic number"):
        pr.Volume(str(p))
    q = tmp_path / "lz4.cvol"
    write_cvol(q, np.zeros((2, 2, 2)), (1, 1, 1), 2, compressed=1)
    with pytest.raises(RuntimeError, match="LZ4"):
        pr.Volume(str(q))
    v = pr.Volume.from_numpy(np.zeros((2, 2, 2), np.float32))
    with pytest.raises(RuntimeError, match="compression"):
        v.save(str(tmp_path / "c.cvol"), compression=3)


def test_reference_volume_file():
    g = json.load(open(os.path.join(GOLDEN, "c60large_cvol.json")))
    if not os.path.exists(C60):
        pytest.skip("the reference's volume file is not on this host")
    v = pr.Volume(C60)
    d = v.getDataCpu(0)[0]
    assert list(d.shape) == g["size"]
    assert abs(v.world_size.x - g["voxel_size"][0] * g["size"][0]) < 1e-12
    assert int((d.double() * 255.0).round().sum().item()) == g["byte_sum"]
    for x, y, z, b in g["points"]:
        assert float(d[x, y, z]) == np.float32(b) / np.float32(255.0)
    bins, mn, mx, nnz = v.extract_histogram()
    assert nnz == g["nonzero"] and mn == g["min_byte"] / 255.0 and mx == np.float32(g["max_byte"]) / np.float32(255)
    assert abs(sum(bins) - 1.0) < 1e-3


def grid_value(eq, res, ix, iy, iz, **kw):
    v = pr.Volume.create_implicit(eq, res, **kw)
    return float(v.getDataCpu(0)[0, ix, iy, iz])


def test_implicit_closed_forms():
    E = pr.ImplicitEquation
    c = 2  # resolution 5 on [lo, hi]: index 2 is the centre
    # sphere / inverse sphere / cube on [-1, 1]
    assert grid_value(E.Sphere, 5, c, c, c) == 1.0
    assert abs(grid_value(E.Sphere, 5, 0, 0, 0) - (1 - math.sqrt(3))) < 1e-6
    assert abs(grid_value(E.InverseSphere, 5, 4, c, c) - 1.0) < 1e-7
    assert grid_value(E.Cube, 5, 3, 3, 3) == 1.0            # |x| = 0.5 <= scale
    assert abs(grid_value(E.Cube, 5, 4, 4, 4, scale=0.5) - (1 - math.sqrt(3 * 0.25))) < 1e-6
    assert abs(grid_value(E.Cube, 5, 4, c, c, scale=0.25) - 0.25) < 1e-7
    # Marschner-Lobb on the axis r = 0: rho = cos(2 pi fM)
    for fM, alpha in ((6.0, 0.25), (3.5, 0.5)):
        for iz in range(5):
            z = -1 + iz * 0.5
            want = ((1 - math.sin(math.pi * z / 2)) + alpha * (1 + math.cos(2 * math.pi * fM))) / (2 * (1 + alpha))
            assert abs(grid_value(E.MarschnerLobb, 5, c, c, iz, fM=fM, alpha=alpha) - want) < 1e-6
    # implicit surfaces mapped to density 0.5 - f, clamped to [0, 1], at the origin
    assert grid_value(E.DingDong, 5, c, c, c) == 0.5
    assert grid_value(E.Steiner, 5, c, c, c) == 0.5
    assert grid_value(E.CrossCap, 5, c, c, c) == 0.5
    assert grid_value(E.Kummer, 5, c, c, c) == 0.0           # f = 1
    assert grid_value(E.Heart, 5, c, c, c) == 1.0            # f = -1
    assert grid_value(E.Blobbly, 5, c, c, c) == 1.0          # f = -2 (the binding's spelling)
    a = 0.3
    assert abs(grid_value(E.Cassini, 5, c, c, c, a=a) - (0.5 - a ** 4)) < 1e-7
    assert abs(grid_value(E.Tube, 5, c, c, c) - 1.0) < 1e-6   # r = 0: (1 - 0.9) * 10
    assert abs(grid_value(E.Kleine, 5, c, c, c) - 1.5) < 1e-6  # not clamped
    s2 = math.sqrt(2)
    endrass = 0.5 + 64 * 4 - (-4 * (1 - s2) - 1 + 12 * s2) ** 2
    assert abs(grid_value(E.Endrass, 5, c, c, c) - endrass) / abs(endrass) < 1e-5
    # Barth on [-1.5, 1.5] (resolution 7: step 0.5); at z = -0.5 the shifted z is 0, at x = y = 0
    # the value is 0.5 - (1 + 2 phi)
    phi = (1 + math.sqrt(5)) / 2
    assert abs(grid_value(E.Barth, 7, 3, 3, 2) - (0.5 - (1 + 2 * phi))) < 1e-5
    # the reference writes z2 = z * 2 in Steiner (renderer/volume_implicit.cpp:161-170): at (0.25, 0, 0.25)
    # on [-0.5, 0.5] with resolution 5, f = x^2 * (2z) = 0.03125
    assert abs(grid_value(E.Steiner, 5, 3, c, 3) - (0.5 - 0.0625 * 0.5)) < 1e-7


def test_histogram_known_values():
    a = np.zeros((4, 4, 4), np.float32)
    a[0, 0, 0] = 1.0
    a[1, 0, 0] = 0.5
    a[2, 0, 0] = 0.25
    a[3, 3, 3] = 0.5
    bins, mn, mx, nnz = pr.Volume.from_numpy(a).extract_histogram()
    assert (mn, mx, nnz) == (0.0, 1.0, 4)
    width = 1.0 / 512
    want = np.zeros(512)
    for d in (1.0, 0.5, 0.25, 0.5):
        want[min(int(d / width), 511)] += 0.25
    assert np.allclose(bins, want)


def test_colour_space():
    lab = pr.TFUtils.rgb_to_lab(pr.real3(1, 1, 1))
    assert abs(lab.x - 100) < 1e-3 and abs(lab.y) < 1e-2 and abs(lab.z) < 1e-2
    red = pr.TFUtils.rgb_to_lab(pr.real3(1, 0, 0))     # published sRGB red: L*a*b* = 53.24, 80.09, 67.20
    assert abs(red.x - 53.24) < 0.01 and abs(red.y - 80.09) < 0.02 and abs(red.z - 67.20) < 0.02
    blk = pr.TFUtils.lab_to_rgb(pr.real3(0, 0, 0))
    assert max(abs(blk.x), abs(blk.y), abs(blk.z)) < 1e-6
    rng = np.random.default_rng(3)
    for rgb in rng.uniform(0, 1, (20, 3)):
        lab = pr.TFUtils.rgb_to_lab(pr.real3(*rgb))
        back = pr.TFUtils.lab_to_rgb(lab)
        assert np.allclose([back.x, back.y, back.z], rgb, atol=2e-4)


def test_tf_assemble_and_tensors():
    white, black = pr.TFUtils.rgb_to_lab(pr.real3(1, 1, 1)), pr.TFUtils.rgb_to_lab(pr.real3(0, 0, 0))
    pts = pr.TFUtils.assemble_from_settings([white, black], [0.0, 1.0], [0.2, 0.6, 1.0], [0.0, 0.5, 1.0],
                                            0.0, 1.0, 2.0, True)
    assert [p.pos for p in pts] == [0.0, 0.5, 1.0]
    one = 1.0 - np.finfo(np.float32).eps
    assert abs(pts[0].val.x - one) < 1e-3 and abs(pts[0].val.w - 0.4) < 1e-6      # opacity * scaling 2
    assert abs(pts[1].val.w - 1.2) < 1e-6
    mid = pr.TFUtils.lab_to_rgb(pr.real3(*[(a + b) / 2 for a, b in zip((white.x, white.y, white.z),
                                                                          (black.x, black.y, black.z))]))
    assert abs(pts[1].val.x - min(max(mid.x, 0), one)) < 1e-6
    assert abs(pts[2].val.w - 2.0 * one) < 1e-6
    # sentinels at -1 / 2 when an axis does not reach 0 / 1, density range mapping
    # (axes -1, 0.5, 2 each: the colour point at 0.5 comes first, then the opacity point at 0.5, which the
    # purge removes as closer than 1e-7 to its predecessor)
    p2 = pr.TFUtils.assemble_from_settings([white], [0.5], [0.5], [0.5], 0.0, 1.0, 1.0, False)
    assert [p.pos for p in p2] == [-1.0, 0.5, 0.5, 2.0]
    p2 = pr.TFUtils.assemble_from_settings([white], [0.5], [0.5], [0.5], 0.0, 1.0, 1.0, True)
    assert [p.pos for p in p2] == [-1.0, 0.5, 2.0]
    p3 = pr.TFUtils.assemble_from_settings([white, black], [0.0, 1.0], [0.0, 1.0], [0.0, 1.0], 0.2, 0.6, 1.0, False)
    # editor positions map to densities 0.2 / 0.6 first, then the sentinels at -1 / 2 are added
    assert [round(p.pos, 6) for p in p3] == [-1.0, 0.2, 0.2, 0.6, 0.6, 2.0]
    # piecewise tensor columns (r, g, b, opacity, position)
    pw = pr.TFUtils.get_piecewise_tensor(pts)
    assert pw.shape == (1, 3, 5) and float(pw[0, 1, 4]) == 0.5 and abs(float(pw[0, 1, 3]) - 1.2) < 1e-6
    # texture: texel centres (i + 0.5) / R interpolate the control points
    tex = pr.TFUtils.get_texture_tensor([pr.TFPoint(0.0, pr.real4(0, 0, 0, 0)), pr.TFPoint(1.0, pr.real4(1, 2, 3, 4))],
                                        8)
    want = (np.arange(8) + 0.5) / 8
    assert np.allclose(tex[0].numpy(), np.outer(want, [1, 2, 3, 4]), atol=1e-6)


def test_settings_files(tmp_path):
    s = st.Settings(os.path.join(GOLDEN, "skull1.json"))
    assert not s.is_synthetic_dataset() and s.get_dataset_name().endswith(".cvol")
    cam = s.get_camera()
    assert abs(cam.pitch_radians - math.radians(5.0)) < 1e-12 and abs(cam.fov_y_radians - math.radians(45)) < 1e-12
    assert abs(cam.distance - 1.100000023841858 ** 4.0) < 1e-12 and int(cam.orientation) == 4
    pts = s.get_tf_points()
    assert len(pts) >= 2 and all(pts[i].pos <= pts[i + 1].pos for i in range(len(pts) - 1))
    assert s.get_stepsize() == 0.2
    tex = pr.TFUtils.get_texture_tensor(pts, 64)
    assert tex.shape == (1, 64, 4) and torch.isfinite(tex).all()
    g = st.Settings(os.path.join(GOLDEN, "tooth3gauss.json"))
    gt = g.get_gaussian_tensor()
    raw = json.load(open(os.path.join(GOLDEN, "tooth3gauss.json")))["tfEditor"]
    e = raw["editorGaussian"]
    assert gt.shape == (1, len(e), 6)
    mn, mx = raw["minDensity"], raw["maxDensity"]
    assert abs(float(gt[0, 0, 4]) - (mn + e[0][4] * (mx - mn))) < 1e-6
    assert abs(float(gt[0, 0, 3]) - e[0][3] * raw["opacityScaling"]) < 1e-5
    # a synthetic dataset and the default renderer inputs built for it
    cfg = {"version": 2, "dataset": {"file": "", "syntheticDataset": 0, "syntheticDatasetResolutionPower": 4}}
    p = tmp_path / "synthetic.json"
    json.dump(cfg, open(p, "w"))
    vol = st.Settings(str(p)).load_dataset()
    assert vol.resolution.x == 16
    inp = st.setup_default_settings(vol, 32, 24, step_size=0.5, run_on_cuda=False)
    assert abs(inp.step_size - 0.5 / 16) < 1e-12
