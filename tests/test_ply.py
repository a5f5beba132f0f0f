"""3DGS PLY loading / writing (the reference's GaussianLoader rules,
src/core/gaussians/gaussian_loader.cu:105-427, 455-562), checked on
hand-built files whose decoded values are fixed by the format's definition:
sigmoid(0) = 0.5, sigmoid(ln 3) = 0.75, exp(ln 2) = 2, normalize((2,0,0,0)) =
(1,0,0,0), and f_rest channel-major -> coefficient-major RGB."""
import math
import os

import numpy as np
import pytest

import gps_amd as G

FIELDS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(45)] +
          ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def write_raw(path, rows, fields=FIELDS, comments=(), extra_uchar=False, types=None):
    types = types or {}
    head = ["ply", "format binary_little_endian 1.0"] + [f"comment {c}" for c in comments]
    head.append(f"element vertex {len(rows)}")
    if extra_uchar:
        head.append("property uchar flags")
    head += [f"property {types.get(f, 'float')} {f}" for f in fields]
    head.append("end_header")
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        for r in rows:
            if extra_uchar:
                fh.write(bytes([7]))
            for f in fields:
                if types.get(f) == "double":
                    fh.write(np.float64(r[f]).tobytes())
                else:
                    fh.write(np.float32(r[f]).tobytes())


def row(**kw):
    r = {f: 0.0 for f in FIELDS}
    r.update(kw)
    return r


def test_activations_and_sh_layout(tmp_path):
    rest = {f"f_rest_{i}": 100.0 + i for i in range(45)}
    rows = [row(x=1.5, y=-2.0, z=3.25, f_dc_0=0.1, f_dc_1=0.2, f_dc_2=0.3, opacity=0.0, scale_0=math.log(2.0),
                scale_1=0.0, scale_2=math.log(0.5), rot_0=2.0, **rest),
            row(opacity=math.log(3.0), rot_0=1.0, rot_1=1.0, rot_2=1.0, rot_3=1.0, **rest)]
    p = str(tmp_path / "a.ply")
    write_raw(p, rows, extra_uchar=True)
    info = G.ply_info(p)
    assert info == {"num_gaussians": 2, "record_bytes": 1 + 4 * len(FIELDS), "has_rest": True, "bvh_sorted": False}
    s = G.load_ply(p)
    assert np.array_equal(s.means[0], np.float32([1.5, -2.0, 3.25]))
    assert s.opacities[0] == 0.5 and abs(s.opacities[1] - 0.75) < 1e-7
    assert np.allclose(s.scales[0], [2.0, 1.0, 0.5], rtol=1e-7)
    assert np.array_equal(s.rotations[0], np.float32([1, 0, 0, 0]))
    assert np.array_equal(s.rotations[1], np.float32([0.5, 0.5, 0.5, 0.5]))
    assert np.array_equal(s.shs[0, :3], np.float32([0.1, 0.2, 0.3]))
    # coefficient j (1..15), channel c comes from f_rest_{(j-1) + 15 c}
    for j in range(1, 16):
        for c in range(3):
            assert s.shs[0, 3 * j + c] == 100.0 + (j - 1) + 15 * c


def test_missing_rest_reads_zero_and_bad_files_fail(tmp_path):
    fields = [f for f in FIELDS if not f.startswith("f_rest")]
    p = str(tmp_path / "b.ply")
    write_raw(p, [row(f_dc_0=0.5, rot_0=1.0)], fields=fields)
    assert G.ply_info(p)["has_rest"] is False
    s = G.load_ply(p)
    assert s.shs[0, 0] == 0.5 and not s.shs[0, 3:].any()
    with pytest.raises(G.GpsError):
        G.load_ply(str(tmp_path / "missing.ply"))
    q = str(tmp_path / "c.ply")
    with open(q, "wb") as fh:
        fh.write(b"not a ply file\n")
    with pytest.raises(G.GpsError):
        G.ply_info(q)
    # a required field of another type is rejected, not misread
    d = str(tmp_path / "d.ply")
    write_raw(d, [row(rot_0=1.0)], types={"x": "double"})
    with pytest.raises(G.GpsError):
        G.load_ply(d)


def test_write_load_round_trip_and_morton(tmp_path):
    a = G.synth_scene(G.synth_params(11, 3000, kind="clustered", sh_degree=3, clusters=10, half_extent=50.0))
    p = str(tmp_path / "r.ply")
    G.write_ply(p, a)
    b = G.load_ply(p)
    assert np.array_equal(b.means, a.means)
    assert np.array_equal(b.shs, a.shs)
    assert np.allclose(b.opacities, a.opacities, rtol=2e-6, atol=1e-7)
    assert np.allclose(b.scales, a.scales, rtol=2e-6)
    assert np.allclose(b.rotations, a.rotations, atol=2e-7)
    # --sort-morton-order: the loaded order permuted by the Morton order
    c = G.load_ply(p, sort_morton_order=True)
    perm = b.morton_order()
    assert np.array_equal(c.means, b.means[perm]) and np.array_equal(c.shs, b.shs[perm])
    # a file marked bvh_sorted_order keeps its order
    q = str(tmp_path / "s.ply")
    G.write_ply(q, a, mark_sorted=True)
    assert G.ply_info(q)["bvh_sorted"] is True
    d = G.load_ply(q, sort_morton_order=True)
    assert np.array_equal(d.means, a.means)
    # lower SH degrees are written zero-padded to the full layout
    e = G.synth_scene(G.synth_params(12, 10, sh_degree=1))
    r = str(tmp_path / "e.ply")
    G.write_ply(r, e)
    f = G.load_ply(r)
    assert np.array_equal(f.shs[:, :12], e.shs) and not f.shs[:, 12:].any()
    assert os.path.getsize(r) > 10 * 62 * 4
