"""Camera-path files (CameraLoader, src/utils/camera_loader.cu:62-150) and the
frames they define, on the first keyframes of the reference's
assets/cameras/truck_camera_path.json (tests/golden/truck_cameras_head.json):

  * the loader returns the file's values (as float32), sorted by id, and
    save / load round-trips them exactly,
  * a camera keyframe whose rotation columns are (s, -u, f) of a look-at
    camera gives that look-at camera's view matrix (the viewer's
    scale(1,-1,-1) * affineInverse(T * mat4_cast(quat_cast(basis))) chain
    against glm::lookAt, which the frame tests already pin),
  * fovy = 2 atan(h / 2 fy) in degrees; the first camera keeps the width and
    sets the height from its aspect ratio,
  * playback passes through every keyframe at multiples of 1/24 s and
    interpolates between them.
"""
import json
import math
import os

import numpy as np
import pytest

import gps_amd as G

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "truck_cameras_head.json")


def test_load_values_and_order(tmp_path):
    raw = json.load(open(FIX))
    cams = G.camera_path_load(FIX)
    assert [c.id for c in cams] == [r["id"] for r in raw] == [0, 1, 2, 3]
    for c, r in zip(cams, raw):
        assert (c.width, c.height) == (r["width"], r["height"])
        assert c.img_name.decode() == r["img_name"]
        assert np.array_equal(np.array(c.position, np.float32), np.float32(r["position"]))
        assert np.array_equal(np.array(c.rotation, np.float32), np.float32(r["rotation"]).reshape(9))
        assert c.fy == np.float32(r["fy"]) and c.fx == np.float32(r["fx"])
    # sorted by id whatever the file order; save / load is exact
    p = str(tmp_path / "shuffled.json")
    json.dump([raw[2], raw[0], raw[3], raw[1]], open(p, "w"))
    assert [c.id for c in G.camera_path_load(p)] == [0, 1, 2, 3]
    q = str(tmp_path / "saved.json")
    G.camera_path_save(q, cams)
    back = G.camera_path_load(q)
    for a, b in zip(cams, back):
        assert bytes(a) == bytes(b)


def test_bad_files_fail(tmp_path):
    with pytest.raises(G.GpsError):
        G.camera_path_load(str(tmp_path / "missing.json"))
    p = str(tmp_path / "bad.json")
    open(p, "w").write('[{"id": 0, "width": 10}]')
    with pytest.raises(G.GpsError):
        G.camera_path_load(p)
    open(p, "w").write('[{"id": 0,')
    with pytest.raises(G.GpsError):
        G.camera_path_load(p)


def _lookat_state(eye, center, up, w=640, h=360, fy=500.0):
    e, c, u = (np.array(v, np.float64) for v in (eye, center, up))
    f = (c - e) / np.linalg.norm(c - e)
    s = np.cross(f, u)
    s /= np.linalg.norm(s)
    uu = np.cross(s, f)
    R = np.stack([s, -uu, f], axis=1)  # columns: camera x (right), y (down), z (forward)
    st = G.CameraState()
    st.id, st.width, st.height, st.fy, st.fx = 0, w, h, fy, fy
    for k in range(3):
        st.position[k] = eye[k]
    for k, v in enumerate(R.reshape(9)):
        st.rotation[k] = v
    return st


@pytest.mark.parametrize("eye,center", [((0.0, 0.0, -30.0), (0.0, 0.0, 0.0)), ((12.0, 5.0, -7.0), (1.0, -2.0, 3.0)),
                                        ((-3.0, 40.0, 2.0), (0.0, 0.0, 0.0))])
def test_camera_frame_equals_look_at(eye, center):
    st = _lookat_state(eye, center, (0.0, 1.0, 0.0))
    a = G.frame_from_camera(st, 640, 360, fovy=45.0, ss=2)
    b = G.frame_params(eye, center, (0.0, 1.0, 0.0), 45.0, 640, 360, ss=2)
    va, vb = np.array(a.view), np.array(b.view)
    scale = max(1.0, float(np.abs(vb).max()))
    assert np.allclose(va, vb, rtol=0, atol=3e-6 * scale), (va.reshape(4, 4), vb.reshape(4, 4))
    assert np.array_equal(np.array(a.proj), np.array(b.proj))
    assert np.allclose(np.array(a.frustum_planes), np.array(b.frustum_planes), rtol=0, atol=1e-5)


def test_fovy_and_resolution_rules():
    cams = G.camera_path_load(FIX)
    c = cams[0]
    want = math.degrees(2 * math.atan(0.5 * c.height / c.fy))
    assert abs(G.camera_fovy(c) - want) < 1e-4
    assert G.camera_resolution(c, 1920) == (1920, round(1920 * 3286 / 4946))
    f = G.frame_from_camera(c, 1920, 1276, ss=2)
    assert abs(f.fovy_degrees - G.camera_fovy(c)) == 0.0
    # the camera sits at its position: view * (p, 1) = (0, 0, 0, 1)
    V = np.array(f.view, np.float64).reshape(4, 4).T
    p = np.array(list(c.position) + [1.0])
    assert np.allclose(V @ p, [0, 0, 0, 1], atol=1e-5)


def test_playback_hits_keyframes_and_interpolates():
    cams = G.camera_path_load(FIX)
    for k in range(4):
        f, fin = G.camera_path_frame(cams, k / 24.0, 640, 360)
        g = G.frame_from_camera(cams[k], 640, 360, fovy=G.camera_fovy(cams[0]))
        assert np.allclose(np.array(f.view), np.array(g.view), rtol=0, atol=2e-5), k
        assert fin == (k == 3)
    f, _ = G.camera_path_frame(cams, 0.5 / 24.0, 640, 360)
    V = np.array(f.view, np.float64).reshape(4, 4).T
    R = V[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-5)
    mid = 0.5 * (np.array(cams[0].position, np.float64) + np.array(cams[1].position, np.float64))
    assert np.allclose(V @ np.append(mid, 1.0), [0, 0, 0, 1], atol=1e-5)
