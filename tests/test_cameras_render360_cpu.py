"""CPU checks of ray generation (cameras.py:34-143 of the reference) and of the
full-frame render driver's host logic (render_360.py:100-170): closed forms,
no GPU needed."""
import math

import numpy as np
import pytest

from tensorf_amd import cameras, render_360, scenes


def _frame(c2w):
    right, down, fwd = c2w[:3, 0], c2w[:3, 1], c2w[:3, 2]
    return right, down, fwd


@pytest.mark.parametrize("w,h,fov", [(33, 21, 0.9), (800, 800, 2 * math.atan(400 / 1111.1)),
                                     (64, 48, 1.7)])
def test_pinhole_rays_closed_form(w, h, fov):
    """Pinhole rays project back onto their own pixel with cx = W/2 - 0.5,
    cy = H/2 - 0.5 and f = (W/2) / tan(fov_x/2) (cameras.py:45-56, 100-121):
    u = f (d.right)/(d.fwd) + cx, v = f (d.down)/(d.fwd) + cy; the edge
    pixels of the centre row are atan((W/2 - 0.5)/f) off the optical axis;
    directions are v / (|v| + 1e-8) (cameras.py:115)."""
    c2w = cameras.look_at(np.array([2.0, -3.0, 1.5]), np.array([0.1, 0.2, -0.3]))
    f = cameras.focal_from_fov(w, fov)
    assert abs(f - (w / 2) / math.tan(fov / 2)) < 1e-9 * f
    rays = cameras.pinhole_rays(c2w, w, h, f, 7, device="cpu")
    d = rays.directions.numpy().astype(np.float64)
    o = rays.origins.numpy().astype(np.float64)
    assert d.shape == (h, w, 3) and rays.camera_indices.shape == (h, w)
    assert np.all(rays.camera_indices.numpy() == 7)
    np.testing.assert_allclose(o, np.broadcast_to(c2w[:3, 3], o.shape), rtol=1e-6)
    right, down, fwd = _frame(c2w)
    vv, uu = np.mgrid[:h, :w]
    z = d @ fwd
    u = f * (d @ right) / z + (w / 2 - 0.5)
    v = f * (d @ down) / z + (h / 2 - 0.5)
    np.testing.assert_allclose(u, uu, atol=2e-3 * max(1.0, f / 100))
    np.testing.assert_allclose(v, vv, atol=2e-3 * max(1.0, f / 100))
    # the +1e-8 normalisation: |d| = |v| / (|v| + 1e-8), within float32 of 1
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1.0, atol=3e-7)
    # off-axis angle of the centre row's edge pixels
    row = h // 2
    ang = math.atan(math.hypot(w / 2 - 0.5, row - (h / 2 - 0.5)) / f)
    for col in (0, w - 1):
        got = math.acos(min(1.0, float(d[row, col] @ fwd) / np.linalg.norm(d[row, col])))
        assert abs(got - ang) < 2e-6


def test_fisheye_rays_equidistant():
    """Equidistant fisheye: a pixel at distance r from the centre
    (cx = W/2 - 0.5) looks theta = r / f off the optical axis, at azimuth
    atan2(v - cy, u - cx) in the image plane (x right, y down)."""
    w, h = 96, 54
    f = (w / 2.0) / (np.pi / 2.0)
    c2w = scenes.ring_cameras(8)[3]
    rays = scenes.fisheye_rays(c2w, w, h, f, 3, device="cpu")
    d = rays.directions.numpy().astype(np.float64)
    right, down, fwd = _frame(c2w)
    vv, uu = np.mgrid[:h, :w]
    x, y = uu - (w / 2 - 0.5), vv - (h / 2 - 0.5)
    theta = np.arctan2(np.linalg.norm(np.cross(d, fwd), axis=-1), d @ fwd)
    np.testing.assert_allclose(theta, np.hypot(x, y) / f, atol=1e-6)
    off = np.hypot(x, y) > 1.0
    az = np.arctan2(d @ down, d @ right)
    dif = np.angle(np.exp(1j * (az - np.arctan2(y, x))))
    assert np.abs(dif[off]).max() < 1e-4
    # 180 degrees across the width: the centre row's edge pixels look ~90 degrees off axis
    r_edge = math.hypot(w / 2 - 0.5, h // 2 - (h / 2 - 0.5))
    assert abs(theta[h // 2, 0] - r_edge / f) < 1e-6 and theta[h // 2, 0] < math.pi / 2


def test_distance_visualisation_closed_form():
    """render_360.py:145-160: inverse distances scaled by the first frame
    (lo = min, hi = 0.9 max), clipped, 3 channels; later frames keep that
    scale; distances below 1e-4 count as 1e-4 (zeros, the median's "sample 0")."""
    vis = render_360.DistanceImages()
    d0 = np.array([[1.0, 2.0], [4.0, np.inf]])
    img = vis(d0)
    inv = np.array([[1.0, 0.5], [0.25, 0.0]])
    lo, hi = 0.0, 0.9
    want = np.clip((inv - lo) / (hi - lo) * 255.0, 0, 255).astype(np.uint8)
    assert img.shape == (2, 2, 3) and img.dtype == np.uint8
    assert np.array_equal(img[:, :, 0], want) and np.array_equal(img[:, :, 2], want)
    assert img[0, 0, 0] == 255 and img[1, 1, 0] == 0
    img2 = vis(np.array([[0.0, 3.0]]))  # the first frame's scale is kept
    assert (vis.lo, vis.hi) == (lo, hi)
    assert img2[0, 0, 0] == 255
    assert img2[0, 1, 0] == np.uint8(np.clip((1 / 3 - lo) / (hi - lo) * 255.0, 0, 255))


def test_rgb_image_and_orbit():
    img = render_360.rgb_image(np.array([[[-0.2, 0.5, 1.3]]]))
    assert img.tolist() == [[[0, 127, 255]]]
    c2w = cameras.look_at(np.array([0.0, -4.0, 1.0]), np.zeros(3))
    axis = render_360.camera_up_axis([c2w])
    # OpenCV up (-y of the camera) of a camera looking at the origin from the
    # -y side with z up is mostly +z
    assert axis[2] > 0.9
    poses = render_360.orbit(c2w, 8, np.array([0.0, 0.0, 1.0]))
    assert len(poses) == 8
    pos = np.array([p[:3, 3] for p in poses])
    # a rotation about the world z axis: height and radius kept, 45 degrees per frame
    np.testing.assert_allclose(pos[:, 2], 1.0, atol=1e-12)
    np.testing.assert_allclose(np.hypot(pos[:, 0], pos[:, 1]), 4.0, atol=1e-12)
    ang = np.unwrap(np.arctan2(pos[:, 1], pos[:, 0]))
    np.testing.assert_allclose(np.abs(np.diff(ang)), 2 * np.pi / 8, atol=1e-12)
    # every pose still looks at the origin
    for p in poses:
        fwd = p[:3, 2]
        np.testing.assert_allclose(fwd, -p[:3, 3] / np.linalg.norm(p[:3, 3]), atol=1e-12)
    # and the rotation is orthonormal
    r = render_360.rotation_about(np.array([1.0, 2.0, 3.0]), 0.7)
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-14)
    np.testing.assert_allclose(r @ np.array([1.0, 2.0, 3.0]), [1.0, 2.0, 3.0], atol=1e-14)
