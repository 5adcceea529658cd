"""Full-frame rendering driver of the reference's render_360.py (render_360.py:59-170).

Frames of an orbiting pinhole camera are rendered with render_rays_batched
(render.py:49-102) in ray chunks; RGB frames become 8-bit images, distance
frames an inverse-distance visualisation scaled by the first frame
(render_360.py:140-160). Host-side orchestration: every ray is rendered by the
HIP library.
"""
from __future__ import annotations

import math
from typing import Iterator, List, Optional

import numpy as np

from . import cameras, render


def rotation_about(axis: np.ndarray, angle: float) -> np.ndarray:
    """3x3 rotation by `angle` about the unit `axis` (SO3.exp of angle * axis)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    k = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + math.sin(angle) * k + (1.0 - math.cos(angle)) * (k @ k)


def camera_up_axis(c2ws: List[np.ndarray]) -> np.ndarray:
    """Mean "camera up" of the given cameras in world coordinates (OpenCV
    cameras look along +z with y down, so up is -y; render_360.py:100-110)."""
    ups = np.array([c[:3, :3] @ np.array([0.0, -1.0, 0.0]) for c in c2ws])
    axis = ups.mean(axis=0)
    return axis / np.linalg.norm(axis)


def orbit(c2w: np.ndarray, frames: int, rotation_axis: np.ndarray) -> List[np.ndarray]:
    """Camera-to-world matrices of the orbit: frame i is the start camera with
    its world-to-camera transform right-multiplied i times by the rotation of
    2 pi / frames about the axis (render_360.py:163-170), i.e. the camera pose
    rotated by the inverse rotation about the world origin."""
    step = np.eye(4)
    step[:3, :3] = rotation_about(rotation_axis, 2.0 * math.pi / frames).T
    out, cur = [], np.asarray(c2w, np.float64)
    for _ in range(frames):
        out.append(cur.copy())
        cur = step @ cur
    return out


def rgb_image(rendered: np.ndarray) -> np.ndarray:
    """(H, W, 3) colours in [0, 1] -> uint8 (render_360.py:141-144)."""
    return np.clip(np.asarray(rendered) * 255.0, 0.0, 255.0).astype(np.uint8)


class DistanceImages:
    """Inverse-distance ("disparity") visualisation of distance frames with
    the scale fixed by the first frame (render_360.py:145-160):
    inv = 1 / max(d, 1e-4); lo = min(inv), hi = 0.9 max(inv) of frame 0;
    pixel = clip((inv - lo) / (hi - lo) * 255), repeated over 3 channels."""

    def __init__(self):
        self.lo: Optional[float] = None
        self.hi: Optional[float] = None

    def __call__(self, distances: np.ndarray) -> np.ndarray:
        inv = 1.0 / np.maximum(np.asarray(distances, np.float64), 1e-4)
        if self.lo is None or self.hi is None:
            self.lo = float(inv.min())
            self.hi = float(inv.max()) * 0.9
        img = (inv - self.lo) / (self.hi - self.lo)
        img = np.clip(img * 255.0, 0.0, 255.0).astype(np.uint8)
        return np.repeat(img[:, :, None], 3, axis=2)


def render_frames(state, start_c2w: np.ndarray, width: int, height: int, fov_x: float,
                  frames: int, mode: render.RenderMode = render.RenderMode.RGB,
                  density_samples_per_ray: int = 512, appearance_samples_per_ray: int = 128,
                  batch_size: int = 16384, rotation_axis: Optional[np.ndarray] = None,
                  camera_index: int = 0, translation_scale: float = 0.8,
                  prng_key: Optional[int] = 0) -> Iterator[np.ndarray]:
    """render_360.main without the experiment harness: the start camera's
    position is scaled by translation_scale (0.8 in the reference), frames of
    the orbit about rotation_axis (default: the start camera's up) are
    rendered in chunks of batch_size rays and yielded as uint8 (H, W, 3)
    images."""
    c2w = np.asarray(start_c2w, np.float64).copy()
    c2w[:3, 3] *= translation_scale
    axis = camera_up_axis([c2w]) if rotation_axis is None else np.asarray(rotation_axis)
    focal = cameras.focal_from_fov(width, fov_x)
    cfg = render.RenderConfig(state.config.render_near, state.config.render_far, mode,
                              density_samples_per_ray, appearance_samples_per_ray,
                              density_shift=state.config.density_shift)
    to_image = rgb_image if mode is render.RenderMode.RGB else DistanceImages()
    for pose in orbit(c2w, frames, axis):
        rays = cameras.pinhole_rays(pose, width, height, focal, camera_index,
                                    device=state.aabb.device)
        out = render.render_rays_batched(state.appearance_mlp, state.learnable_params, state.aabb,
                                         rays, prng_key, cfg, batch_size=batch_size)
        yield to_image(out)


def save_frames(images: Iterator[np.ndarray], output_dir) -> List[str]:
    """Writes image_000.png, image_001.png, ... (render_360.py:162)."""
    import pathlib

    from PIL import Image
    out = pathlib.Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)
    paths = []
    for i, img in enumerate(images):
        p = out / f"image_{i:03}.png"
        Image.fromarray(img).save(p)
        paths.append(str(p))
    return paths
