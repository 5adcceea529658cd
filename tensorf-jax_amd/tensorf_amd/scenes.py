"""Procedural scenes, cameras and ray datasets for training and evaluation.

The reference trains on image datasets (data.py); this build's workloads are
procedural: a density field made of isotropic Gaussian blobs with per-blob
albedo, ground truth rendered by FP64 emission-absorption quadrature on the
GPU (tfh_blob_scene_render), white background. Rays follow the reference's
pinhole convention (cameras.py:34-143: cx = W/2 - 0.5, OpenCV axes, unit
directions with +1e-8).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np
import torch

from . import _lib, cameras


@dataclasses.dataclass
class BlobScene:
    """blobs: (B, 8) float64 rows [cx, cy, cz, sigma, peak, r, g, b]."""
    blobs: np.ndarray
    unbounded: bool = False

    def density(self, pts: np.ndarray) -> np.ndarray:
        b = self.blobs
        d2 = ((pts[:, None, :] - b[None, :, :3]) ** 2).sum(-1)
        return (b[None, :, 4] * np.exp(-d2 / (2 * b[None, :, 3] ** 2))).sum(-1)


def bounded_blob_scene(seed: int = 0, n: int = 8) -> BlobScene:
    """8 blobs: centres U[-0.8,0.8]^3, sigma U[0.1,0.3], peak 50, albedo U[0,1]^3."""
    rng = np.random.default_rng(seed)
    blobs = np.concatenate([rng.uniform(-0.8, 0.8, (n, 3)), rng.uniform(0.1, 0.3, (n, 1)),
                            np.full((n, 1), 50.0), rng.uniform(0, 1, (n, 3))], axis=1)
    return BlobScene(blobs)


def unbounded_blob_scene(seed: int = 0, n: int = 32) -> BlobScene:
    """32 blobs: centre radius log-uniform in [0.2, 20] (uniform direction),
    sigma = radius * U(0.05, 0.15), peak 50, albedo U[0,1]^3."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rad = np.exp(rng.uniform(np.log(0.2), np.log(20.0), n))
    centres = v * rad[:, None]
    sig = rad * rng.uniform(0.05, 0.15, n)
    blobs = np.concatenate([centres, sig[:, None], np.full((n, 1), 50.0),
                            rng.uniform(0, 1, (n, 3))], axis=1)
    return BlobScene(blobs, unbounded=True)


def hemisphere_cameras(n_views: int, radius: float = 4.0, seed: int = 1) -> List[np.ndarray]:
    """Camera-to-world matrices on the upper hemisphere (z > 0) looking at the
    origin, spread by a Fibonacci spiral with a seeded phase."""
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 2 * np.pi)
    out = []
    for i in range(n_views):
        z = 0.05 + 0.9 * (i + 0.5) / n_views
        phi = phase + i * np.pi * (3.0 - np.sqrt(5.0))
        rxy = np.sqrt(1 - z * z)
        eye = radius * np.array([rxy * np.cos(phi), rxy * np.sin(phi), z])
        out.append(cameras.look_at(eye, np.zeros(3)))
    return out


def ring_cameras(n_views: int, radius: float = 3.0, height_jitter: float = 0.3,
                 seed: int = 2) -> List[np.ndarray]:
    """Outward-free ring: cameras on a circle of the given radius (heights
    jittered by +-height_jitter) looking at the origin."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_views):
        a = 2 * np.pi * i / n_views
        eye = np.array([radius * np.cos(a), radius * np.sin(a), rng.uniform(-height_jitter,
                                                                             height_jitter)])
        out.append(cameras.look_at(eye, np.zeros(3)))
    return out


def fisheye_rays(c2w: np.ndarray, width: int, height: int, focal: float, camera_index: int,
                 device="cuda") -> cameras.Rays3D:
    """Equidistant fisheye (r_img = f * theta) rays in OpenCV camera axes."""
    v, u = np.mgrid[:height, :width].astype(np.float64)
    x = u - (width / 2.0 - 0.5)
    y = v - (height / 2.0 - 0.5)
    r = np.sqrt(x * x + y * y)
    theta = r / focal
    phi = np.arctan2(y, x)
    dirs_c = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)],
                      -1)
    dirs_w = dirs_c @ c2w[:3, :3].T
    dirs_w /= np.linalg.norm(dirs_w, axis=-1, keepdims=True) + 1e-8
    origins = np.broadcast_to(c2w[:3, 3], dirs_w.shape)
    cam = np.full((height, width), camera_index, np.int32)
    return cameras.Rays3D(torch.tensor(origins, dtype=torch.float32, device=device),
                          torch.tensor(dirs_w, dtype=torch.float32, device=device),
                          torch.tensor(cam, device=device))


def render_ground_truth(scene: BlobScene, rays: cameras.Rays3D, samples_per_ray: int = 4096,
                        t_near: Optional[float] = None, t_far: Optional[float] = None,
                        chunk: int = 1 << 20, with_depth: bool = False):
    """FP64 quadrature of the scene along the rays: (colours (*, 3), opacity (*,))
    and, with_depth, the median termination depth (*,) (+inf if never reached)."""
    flat = rays.reshape_flat()
    n = flat.origins.shape[0]
    dev = flat.origins.device
    blobs = torch.tensor(scene.blobs, dtype=torch.float64, device=dev).contiguous()
    rgb = torch.empty((n, 3), dtype=torch.float32, device=dev)
    opa = torch.empty((n,), dtype=torch.float32, device=dev)
    depth = torch.empty((n,), dtype=torch.float32, device=dev) if with_depth else None
    if scene.unbounded:
        tn, tf = (t_near or 0.05), (t_far or 200.0)
    else:
        tn, tf = (t_near or 1.0), (t_far or 7.0)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        o = flat.origins[s:e].contiguous()
        d = flat.directions[s:e].contiguous()
        _lib.call("tfh_blob_scene_render", _lib.ptr(blobs), scene.blobs.shape[0], _lib.ptr(o),
                  _lib.ptr(d), e - s, float(tn), float(tf), int(samples_per_ray),
                  int(scene.unbounded), _lib.ptr(rgb[s:e]), _lib.ptr(opa[s:e]),
                  _lib.ptr(depth[s:e]) if with_depth else None, _lib.stream())
        torch.cuda.synchronize()
    shape = rays.get_batch_axes()
    if with_depth:
        return rgb.view(*shape, 3), opa.view(*shape), depth.view(*shape)
    return rgb.view(*shape, 3), opa.view(*shape)


@dataclasses.dataclass
class RayDataset:
    """Flattened training rays and colours on the device (data.py:301-337)."""
    rays: cameras.Rays3D
    colors: torch.Tensor
    image_shape: tuple
    n_views: int

    def minibatches(self, batch: int, seed: int = 0):
        """Cycled shuffled minibatches (the role of fifteen's cycled_minibatches)."""
        n = self.colors.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        while True:
            perm = torch.randperm(n, generator=g).to(self.colors.device)
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                yield cameras.RenderedRays(
                    self.colors[idx].contiguous(),
                    cameras.Rays3D(self.rays.origins[idx].contiguous(),
                                   self.rays.directions[idx].contiguous(),
                                   self.rays.camera_indices[idx].contiguous()))

    def view(self, i: int) -> cameras.RenderedRays:
        h, w = self.image_shape
        s, e = i * h * w, (i + 1) * h * w
        return cameras.RenderedRays(self.colors[s:e].view(h, w, 3),
                                    cameras.Rays3D(self.rays.origins[s:e].view(h, w, 3),
                                                   self.rays.directions[s:e].view(h, w, 3),
                                                   self.rays.camera_indices[s:e].view(h, w)))


def make_dataset(scene: BlobScene, c2ws: List[np.ndarray], width: int, height: int, focal: float,
                 fisheye: bool = False, gt_samples: int = 4096, device="cuda") -> RayDataset:
    """Rays of every view with their ground-truth colours."""
    all_rays = []
    for i, c2w in enumerate(c2ws):
        if fisheye:
            all_rays.append(fisheye_rays(c2w, width, height, focal, i, device))
        else:
            all_rays.append(cameras.pinhole_rays(c2w, width, height, focal, i, device))
    o = torch.cat([r.origins.reshape(-1, 3) for r in all_rays]).contiguous()
    d = torch.cat([r.directions.reshape(-1, 3) for r in all_rays]).contiguous()
    c = torch.cat([r.camera_indices.reshape(-1) for r in all_rays]).contiguous()
    rays = cameras.Rays3D(o, d, c)
    rgb, _ = render_ground_truth(scene, rays, gt_samples)
    return RayDataset(rays, rgb, (height, width), len(c2ws))


def focal_800() -> float:
    """f = 1111.1 px at 800x800 (fov_x = 2 atan(400 / 1111.1))."""
    return 1111.1


def psnr(pred: torch.Tensor, target: torch.Tensor) -> float:
    mse = float(((pred - target) ** 2).mean())
    return -10.0 * math.log10(max(mse, 1e-10))
