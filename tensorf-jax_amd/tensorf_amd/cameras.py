"""Ray containers and synthetic cameras (cameras.py:10-143 of the reference)."""
from __future__ import annotations

import dataclasses
import math
from typing import Tuple

import numpy as np
import torch


@dataclasses.dataclass
class Rays3D:
    """cameras.Rays3D (cameras.py:10-20): origins (*, 3), directions (*, 3),
    camera_indices (*,) uint32 (stored as int32 tensors on the device)."""

    origins: torch.Tensor
    directions: torch.Tensor
    camera_indices: torch.Tensor

    def get_batch_axes(self) -> Tuple[int, ...]:
        return tuple(self.origins.shape[:-1])

    def reshape_flat(self) -> "Rays3D":
        return Rays3D(self.origins.reshape(-1, 3).contiguous(),
                      self.directions.reshape(-1, 3).contiguous(),
                      self.camera_indices.reshape(-1).contiguous())

    def slice(self, start: int, stop: int) -> "Rays3D":
        return Rays3D(self.origins[start:stop], self.directions[start:stop],
                      self.camera_indices[start:stop])


@dataclasses.dataclass
class RenderedRays:
    """data.RenderedRays: target colours (*, 3) for a batch of rays."""

    colors: torch.Tensor
    rays_wrt_world: Rays3D

    def get_batch_axes(self) -> Tuple[int, ...]:
        return tuple(self.colors.shape[:-1])


def pinhole_rays(c2w: np.ndarray, width: int, height: int, focal: float, camera_index: int,
                 device="cuda") -> Rays3D:
    """Per-pixel rays of a pinhole camera in the OpenCV convention
    (cameras.py:34-76, 100-143): cx = W/2 - 0.5, cy = H/2 - 0.5, direction
    R_world_camera K^-1 [u v 1]^T normalised with +1e-8 (cameras.py:115).
    c2w: (4, 4) camera-to-world transform whose rotation maps OpenCV camera
    axes (x right, y down, z forward) to the world."""
    v, u = np.mgrid[:height, :width].astype(np.float64)
    cx, cy = width / 2.0 - 0.5, height / 2.0 - 0.5
    dirs_c = np.stack([(u - cx) / focal, (v - cy) / focal, np.ones_like(u)], -1)
    dirs_w = dirs_c @ c2w[:3, :3].T
    dirs_w = dirs_w / (np.linalg.norm(dirs_w, axis=-1, keepdims=True) + 1e-8)
    origins = np.broadcast_to(c2w[:3, 3], dirs_w.shape)
    cam = np.full((height, width), camera_index, np.int32)
    return Rays3D(torch.tensor(origins, dtype=torch.float32, device=device),
                  torch.tensor(dirs_w, dtype=torch.float32, device=device),
                  torch.tensor(cam, device=device))


def look_at(eye: np.ndarray, target: np.ndarray, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Camera-to-world matrix (OpenCV axes) looking from eye to target."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    if np.linalg.norm(right) < 1e-9:
        right = np.cross(fwd, np.array([0.0, 1.0, 0.0]))
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, down, fwd, eye
    return m


def focal_from_fov(width: int, fov_x_radians: float) -> float:
    """fx = (W/2) / tan(fov_x/2) (cameras.py:52-53)."""
    return (width / 2.0) / math.tan(fov_x_radians / 2.0)
