"""Pose -> VirtualToMoving matrix on the host, for Python callers that need a matrix
without an engine pair (data simulation, load generators). Same conventions as the
engine's C++ geometry (csrc/lc2e_geometry.hpp):

  virtual grid: origin = fixed origin, spacing = resampled, dims = ceil(dim*sp/resampled)
                (interpolator.cu:320, :337-339)
  M(x) = A (x - C + shift) + C + T, C = fixed centre, shift = c_moving - c_virtual when the
  centres are overlaid (interpolator.cu:495-529); A = Rz Ry Rx Sh diag(S) in 3D
  (spatialMappers.cuh:637-850), A = R Sh diag(S) in 2D (spatialMappers.cuh:999-1033).
Returned matrices are float32, 3x4 (3D) or 3x3 homogeneous (2D).
"""
from __future__ import annotations

import math

import numpy as np


def center(dims, spacing, origin) -> np.ndarray:
    """origin + spacing * dim / 2 evaluated in float32 (PhysioSpatial*Mapper::SetCenter)."""
    d = np.asarray(dims, np.float32)
    s = np.asarray(spacing, np.float32)
    o = np.asarray(origin, np.float32)
    return (o + (s * d) / np.float32(2.0)).astype(np.float32)


def virtual_dims(dims, spacing, resampled):
    out = []
    for d, s, r in zip(dims, spacing, resampled):
        ext = np.float32(d) * np.float32(s)
        out.append(int(math.ceil(float(np.float32(ext / np.float32(r))))))
    return tuple(out)


def _rot3(r):
    cx, sx = math.cos(math.radians(r[0])), math.sin(math.radians(r[0]))
    cy, sy = math.cos(math.radians(r[1])), math.sin(math.radians(r[1]))
    cz, sz = math.cos(math.radians(r[2])), math.sin(math.radians(r[2]))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def affine_3d(params, C, shift) -> np.ndarray:
    p = np.asarray(params, np.float32).astype(np.float64)
    T, R = p[0:3], p[3:6]
    S = p[6:9] if p.size == 12 else np.ones(3)
    H = p[9:12] if p.size == 12 else np.zeros(3)
    Sh = np.array([[1, H[0], H[1]], [0, 1, H[2]], [0, 0, 1]])
    A = _rot3([float(np.float32(v)) for v in R]) @ Sh @ np.diag(S)
    C = np.asarray(C, np.float64)
    t = A @ (np.asarray(shift, np.float64) - C) + C + T
    return np.concatenate([A, t[:, None]], 1).astype(np.float32)


def affine_2d(params, C, shift) -> np.ndarray:
    p = np.asarray(params, np.float32).astype(np.float64)
    T, r = p[0:2], p[2]
    S = p[3:5] if p.size == 6 else np.ones(2)
    h = p[5] if p.size == 6 else 0.0
    c, s = math.cos(math.radians(r)), math.sin(math.radians(r))
    A = np.array([[c, -s], [s, c]]) @ np.array([[1, h], [0, 1]]) @ np.diag(S)
    C = np.asarray(C, np.float64)
    t = A @ (np.asarray(shift, np.float64) - C) + C + T
    M = np.eye(3)
    M[:2, :2], M[:2, 2] = A, t
    return M.astype(np.float32)


def embed_2d(M3) -> np.ndarray:
    """3x3 homogeneous 2D matrix -> the engine's 3x4 layout [[a,b,0,tx],[c,d,0,ty],[0,0,1,0]]."""
    M3 = np.asarray(M3, np.float32)
    return np.array([[M3[0, 0], M3[0, 1], 0, M3[0, 2]], [M3[1, 0], M3[1, 1], 0, M3[1, 2]],
                     [0, 0, 1, 0]], np.float32)


def _pair_consts(fixed_dims, fixed_sp, fixed_org, mov_dims, mov_sp, mov_org, resampled,
                 center_overlaid):
    vd = virtual_dims(fixed_dims, fixed_sp, resampled)
    C = center(fixed_dims, fixed_sp, fixed_org)
    cv = center(vd, resampled, fixed_org)
    cm = center(mov_dims, mov_sp, mov_org)
    shift = (cm - cv) if center_overlaid else np.zeros(len(fixed_dims), np.float32)
    return C, shift


def pose_to_matrix_3d(params, fixed_dims, fixed_sp, fixed_org, mov_dims, mov_sp, mov_org,
                      resampled, center_overlaid=False) -> np.ndarray:
    C, shift = _pair_consts(fixed_dims, fixed_sp, fixed_org, mov_dims, mov_sp, mov_org,
                            resampled, center_overlaid)
    return affine_3d(params, C, shift)


def pose_to_matrix_2d(params, fixed_dims, fixed_sp, fixed_org, mov_dims, mov_sp, mov_org,
                      resampled, center_overlaid=False) -> np.ndarray:
    C, shift = _pair_consts(fixed_dims, fixed_sp, fixed_org, mov_dims, mov_sp, mov_org,
                            resampled, center_overlaid)
    return affine_2d(params, C, shift)


def invert(M) -> np.ndarray:
    """Inverse of an affine matrix (3x4 / 3x3 homogeneous), bottom row forced."""
    M = np.asarray(M, np.float64)
    n = M.shape[1] - 1
    A, t = M[:n, :n], M[:n, n]
    Ai = np.linalg.inv(A)
    out = np.eye(n + 1)
    out[:n, :n], out[:n, n] = Ai, -Ai @ t
    return out.astype(np.float32)
