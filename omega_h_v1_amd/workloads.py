"""Synthetic workloads of the adapt hot path (host-side field evaluation).

Fields are evaluated on the host in float64 with numpy from the current
vertex coordinates, exactly as the reference evaluates analytic fields on
the host (eval_field.cpp:15-49); the arrays are then handed to the mesh.
"""
import math

import numpy as np


def sine_warp_2d(coords, amp=0.1):
    """u(x, y) = amp*sin(2 pi x)*cos(2 pi y) in both x and y components."""
    x, y = coords[:, 0], coords[:, 1]
    u = amp * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    w = np.zeros_like(coords)
    w[:, 0] = u
    w[:, 1] = u
    return w


def rotation_warp(coords, rotation, mid=(0.5, 0.5, 0.0), flatten_z=True):
    """The swirl of test_warp.cpp:34-50 / test_warp_perf.cpp:47-64: points
    within radius 0.5 of the axis rotate by rotation*2*(0.5-r)."""
    x = coords - np.asarray(mid)[None, :]
    if flatten_z:
        x[:, 2] = 0
    polar_a = np.arctan2(x[:, 1], x[:, 0])
    polar_r = np.sqrt(np.sum(x * x, axis=1))
    rot_a = np.where(polar_r < 0.5, rotation * (2 * (0.5 - polar_r)), 0.0)
    dest_a = polar_a + rot_a
    dst = np.zeros_like(x)
    dst[:, 0] = np.cos(dest_a) * polar_r
    dst[:, 1] = np.sin(dest_a) * polar_r
    return dst - x


def uniform_size(nverts, h):
    return np.full(nverts, float(h))


PI_REF = 3.14159265359  # algebra.hpp:8


def tanh_front_hessian(coords, a=50.0, s0=1.5, dim=3):
    """Hessian of f = tanh(a (x + y [+ z] - s0)): f''(s) times the all-ones
    matrix over the active coordinates, as 6 components per vertex."""
    s = coords[:, 0] + coords[:, 1] + (coords[:, 2] if dim == 3 else 0.0) - s0
    th = np.tanh(a * s)
    f2 = -2.0 * a * a * th * (1.0 - th * th)
    h = np.zeros((len(coords), 6))
    if dim == 3:
        h[:] = f2[:, None]  # all second partials are f''
    else:
        h[:, 0] = f2  # xx
        h[:, 1] = f2  # xy
        h[:, 3] = f2  # yy
    return h


def rotating_metric(coords, h_short, aspect, dim=3):
    """Metric whose short axis (size h_short) lies in the xy plane at angle
    2 pi z, long axes aspect * h_short (6 components per vertex)."""
    ang = 2.0 * np.pi * coords[:, 2] if dim == 3 else 2.0 * np.pi * coords[:, 0]
    c, s = np.cos(ang), np.sin(ang)
    l1 = 1.0 / (h_short * h_short)
    l2 = 1.0 / ((aspect * h_short) ** 2)
    m = np.zeros((len(coords), 6))
    m[:, 0] = l1 * c * c + l2 * s * s
    m[:, 1] = (l1 - l2) * c * s
    m[:, 3] = l1 * s * s + l2 * c * c
    m[:, 5] = l2
    return m


def random_rotation(rng):
    """Uniform rotation on SO(3): a normalised 4D Gaussian quaternion."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def constant_metric(nverts, rot, lengths):
    """The same SPD metric R diag(1/h^2) R^T at every vertex (6 components)."""
    m = rot @ np.diag(1.0 / np.asarray(lengths, dtype=np.float64) ** 2) @ rot.T
    row = np.array([m[0, 0], m[0, 1], m[0, 2], m[1, 1], m[1, 2], m[2, 2]])
    return np.tile(row, (nverts, 1))


def batch_specs(count, seed, n_lo=8, n_hi=32, h_lo=0.005, h_hi=0.1):
    """Meshes of a batch: cubes per axis n log-uniform in [n_lo, n_hi]
    (rounded), a uniform random rotation and three axis lengths log-uniform
    in [h_lo, h_hi] per mesh, from numpy's PCG64 generator seeded with seed.
    Each mesh is the Kuhn box [0, L]^3 with L = n (h1 h2 h3)^(1/3), so its
    mesh spacing is the metric's mean length."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        rot = random_rotation(rng)
        h = np.exp(rng.uniform(np.log(h_lo), np.log(h_hi), size=3))
        out.append({"n": n, "rot": rot, "lengths": h, "side": n * float(np.prod(h)) ** (1.0 / 3.0)})
    return out
