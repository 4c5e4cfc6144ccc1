"""Synthetic rendering workloads (seeded, identical on every host).

Random numbers come from a SplitMix64 stream: value i of a stream with seed
s is mix(s + (i+1)*0x9E3779B97F4A7C15), and a uniform in [0,1) is the top 53
bits times 2^-53.  Volumes are sums of isotropic Gaussian blobs evaluated at
voxel centres p=(i,j,k)/(size-1) in [0,1]^3 and clamped to [0,1].

Conventions (DiffDVR): box [-0.5,0.5]^3 (box_min=-0.5, box_size=1,
pytests/tests/tf/test_tf_optimization.py:94-95); a step of f voxels is f/X in
world units (pytests/diffdvr/settings.py:181); sphere cameras are
viewport_from_sphere(center 0, yaw=lon, pitch=lat, r, Orientation.Yp) with a
45 degree vertical field of view, rendered in ReferenceFrame mode;
fibonacci_sphere follows pytests/diffdvr/utils.py:76-89.

The five workloads:
  C1 64^3 x 8 blobs, linear TF (5 points), 1 view 128^2, step 0.5 voxel, forward only
  C2 128^3 x 32 blobs, 64x4 texture TF, 8 views 256^2 at lat 20, step 0.25 voxel, L1 vs a 2nd TF
  C3 256^3 x 64 blobs + noise 0.02, identity TF x10, 64 Fibonacci views 256^2, step 0.5 voxel, L2 vs a 2nd volume
  C4 128^3 blobs, 4-component Gaussian TF, 512 views of sizes {64..512}^2, camera gradients
  C5 16 volumes of sizes {64,96,128,192,256}^3, 4-32 views each of 128..512, joint density + TF
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, n: int) -> np.ndarray:
    """Values start..start+n-1 of the SplitMix64 stream with this seed (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Sequential uniform draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        v = splitmix64(self.seed, self.pos, n)
        self.pos += n
        u = (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def choice(self, values, n: int) -> np.ndarray:
        idx = np.minimum((self.uniform(n) * len(values)).astype(np.int64), len(values) - 1)
        return np.asarray(values)[idx]


# ---------------------------------------------------------------------------
# volumes and transfer functions
# ---------------------------------------------------------------------------
def blob_params(rng: Rng, n_blobs: int):
    c = rng.uniform(3 * n_blobs, 0.2, 0.8).reshape(n_blobs, 3)
    s = rng.uniform(n_blobs, 0.05, 0.15)
    a = rng.uniform(n_blobs, 0.3, 1.0)
    return c, s, a


def blob_volume(size, n_blobs: int, seed: int, noise: float = 0.0, device="cpu") -> torch.Tensor:
    """1 x X x Y x Z float32 density: clamp(sum_k a_k exp(-|p-c_k|^2/(2 s_k^2)) + noise, 0, 1)."""
    X, Y, Z = (size, size, size) if isinstance(size, int) else size
    rng = Rng(seed)
    c, s, a = blob_params(rng, n_blobs)
    dev = torch.device(device)
    px = torch.arange(X, dtype=torch.float64, device=dev) / (X - 1)
    py = torch.arange(Y, dtype=torch.float64, device=dev) / (Y - 1)
    pz = torch.arange(Z, dtype=torch.float64, device=dev) / (Z - 1)
    vol = torch.zeros(X, Y, Z, dtype=torch.float64, device=dev)
    for k in range(n_blobs):  # separable Gaussian: outer product of three 1-D factors
        inv = 1.0 / (2.0 * s[k] * s[k])
        gx = a[k] * torch.exp(-(px - c[k, 0]) ** 2 * inv)
        gy = torch.exp(-(py - c[k, 1]) ** 2 * inv)
        gz = torch.exp(-(pz - c[k, 2]) ** 2 * inv)
        vol += gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    if noise > 0:
        nz = rng.uniform(X * Y * Z, -noise, noise).reshape(X, Y, Z)
        vol += torch.from_numpy(nz).to(dev)
    return vol.clamp_(0.0, 1.0).to(torch.float32).unsqueeze(0)


def linear_tf(n_points: int, seed: int) -> torch.Tensor:
    """1 x R x 5: positions {0, sorted U(0.1,0.9), 1}, RGBA U[0,1] (unittests/testAdjointTF.cpp:36-55)."""
    rng = Rng(seed)
    inner = np.sort(rng.uniform(n_points - 2, 0.1, 0.9))
    pos = np.concatenate([[0.0], inner, [1.0]])
    rgba = rng.uniform(4 * n_points).reshape(n_points, 4)
    return torch.from_numpy(np.concatenate([rgba, pos[:, None]], axis=1)).float().unsqueeze(0)


def texture_tf(R: int, seed: int) -> torch.Tensor:
    """1 x R x 4, entries U[0,1]."""
    return torch.from_numpy(Rng(seed).uniform(4 * R).reshape(R, 4)).float().unsqueeze(0)


def gaussian_tf(n: int, seed: int) -> torch.Tensor:
    """1 x n x 6: RGBA U[0,1], mean U[0,1], sigma U[0.02,0.1] (column 5 is sigma)."""
    rng = Rng(seed)
    rgba = rng.uniform(4 * n).reshape(n, 4)
    mu = rng.uniform(n)
    sg = rng.uniform(n, 0.02, 0.1)
    return torch.from_numpy(np.concatenate([rgba, mu[:, None], sg[:, None]], axis=1)).float().unsqueeze(0)


def identity_tf(absorption_scale: float, color_scale: float) -> torch.Tensor:
    return torch.tensor([[[absorption_scale, color_scale]]], dtype=torch.float32)


def fibonacci_sphere(N: int, dtype=np.float32):
    """(latitude, longitude) of N points on a Fibonacci spiral (pytests/diffdvr/utils.py:76-89)."""
    gr = (np.sqrt(5.0) + 1.0) / 2.0
    ga = (2 - gr) * (2 * np.pi)
    i = np.arange(1, N + 1, dtype=dtype)
    lat = np.arcsin(-1 + 2 * i / (N + 1))
    lon = np.remainder(ga * i, 2 * np.pi)
    return lat, lon


def sphere_viewports(lat_rad, lon_rad, radius: float = 2.5, device="cpu") -> torch.Tensor:
    """B x 3 x 3 float32 viewports of cameras on a sphere around the origin (Orientation.Yp)."""
    from . import pyrenderer as pr

    lat = torch.as_tensor(np.asarray(lat_rad, dtype=np.float64)).reshape(-1, 1)
    lon = torch.as_tensor(np.asarray(lon_rad, dtype=np.float64)).reshape(-1, 1)
    B = lat.shape[0]
    vp = pr.Camera.viewport_from_sphere(torch.zeros(B, 3, dtype=torch.float64), lon, lat,
                                        torch.full((B, 1), float(radius), dtype=torch.float64), pr.Orientation.Yp)
    return vp.to(torch.float32).to(device)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
FOV_Y = math.radians(45.0)


@dataclass
class Workload:
    """One volume, B views of equal size, one TF (broadcast over the views)."""
    name: str
    volume: torch.Tensor              # 1 x X x Y x Z
    tf: torch.Tensor                  # 1 x R x C
    tf_mode: int                      # pyrenderer.TFMode value
    viewports: torch.Tensor           # B x 3 x 3
    width: int
    height: int
    step: float                       # world units
    filter: int = 1                   # trilinear
    blend: int = 0                    # Beer-Lambert
    fov_y: float = FOV_Y
    loss: str | None = None           # "l1" / "l2"
    target_volume: torch.Tensor | None = None
    target_tf: torch.Tensor | None = None
    grads: tuple = ()                 # subset of ("tf", "volume", "camera")
    seeds: dict = field(default_factory=dict)

    @property
    def B(self):
        return self.viewports.shape[0]

    def to(self, device):
        for k in ("volume", "tf", "viewports", "target_volume", "target_tf"):
            v = getattr(self, k)
            if v is not None:
                setattr(self, k, v.to(device))
        return self


def workload_c1(device="cpu") -> Workload:
    seed = 1001
    vol = blob_volume(64, 8, seed, device=device)
    tf = linear_tf(5, seed + 1)
    vp = sphere_viewports([math.radians(30.0)], [math.radians(45.0)], 2.5)
    return Workload("C1", vol, tf, 2, vp, 128, 128, 0.5 / 64, seeds={"volume": seed, "tf": seed + 1}).to(device)


def workload_c2(device="cpu", views: int = 8, size: int = 256) -> Workload:
    seed = 2002
    vol = blob_volume(128, 32, seed, device=device)
    tf = texture_tf(64, seed + 1)
    tf_target = texture_tf(64, seed + 2)
    lon = [2 * math.pi * j / views for j in range(views)]
    vp = sphere_viewports([math.radians(20.0)] * views, lon, 2.5)
    return Workload("C2", vol, tf, 1, vp, size, size, 0.25 / 128, loss="l1", target_tf=tf_target, grads=("tf",),
                    seeds={"volume": seed, "tf": seed + 1, "target_tf": seed + 2}).to(device)


def workload_c3(device="cpu", views: int = 64, size: int = 256, res: int = 256) -> Workload:
    seed = 3003
    vol = blob_volume(res, 64, seed, noise=0.02, device=device)
    target = blob_volume(res, 64, seed + 7, noise=0.02, device=device)
    lat, lon = fibonacci_sphere(views)
    vp = sphere_viewports(lat, lon, 2.5)
    return Workload("C3", vol, identity_tf(10.0, 10.0), 0, vp, size, size, 0.5 / res, loss="l2",
                    target_volume=target, grads=("volume",),
                    seeds={"volume": seed, "target_volume": seed + 7}).to(device)


@dataclass
class CameraBatch:
    """Sphere cameras of uneven image sizes around the origin (Orientation.Yp)."""
    lat: torch.Tensor                 # N x 1 float64 (pitch)
    lon: torch.Tensor                 # N x 1 float64 (yaw)
    radius: torch.Tensor              # N x 1 float64
    sizes: list                       # [(W, H)] per camera

    def viewports(self, device="cpu", dtype=torch.float32) -> torch.Tensor:
        from . import pyrenderer as pr
        vp = pr.Camera.viewport_from_sphere(torch.zeros(self.lat.shape[0], 3, dtype=torch.float64), self.lon,
                                            self.lat, self.radius, pr.Orientation.Yp)
        return vp.to(dtype).to(device)


@dataclass
class MultiWorkload:
    """Several volumes, each with its own TF and cameras of uneven image sizes."""
    name: str
    volumes: list                     # [X x Y x Z] float32
    tfs: list                         # [R x C] float32
    tf_mode: int
    cameras: list                     # [CameraBatch] per volume
    steps: list                       # world step per volume
    target_volumes: list | None = None
    target_cameras: list | None = None
    loss: str | None = None
    grads: tuple = ()
    filter: int = 1
    blend: int = 0
    fov_y: float = FOV_Y
    seeds: dict = field(default_factory=dict)


def workload_c4(device="cpu", cameras: int = 512) -> MultiWorkload:
    """C4: 128^3 blobs, 4-component Gaussian TF, `cameras` sphere cameras with seeded
    latitude / longitude and image sizes from {64,128,256,512}^2; gradient w.r.t. every
    camera's latitude, longitude and radius (L2 against images from perturbed cameras)."""
    seed = 4004
    vol = blob_volume(128, 32, seed, device=device)[0]
    tf = gaussian_tf(4, seed + 1)[0].to(device)
    rng = Rng(seed + 2)
    lat = np.arcsin(rng.uniform(cameras, -0.95, 0.95))
    lon = rng.uniform(cameras, 0.0, 2.0 * math.pi)
    side = rng.choice([64, 128, 256, 512], cameras)
    dlat = rng.uniform(cameras, -0.05, 0.05)
    dlon = rng.uniform(cameras, -0.05, 0.05)
    drad = rng.uniform(cameras, -0.1, 0.1)
    col = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).reshape(-1, 1)
    sizes = [(int(s), int(s)) for s in side]
    cams = CameraBatch(col(lat), col(lon), col(np.full(cameras, 2.5)), sizes)
    target = CameraBatch(col(lat + dlat), col(lon + dlon), col(2.5 + drad), sizes)
    return MultiWorkload("C4", [vol], [tf], 3, [cams], [0.5 / 128], target_cameras=[target], loss="l2",
                         grads=("camera",), seeds={"volume": seed, "tf": seed + 1, "cameras": seed + 2})


def workload_c5(device="cpu", volumes: int = 16) -> MultiWorkload:
    """C5: `volumes` blob volumes of seeded sizes {64,96,128,192,256}^3, each with its own
    5-point linear TF and 4-32 sphere cameras at seeded square resolutions 128..512
    (multiples of 32), step 0.5 voxel; L2 against images of a second seeded volume
    through the same TF; gradient w.r.t. density and TF jointly."""
    seed = 5005
    rng = Rng(seed)
    sides = rng.choice([64, 96, 128, 192, 256], volumes)
    vols, tgts, tfs, cams, steps = [], [], [], [], []
    for v in range(volumes):
        n = int(sides[v])
        vols.append(blob_volume(n, 16, seed + 100 * v + 1, device=device)[0])
        tgts.append(blob_volume(n, 16, seed + 100 * v + 2, device=device)[0])
        tfs.append(linear_tf(5, seed + 100 * v + 3)[0].to(device))
        nc = int(rng.choice(np.arange(4, 33), 1)[0])
        lat = np.arcsin(rng.uniform(nc, -0.95, 0.95))
        lon = rng.uniform(nc, 0.0, 2.0 * math.pi)
        res = rng.choice(np.arange(128, 513, 32), nc)
        col = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).reshape(-1, 1)
        cams.append(CameraBatch(col(lat), col(lon), col(np.full(nc, 2.5)), [(int(r), int(r)) for r in res]))
        steps.append(0.5 / n)
    return MultiWorkload("C5", vols, tfs, 2, cams, steps, target_volumes=tgts, loss="l2", grads=("volume", "tf"),
                         seeds={"base": seed})


__all__ = ["Rng", "splitmix64", "blob_volume", "linear_tf", "texture_tf", "gaussian_tf", "identity_tf",
           "fibonacci_sphere", "sphere_viewports", "Workload", "workload_c1", "workload_c2", "workload_c3", "FOV_Y",
           "CameraBatch", "MultiWorkload", "workload_c4", "workload_c5"]
