"""Settings files of the renderer GUI (version 2 JSON), as pytests/diffdvr/settings.py:14-184.

``Settings(file)`` reads the dataset (a .cvol file or an implicit dataset), the
camera on a sphere, the TF editor state (piecewise-linear, texture or Gaussian
TF) and the step size; ``setup_default_settings`` builds ``RendererInputs`` for a
volume the way the GUI does (renderer box centred on the origin, half a voxel
larger than the grid).
"""
from __future__ import annotations

import json
import os
from typing import NamedTuple, Optional

import numpy as np
import torch

from . import pyrenderer


class Settings:

    def __init__(self, file):
        self._filepath = os.path.split(file)[0]
        with open(file) as fb:
            self._data = json.load(fb)
        if self._data["version"] != 2:
            raise Exception("incorrect file version, expected 2 but got " + str(self._data["version"]))

    # -- dataset ------------------------------------------------------------
    def is_synthetic_dataset(self) -> bool:
        return self._data["dataset"]["syntheticDataset"] >= 0

    def get_synthetic_dataset_type(self) -> pyrenderer.ImplicitEquation:
        assert self.is_synthetic_dataset()
        return pyrenderer.ImplicitEquation(int(self._data["dataset"]["syntheticDataset"]))

    def get_dataset_name(self) -> str:
        if self.is_synthetic_dataset():
            return self.get_synthetic_dataset_type().name
        return os.path.split(self._data["dataset"]["file"])[-1]

    def dataset_path(self) -> str:
        """Absolute path of a file dataset (relative paths are relative to the settings file)."""
        path = self._data["dataset"]["file"].replace("\\", "/")
        if not os.path.isabs(path):
            path = os.path.abspath(os.path.join(self._filepath, path))
        return path

    def load_dataset(self, resolution: Optional[int] = None) -> pyrenderer.Volume:
        """The dataset of the settings file, on the CPU; `resolution` overrides the synthetic one."""
        if self.is_synthetic_dataset():
            if resolution is None:
                resolution = 1 << int(self._data["dataset"]["syntheticDatasetResolutionPower"])
            return pyrenderer.Volume.create_implicit(self.get_synthetic_dataset_type(), resolution)
        return pyrenderer.Volume(self.dataset_path())

    # -- camera -------------------------------------------------------------
    class CameraConfig(NamedTuple):
        pitch_radians: float
        yaw_radians: float
        fov_y_radians: float
        center: np.ndarray  # shape (3,)
        distance: float
        orientation: pyrenderer.Orientation

    def get_camera(self) -> "Settings.CameraConfig":
        """Camera on a sphere (for parametrizations.CameraOnASphere); distance = zoomSpeed ** zoomValue."""
        c = self._data["camera"]
        return Settings.CameraConfig(np.deg2rad(c["currentPitch"]), np.deg2rad(c["currentYaw"]),
                                     np.deg2rad(c["fov"]), np.array(c["lookAt"]),
                                     1.0 * (c["zoomSpeed"] ** c["zoomValue"]),
                                     pyrenderer.Orientation(int(c["orientation"])))

    # -- transfer functions ---------------------------------------------------
    def _tf_range(self, min_density, max_density, opacity_scaling):
        e = self._data["tfEditor"]
        return (e["minDensity"] if min_density is None else min_density,
                e["maxDensity"] if max_density is None else max_density,
                e["opacityScaling"] if opacity_scaling is None else opacity_scaling)

    def get_tf_points(self, min_density: Optional[float] = None, max_density: Optional[float] = None,
                      opacity_scaling: Optional[float] = None, purge_zero_regions: bool = True):
        """Control points of the piecewise-linear editor, for TFUtils.get_piecewise_tensor /
        get_texture_tensor."""
        mn, mx, sc = self._tf_range(min_density, max_density, opacity_scaling)
        e = self._data["tfEditor"]
        g = e["editor"] if "editor" in e else e["editorLinear"]
        return pyrenderer.TFUtils.assemble_from_settings(
            [pyrenderer.real3(v[0], v[1], v[2]) for v in g["colorAxis"]], g["densityAxisColor"], g["opacityAxis"],
            g["densityAxisOpacity"], mn, mx, sc, purge_zero_regions)

    def get_tf_points_texture(self, min_density: Optional[float] = None, max_density: Optional[float] = None,
                              opacity_scaling: Optional[float] = None):
        """Control points of the texture editor (opacities evenly spaced on [0, 1])."""
        mn, mx, sc = self._tf_range(min_density, max_density, opacity_scaling)
        g = self._data["tfEditor"]["editorTexture"]
        opacity = g["opacities"]
        return pyrenderer.TFUtils.assemble_from_settings(
            [pyrenderer.real3(v[0], v[1], v[2]) for v in g["colorAxis"]], g["densityAxisColor"], opacity,
            list(np.linspace(0, 1, len(opacity), endpoint=True)), mn, mx, sc, False)

    def get_gaussian_tensor(self, min_density: Optional[float] = None, max_density: Optional[float] = None,
                            opacity_scaling: Optional[float] = None) -> torch.Tensor:
        """1*R*6 Gaussian TF (r, g, b, opacity * scaling, mean, sigma) in density units."""
        mn, mx, sc = self._tf_range(min_density, max_density, opacity_scaling)
        g = self._data["tfEditor"]["editorGaussian"]
        tf = np.empty((1, len(g), 6), dtype=np.float32)
        for r, row in enumerate(g):
            tf[0, r, 0:3] = row[0:3]
            tf[0, r, 3] = row[3] * sc
            tf[0, r, 4] = mn + row[4] * (mx - mn)
            tf[0, r, 5] = row[5] * (mx - mn)
        return torch.from_numpy(tf)

    def get_stepsize(self) -> float:
        return self._data["renderer"]["stepsize"]


def setup_default_settings(volume: pyrenderer.Volume, screen_width: int, screen_height: int,
                           step_size: float = 0.1, run_on_cuda: bool = True) -> pyrenderer.RendererInputs:
    """RendererInputs for `volume` as the GUI's setupRendererArgs: box of the volume's world size centred
    on the origin (shifted by half a voxel), trilinear, Beer-Lambert, step = step_size / X."""
    volume_data = volume.getDataGpu(0) if run_on_cuda else volume.getDataCpu(0)
    ws, res = volume.world_size, volume.resolution
    box_size = np.array([ws.x, ws.y, ws.z])
    resolution = np.array([res.x, res.y, res.z])
    voxel_size = box_size / (resolution - np.array([1, 1, 1]))
    box_min = (-box_size * 0.5) - (voxel_size * 0.5)
    s = pyrenderer.RendererInputs()
    s.volume = volume_data
    s.volume_filter_mode = pyrenderer.VolumeFilterMode.Trilinear
    s.screen_size = pyrenderer.int2(screen_width, screen_height)
    s.box_min = pyrenderer.real3(*[float(v) for v in box_min])
    s.box_size = pyrenderer.real3(*[float(v) for v in box_size])
    s.step_size = float(step_size / resolution[0])
    s.blend_mode = pyrenderer.BlendMode.BeerLambert
    return s
