"""Constructors of options and integrators (pyTensorRay/fwd.py), without gin."""
import dataclasses
import os
from typing import List

from tensorray_amd import TensorRay as TR


class _BatchOptions:
    spp_batch = 1
    sppe_batch = 1
    sppse0_batch = 1
    sppse1_batch = 1
    sppe0_batch = 1


class _GuidingOptions:
    g_direct = False
    g_direct_max_size = 100000
    g_direct_spp = 16
    g_eps = 0.0


def RenderBatchOptions(spp_batch=1, sppe_batch=1, sppse0_batch=1, sppse1_batch=1, sppe0_batch=1):
    """Batch sizes to copy into RenderOptions (utils.update_render_batch_options)."""
    o = _BatchOptions()
    o.spp_batch, o.sppe_batch, o.sppse0_batch = spp_batch, sppe_batch, sppse0_batch
    o.sppse1_batch, o.sppe0_batch = sppse1_batch, sppe0_batch
    return o


def GuidingOptions(g_direct=False, g_direct_max_size=100000, g_direct_spp=16, g_eps=0.0):
    """Guiding settings to copy into RenderOptions (utils.update_guiding_options)."""
    o = _GuidingOptions()
    o.g_direct, o.g_direct_max_size, o.g_direct_spp, o.g_eps = g_direct, g_direct_max_size, g_direct_spp, g_eps
    return o


def RenderOptions(seed, max_bounce, spp, sppe, sppse0, sppse1, sppe0, spp_batch=1, sppe_batch=1, sppse0_batch=1,
                  sppse1_batch=1, sppe0_batch=1, g_direct=False, g_direct_max_size=100000, g_direct_spp=16,
                  g_eps=0.0, quiet=False, checkpoint=False):
    o = TR.RenderOptions(seed, max_bounce, spp, sppe, sppse0, sppse1, sppe0)
    o.spp_batch, o.sppe_batch, o.sppse0_batch = spp_batch, sppe_batch, sppse0_batch
    o.sppse1_batch, o.sppe0_batch = sppse1_batch, sppe0_batch
    o.g_direct, o.g_direct_max_size, o.g_direct_spp, o.g_eps = g_direct, g_direct_max_size, g_direct_spp, g_eps
    o.quiet = quiet
    o.checkpoint = checkpoint
    return o


def Scene(filename):
    """Loads a scene file; relative paths inside it resolve against its
    directory (fwd.py:127-135 of the reference changes into it)."""
    cur = os.getcwd()
    d = os.path.dirname(filename)
    try:
        if d:
            os.chdir(d)
        s = TR.Scene()
        s.load_file(os.path.basename(filename))
    finally:
        os.chdir(cur)
    return s


def Path():
    return TR.PathTracer()


def PrimaryEdgeIntegrator():
    return TR.PrimaryEdgeIntegrator()


def DirectEdgeIntegrator():
    return TR.DirectEdgeIntegrator()


def IndirectEdgeIntegrator():
    return TR.IndirectEdgeIntegrator()


def PixelBoundaryIntegrator():
    return TR.PixelBoundaryIntegrator()


@dataclasses.dataclass
class Camera:
    origin: List = dataclasses.field(default_factory=lambda: [1.0, 1.0, 1.0])
    target: List = dataclasses.field(default_factory=lambda: [0.0, 0.0, 0.0])
    up: List = dataclasses.field(default_factory=lambda: [0.0, 0.0, 1.0])
    fov: float = 10.0


class Transformation:
    pass


@dataclasses.dataclass
class Translation(Transformation):
    translation: list
    xform_type = "translate"


@dataclasses.dataclass
class Rotation(Transformation):
    rotation: list
    xform_type = "rotate"


@dataclasses.dataclass
class Transform:
    shape_id: int = -1
    vertex_id: int = -1  # -1: all vertices
    transformation: Transformation = None
