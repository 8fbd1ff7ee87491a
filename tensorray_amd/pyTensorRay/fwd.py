"""Constructors of options and integrators (pyTensorRay/fwd.py), without gin."""
from tensorray_amd import TensorRay as TR


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
    s = TR.Scene()
    s.load_file(filename)
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
