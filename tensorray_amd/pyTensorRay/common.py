"""Pointer helpers (pyTensorRay/common.py of the reference)."""
from pathlib import Path

from tensorray_amd import TensorRay as TR


def get_ptr(t):
    """Address of a float tensor's storage (host or device)."""
    return TR.float_ptr(t.data_ptr())


def get_int_ptr(t):
    return TR.int_ptr(t.data_ptr())


def mkdir(path):
    Path(path).mkdir(parents=True, exist_ok=True)
