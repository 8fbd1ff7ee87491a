"""Image losses of the inverse-rendering examples (pyTensorRay/loss.py)."""
import math

import torch


def compute_render_loss_L2(img, target_img, weight):
    d = (img - target_img).nan_to_num(nan=0)
    return 0.5 * torch.square(d).sum() * weight


def compute_render_loss_L1(img, target_img, weight):
    d = (img - target_img).nan_to_num(nan=0)
    return d.abs().sum() * weight


def downsample(x):
    """2x2 box downsampling of an (H, W, C) image (odd sizes repeat the last row / column)."""
    if x.size(0) % 2 == 1:
        x = torch.cat((x, x[-1:, :]), dim=0)
    if x.size(1) % 2 == 1:
        x = torch.cat((x, x[:, -1:]), dim=1)
    return (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2]) * 0.25


def build_pyramid(img):
    levels = min(5, int(min(math.log2(img.shape[0]), math.log2(img.shape[1]))) + 1)
    out = [img]
    for _ in range(levels - 1):
        out.append(downsample(out[-1]))
    return out


def compute_render_loss_pyramid_L1(img, target_pyramid, weight):
    return sum(compute_render_loss_L1(a, b, weight) * (4.0 ** i)
               for i, (a, b) in enumerate(zip(build_pyramid(img), target_pyramid)))


def compute_render_loss_pyramid_L2(img, target_pyramid, weight):
    return sum(compute_render_loss_L2(a, b, weight) * (4.0 ** i)
               for i, (a, b) in enumerate(zip(build_pyramid(img), target_pyramid)))
