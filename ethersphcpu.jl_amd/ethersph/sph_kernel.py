"""Smoothing kernels: the host-side mirror of SPHKernel/SmoothKernel/*.jl.

``SmoothKernel(R, dim, CubicSpline)`` builds the kernel record exactly as
SmoothKernel.jl:37-48 (h = R / radius_ratio, sigma = c_dim / h^dim,
kernel_0 = sigma).  ``kernelValue`` / ``kernelGradient`` evaluate W and dW/dr
on the GPU through ``esph_kernel_eval``.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib


class SmoothKernel_:
    """Common fields of the reference's kernel structs (CubicSpline.jl:8-15)."""

    kind: int = -1
    radius_ratio: float = 2.0
    c_dim: tuple = ()

    def __init__(self, h_, dim_, radius_ratio_, influence_radius_, sigma_, kernel_0_):
        self.h_ = h_
        self.dim_ = dim_
        self.radius_ratio_ = radius_ratio_
        self.influence_radius_ = influence_radius_
        self.sigma_ = sigma_
        self.kernel_0_ = kernel_0_

    def __repr__(self):
        return (f"{type(self).__name__}(h_={self.h_!r}, dim_={self.dim_}, radius_ratio_={self.radius_ratio_}, "
                f"influence_radius_={self.influence_radius_!r}, sigma_={self.sigma_!r})")


# kSmoothKernelParametersDict, SmoothKernel.jl:15-20
class CubicSpline(SmoothKernel_):
    kind = _lib.CUBIC_SPLINE
    radius_ratio = 2.0
    c_dim = (2.0 / 3.0, 10.0 / 7.0 / math.pi, 1.0 / math.pi)


class Gaussian(SmoothKernel_):
    kind = _lib.GAUSSIAN
    radius_ratio = 3.0
    c_dim = (1.0 / math.sqrt(math.pi), 1.0 / math.pi, 1.0 / math.sqrt(math.pi ** 3))


class WendlandC2(SmoothKernel_):
    kind = _lib.WENDLAND_C2
    radius_ratio = 2.0
    c_dim = (0.0, 7.0 / 4.0 / math.pi, 21.0 / 16.0 / math.pi)


class WendlandC4(SmoothKernel_):
    kind = _lib.WENDLAND_C4
    radius_ratio = 2.0
    c_dim = (5.0 / 8.0, 9.0 / 4.0 / math.pi, 495.0 / 256.0 / math.pi)


KERNELS = {"CubicSpline": CubicSpline, "Gaussian": Gaussian, "WendlandC2": WendlandC2, "WendlandC4": WendlandC4}


def SmoothKernel(influence_radius: float, dim: int, smooth_kernel: type) -> SmoothKernel_:
    """SmoothKernel(R, dim, K), SmoothKernel.jl:37-48."""
    assert dim in (1, 2, 3)
    ratio = smooth_kernel.radius_ratio
    h = influence_radius / ratio
    hd = h if dim == 1 else (h * h if dim == 2 else h * h * h)
    sigma = smooth_kernel.c_dim[dim - 1] / hd
    return smooth_kernel(h, dim, ratio, influence_radius, sigma, sigma)


def _eval(r, kernel: SmoothKernel_, ctx=None):
    ctx = ctx or _lib.default_context()
    scalar = np.ndim(r) == 0
    ra = _lib.f64(np.atleast_1d(r)).ravel()
    w = np.empty_like(ra)
    dw = np.empty_like(ra)
    ctx.check(ctx._L.esph_kernel_eval(ctx.handle, kernel.kind, kernel.dim_, kernel.influence_radius_, ra.size,
                                      _lib.dptr(ra), _lib.dptr(w), _lib.dptr(dw)))
    if scalar:
        return float(w[0]), float(dw[0])
    return w.reshape(np.shape(r)), dw.reshape(np.shape(r))


def kernelValue(r, kernel: SmoothKernel_, ctx=None):
    """W(r) on the device (CubicSpline.jl:39-48 and the other kernels)."""
    return _eval(r, kernel, ctx)[0]


def kernelGradient(r, kernel: SmoothKernel_, ctx=None):
    """dW/dr on the device (CubicSpline.jl:64-73 and the other kernels)."""
    return _eval(r, kernel, ctx)[1]
