"""ethersph — MI355X (gfx950) per-step SPH hot path with EtherSPHCPU.jl's API.

Host-side mirror of the reference's types and solve drivers; the per-step
work runs in libethersph_hip.so (HIP kernels behind the C ABI of
include/ethersph.h).
"""
from ._lib import (CUBIC_SPLINE, FP32_BULK, FP64, GAUSSIAN, WENDLAND_C2, WENDLAND_C4, Context, EsphError,
                   default_context)
from .models import (CommonWeaklyCompressibleLiquidModel, ThermalWeaklyCompressibleLiquidModel,
                     WeaklyCompressibleLiquidModel)
from .neighbour import neighborlist
from .particles import (CommonLiquidParticle, CompulsiveWallParticle, LiquidParticles, StaticVelocityWallParticle,
                        ThermalLiquidParticle, ThermostaticWallParticle, WallParticles)
from .solve import AABBCheck, Batch, Scene, solve, solve_
from .sph_kernel import (CubicSpline, Gaussian, SmoothKernel, WendlandC2, WendlandC4, kernelGradient,
                         kernelValue)
from .time_discretization import (DensityReinitializedFixedTimeIntervalForwardEuler, isCheckStep,
                                  isDensityReinitializedStep, isOutputStep)

__all__ = [n for n in dir() if not n.startswith("_")]
