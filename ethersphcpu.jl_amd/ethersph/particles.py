"""Particle types: mirror of Particle/*.jl.

The reference keeps each particle as a mutable struct (an array of heap
objects).  The per-particle classes below keep those names and zero-initialised
fields (LiquidParticle.jl:27-41, :66-101; WallParticle.jl:14-51).  The
``*Particles`` containers hold the same fields as structure-of-arrays, the
form the device path packs; ``solve`` accepts either.
"""
from __future__ import annotations

import numpy as np

from . import _lib


class AbstractParticle:
    pass


class MovableParticle(AbstractParticle):
    pass


class FluidParticle(MovableParticle):
    pass


class LiquidParticle(FluidParticle):
    pass


class FixedParticle(AbstractParticle):
    pass


class WallParticle(FixedParticle):
    pass


class CommonLiquidParticle(LiquidParticle):
    """LiquidParticle.jl:8-41."""

    thermal = False

    def __init__(self, dim: int):
        self.x_vec_ = np.zeros(dim)
        self.v_vec_ = np.zeros(dim)
        self.dv_vec_ = np.zeros(dim)
        self.rho_ = 0.0
        self.drho_ = 0.0
        self.p_ = 0.0
        self.c_ = 0.0
        self.mass_ = 0.0
        self.gap_ = 0.0
        self.sum_kernel_weight_ = 0.0
        self.sum_weighted_scalar_ = 0.0
        self.sum_weighted_vector_ = np.zeros(dim)


class ThermalLiquidParticle(CommonLiquidParticle):
    """LiquidParticle.jl:43-101 (adds t_, dt_, kappa_, cp_)."""

    thermal = True

    def __init__(self, dim: int):
        super().__init__(dim)
        self.t_ = 0.0
        self.dt_ = 0.0
        self.kappa_ = 0.0
        self.cp_ = 0.0


class CompulsiveWallParticle(WallParticle):
    """WallParticle.jl:8-19."""

    kind = _lib.WALL_COMPULSIVE

    def __init__(self, dim: int):
        self.x_vec_ = np.zeros(dim)
        self.normal_vec_ = np.zeros(dim)
        self.gap_ = 0.0


class StaticVelocityWallParticle(CompulsiveWallParticle):
    """WallParticle.jl:21-34 (adds v_vec_)."""

    kind = _lib.WALL_STATIC_VELOCITY

    def __init__(self, dim: int):
        super().__init__(dim)
        self.v_vec_ = np.zeros(dim)


class ThermostaticWallParticle(CompulsiveWallParticle):
    """WallParticle.jl:36-51 (adds t_, kappa_)."""

    kind = _lib.WALL_THERMOSTATIC

    def __init__(self, dim: int):
        super().__init__(dim)
        self.t_ = 0.0
        self.kappa_ = 0.0


# ---- structure-of-arrays containers -------------------------------------------

class LiquidParticles:
    """n fluid particles as arrays: x_vec_ (n, dim), v_vec_ (n, dim), rho_, p_,
    c_, mass_, gap_ and, when thermal, t_, kappa_, cp_."""

    def __init__(self, dim: int, n: int, thermal: bool = False):
        self.dim = dim
        self.thermal = thermal
        self.x_vec_ = np.zeros((n, dim))
        self.v_vec_ = np.zeros((n, dim))
        self.rho_ = np.zeros(n)
        self.p_ = np.zeros(n)
        self.c_ = np.zeros(n)
        self.mass_ = np.zeros(n)
        self.gap_ = np.zeros(n)
        if thermal:
            self.t_ = np.zeros(n)
            self.kappa_ = np.zeros(n)
            self.cp_ = np.zeros(n)

    def __len__(self):
        return self.x_vec_.shape[0]

    SCALARS = ("rho_", "p_", "c_", "mass_", "gap_")
    THERMAL = ("t_", "kappa_", "cp_")

    @classmethod
    def from_particles(cls, ps) -> "LiquidParticles":
        ps = list(ps)
        dim = len(ps[0].x_vec_) if ps else 2
        thermal = bool(ps) and getattr(ps[0], "thermal", False)
        a = cls(dim, len(ps), thermal)
        if ps:
            a.x_vec_[:] = [p.x_vec_ for p in ps]
            a.v_vec_[:] = [p.v_vec_ for p in ps]
            for f in cls.SCALARS + (cls.THERMAL if thermal else ()):
                getattr(a, f)[:] = [getattr(p, f) for p in ps]
        return a

    def write_back(self, ps: list) -> None:
        """Writes the arrays into particle objects (len(ps) == len(self))."""
        for i, p in enumerate(ps):
            p.x_vec_[:] = self.x_vec_[i]
            p.v_vec_[:] = self.v_vec_[i]
            for f in self.SCALARS + (self.THERMAL if self.thermal else ()):
                setattr(p, f, float(getattr(self, f)[i]))

    def subset(self, idx) -> "LiquidParticles":
        a = LiquidParticles(self.dim, len(idx), self.thermal)
        for f in ("x_vec_", "v_vec_") + self.SCALARS + (self.THERMAL if self.thermal else ()):
            getattr(a, f)[:] = getattr(self, f)[idx]
        return a


class WallParticles:
    """A group of wall particles of one kind as arrays (x_vec_, normal_vec_,
    gap_, and v_vec_ or t_/kappa_ as the kind has them)."""

    def __init__(self, kind: int, dim: int, n: int):
        self.kind = kind
        self.dim = dim
        self.x_vec_ = np.zeros((n, dim))
        self.normal_vec_ = np.zeros((n, dim))
        self.gap_ = np.zeros(n)
        self.v_vec_ = np.zeros((n, dim))
        self.t_ = np.zeros(n)
        self.kappa_ = np.zeros(n)

    def __len__(self):
        return self.x_vec_.shape[0]

    @classmethod
    def from_particles(cls, ps) -> "WallParticles":
        ps = list(ps)
        kinds = {getattr(p, "kind", _lib.WALL_COMPULSIVE) for p in ps}
        if len(kinds) > 1:
            raise TypeError("a wall group holds one wall particle type")
        kind = kinds.pop() if kinds else _lib.WALL_COMPULSIVE
        dim = len(ps[0].x_vec_) if ps else 2
        a = cls(kind, dim, len(ps))
        if ps:
            a.x_vec_[:] = [p.x_vec_ for p in ps]
            a.normal_vec_[:] = [p.normal_vec_ for p in ps]
            a.gap_[:] = [p.gap_ for p in ps]
            if kind == _lib.WALL_STATIC_VELOCITY:
                a.v_vec_[:] = [p.v_vec_ for p in ps]
            if kind == _lib.WALL_THERMOSTATIC:
                a.t_[:] = [p.t_ for p in ps]
                a.kappa_[:] = [p.kappa_ for p in ps]
        return a

    @staticmethod
    def concat(groups) -> "WallParticles":
        groups = list(groups)
        kinds = {g.kind for g in groups}
        if len(kinds) != 1:
            raise TypeError("concat needs groups of one kind")
        dim = groups[0].dim
        a = WallParticles(groups[0].kind, dim, sum(len(g) for g in groups))
        for f in ("x_vec_", "normal_vec_", "gap_", "v_vec_", "t_", "kappa_"):
            setattr(a, f, np.concatenate([getattr(g, f) for g in groups]))
        return a
