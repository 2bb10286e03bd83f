"""Equation models: mirror of EquationModel/LiquidModel.jl:10-66."""
from __future__ import annotations

import numpy as np


class WeaklyCompressibleLiquidModel:
    pass


class CommonWeaklyCompressibleLiquidModel(WeaklyCompressibleLiquidModel):
    """LiquidModel.jl:10-33; b = c0^2 rho0 / gamma, nu0 = mu0 / rho0."""

    def __init__(self, rho_0, c_0, p_0, gamma: int, mu_0, body_force_vec):
        self.rho_0_ = float(rho_0)
        self.c_0_ = float(c_0)
        self.p_0_ = float(p_0)
        self.gamma_ = int(gamma)
        self.b_ = self.c_0_ * self.c_0_ * self.rho_0_ / self.gamma_
        self.mu_0_ = float(mu_0)
        self.nu_0_ = self.mu_0_ / self.rho_0_
        self.body_force_vec_ = np.asarray(body_force_vec, dtype=np.float64)


class ThermalWeaklyCompressibleLiquidModel(CommonWeaklyCompressibleLiquidModel):
    """LiquidModel.jl:35-66; adds t_0 (reference temperature), kappa_0, cp, beta."""

    def __init__(self, rho_0, c_0, p_0, gamma: int, mu_0, body_force_vec, t_0, kappa_0, cp, beta):
        super().__init__(rho_0, c_0, p_0, gamma, mu_0, body_force_vec)
        self.t_0_ = float(t_0)
        self.kappa_0_ = float(kappa_0)
        self.cp_ = float(cp)
        self.beta_ = float(beta)
