"""Time discretisation: mirror of TimeDiscretization/*.jl."""
from __future__ import annotations


class DensityReinitializedFixedTimeIntervalForwardEuler:
    """FixedTimeIntervalDiscretization.jl:10-16.

    Positional order is the reference's: (dt, total_step, output_interval,
    check_interval, density_reinitialized_interval).  Several reference
    examples pass the last two swapped; keyword arguments avoid that.
    """

    def __init__(self, dt, total_step, output_interval, check_interval, density_reinitialized_interval):
        self.dt_ = float(dt)
        self.total_step_ = int(total_step)
        self.output_interval_ = int(output_interval)
        self.check_interval_ = int(check_interval)
        self.density_reinitialized_interval_ = int(density_reinitialized_interval)


def isDensityReinitializedStep(step: int, td) -> bool:
    """FixedTimeIntervalDiscretization.jl:18-23."""
    return step % td.density_reinitialized_interval_ == 0


def isOutputStep(step: int, td) -> bool:
    """TimeDiscretization.jl:13-18."""
    return step % td.output_interval_ == 0


def isCheckStep(step: int, td) -> bool:
    """TimeDiscretization.jl:20-25."""
    return step % td.check_interval_ == 0
