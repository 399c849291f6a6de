from .forwardEuler import ForwardEuler
from .fused import HeatStep, WaveStep

__all__ = ["ForwardEuler", "HeatStep", "WaveStep"]
