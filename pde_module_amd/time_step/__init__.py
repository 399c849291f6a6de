from .forwardEuler import ForwardEuler
from .fused import HeatStep, WaveStep
from .rungeKatta import RungeKatta1, RungeKatta2, field_dtype

__all__ = ["ForwardEuler", "HeatStep", "WaveStep", "RungeKatta1", "RungeKatta2", "field_dtype"]
