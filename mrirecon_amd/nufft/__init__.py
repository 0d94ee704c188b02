from .layout import AxisSpec
from .api import NUFFT, NUFFTConfig, NSNUFFT

__all__ = ["NUFFT", "NSNUFFT", "NUFFTConfig", "AxisSpec"]
