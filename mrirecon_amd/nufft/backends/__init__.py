from .hip import HipNUFFTAdapter, NufftPlan

__all__ = ["HipNUFFTAdapter", "NufftPlan"]
