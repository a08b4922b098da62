"""Learned warm start for the PDHG solvers: a small 3-D CNN on MFMA (bf16, fp32 accumulation)."""

from .net import WarmStartNet, default_weights_path
from .warm import DEFAULT_CG_STEPS, cg_refine, cnn_warm_start, warm_start

__all__ = ["WarmStartNet", "default_weights_path", "cnn_warm_start", "cg_refine", "warm_start",
           "DEFAULT_CG_STEPS"]
