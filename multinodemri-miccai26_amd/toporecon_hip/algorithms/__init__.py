"""HIP PDHG reconstructors (TVM, TVME) with the reference's constructor signatures."""

from .tvm import TvmReconstructor
from .tvme import TvmeReconstructor

__all__ = ["TvmReconstructor", "TvmeReconstructor"]
