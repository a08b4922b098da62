"""HIP PDHG reconstructors (TVM, TVME, TVMW) with the reference's constructor signatures."""

from .tvm import TvmReconstructor
from .tvme import TvmeReconstructor
from .tvmw import TvmwReconstructor

__all__ = ["TvmReconstructor", "TvmeReconstructor", "TvmwReconstructor"]
