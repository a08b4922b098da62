"""toporecon_hip: MI355X-native PDHG reconstruction hot path (TVM / TVME / TVMW).

The NUFFT plugin (``nufft_backend``), the HIP reconstructors (``algorithms``)
and the host-side conditioning / synthetic-data helpers.  Compute runs in the
hand-written HIP library ``libtoporecon_hip.so`` through a C ABI; PyTorch only
provides device buffers and streams.
"""

__version__ = "0.1.0"
