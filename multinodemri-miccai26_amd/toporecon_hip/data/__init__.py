"""Shard file formats of the reconstruction path (BART-style CFL/HDR)."""
