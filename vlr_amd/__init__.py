"""VLR light-transport path on MI355X (gfx950).

The product is ``libvlr_amd.so``: the VLR C API (include/vlr/vlr.h) over a wavefront path
tracer of hand-written HIP kernels.  Python modules here are thin helpers:
``vlr_amd.api`` (ctypes binding of the C API), ``vlr_amd.scenes`` (procedural benchmark
scenes) and ``vlr_amd.build`` (in-tree hipcc build).
"""
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvlr_amd.so"

__all__ = ["LIB_PATH"]
