"""dtcf_amd -- MI355X (gfx950) engine for DTCF's per-rank local contraction path.

Layout:
  csrc/          HIP kernels, the C ABI (include/dtcf_local.h), the C++ host
                 planner for the RRR local step and the C++ drop-in symbols
                 (include/dtcf_dropin.h)
  _lib.py        ctypes binding of libdtcf.so (raises if it is not built)
  ops.py         torch-tensor wrappers of the C ABI
  rrr.py         RRR tile tensors and the local step (contraction.cc)
  dltc.py        DLTC index split, transpose and tile tasks (DLTC/)
  synth.py       seeded synthetic inputs
"""
from ._lib import DtcfError, lib  # noqa: F401

__all__ = ["DtcfError", "lib"]
