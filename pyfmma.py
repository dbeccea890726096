"""pyfmma -- the module name of fockl/FMMA's Python binding (setup.py:6).

``import pyfmma`` gives the six classes fmmad1, fmmaf1, fmmad2, fmmaf2,
fmmad3, fmmaf3 with the reference's methods (cpython/pyfmma.cpp:140-168);
every solve runs on the GPU through fmma_amd's C-ABI, and set_fn1 / set_fn2
take Python callables (traced and compiled for the GPU by fmma_amd.jit).
"""
from fmma_amd.pyfmma import fmmad1, fmmad2, fmmad3, fmmaf1, fmmaf2, fmmaf3  # noqa: F401

__all__ = ["fmmad1", "fmmaf1", "fmmad2", "fmmaf2", "fmmad3", "fmmaf3"]
