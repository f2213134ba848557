"""MI355X photon random-walk engine (xmcml's ComputePhoton/LaunchOMP path).

The engine itself is the C-ABI library ``libxmcml_amd.so`` (include/xmcml.h)
built from ``xmcml_amd/csrc``; this package only locates and binds it and
builds xmcml-shaped inputs.  Nothing here computes photon transport on the CPU.
"""
from .abi import *  # noqa: F401,F403
