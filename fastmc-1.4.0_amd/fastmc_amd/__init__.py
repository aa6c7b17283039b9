"""Host-side mirror of the FastMC trial-move path over the MI355X C ABI.

``lib()`` loads ``lib/libfastmc_amd.so`` (built in-tree by
``__graft_entry__.build``) and fails loudly when it is missing: there is no
CPU fallback on the product path.
"""
from .capi import (lib, FmcError, LIB_PATH, Context, context_from_inputs,  # noqa: F401
                   INSERT, DELETE, DISPLACE, PREC_FP64, PREC_MIXED)
