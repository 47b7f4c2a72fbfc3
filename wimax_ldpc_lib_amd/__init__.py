"""MI355X-native WiMAX (802.16e) QC-LDPC codec: gfx950 HIP kernels behind a C ABI.

The compute path is libwimax_ldpc_amd.so (wimax_ldpc_lib_amd/lib/), built by
wimax_ldpc_lib_amd/build.py; this package only binds it.
"""
from .capi import WLError, code_params, device_count, lib  # noqa: F401
from .codec import (FIVESIXTHS, HALFRATE, THREEQUARTERSA, THREEQUARTERSB, TWOTHIRDSA,  # noqa: F401
                    TWOTHIRDSB, ldpc, ldpc_decoder, ldpc_encoder, qc_code, unpack_hard)

__all__ = ["WLError", "code_params", "device_count", "lib", "ldpc", "ldpc_decoder", "ldpc_encoder",
           "qc_code", "unpack_hard", "HALFRATE", "TWOTHIRDSA", "TWOTHIRDSB", "THREEQUARTERSA", "THREEQUARTERSB",
           "FIVESIXTHS"]
