"""hmpc_amd: MI355X (gfx950) implementation of hmpc's NTT / BGV / ChaCha20 core.

The product is the C-ABI HIP library ``libhmpc_hip.so`` (include/hmpc_hip.h); this
package holds its sources (``csrc/``), its build step (``build.py``) and thin ctypes
bindings (``runtime.py``).
"""

from .runtime import (Batch, Buffer, Context, Event, HmpcError, HostBuffer, Modulus, Server, Stream, View, LIB_PATH,  # noqa: F401
                      BITS_TERNARY, SIGNATURES,
                      int_to_words, load, words_to_int)

Q64 = (1 << 64) - (1 << 32) + 1
Q128A = (1 << 124) + (1 << 20) + (1 << 18) + 1
Q128B = 0x10000000000000000000000000078001
