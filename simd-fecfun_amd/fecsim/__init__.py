"""fecsim: MI355X-native LDPC simulator hot path (host side).

Mirrors the reference simulator's decoder interface (tomt-1/simd-fecfun:
decode_matrix_allocate / LDPCdecode_set and the .dat table files) over the C ABI of
libfecsim.so, whose kernels are hand-written HIP for gfx950.
"""
from . import codes, minprod_tree  # noqa: F401  (pure host code, importable without the GPU library)
from .lib import F32, F64, I8, I16, I32, MIN, PROD, FecsimError  # noqa: F401


def __getattr__(name):
    # Decoder/Hdec load libfecsim.so on first use, so `import fecsim` works on a
    # machine where only the table tools are needed.
    if name in ("Decoder", "Hdec"):
        from . import decoder
        return getattr(decoder, name)
    raise AttributeError(name)
