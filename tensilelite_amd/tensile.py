"""The Tensile-named kernels as a client loads them.

``lib/ffgemm_tensile_gfx950.co`` holds kernels under the names and 148-byte
argument list of the reference's Tensile-layout kernels
(``Cijk_Alik_Bljk_{SB,HBH,BBH}_MT{MT0}x{MT1}x{DU}_SE_K1``, src/gemmMfma.cpp:
70-103).  A Tensile client loads such a code object with hipModuleLoad and
launches the kernel with a packed argument buffer; :class:`TensileModule`
does exactly that through the HIP runtime's module API (ctypes), and
:func:`pack_kernarg` builds the buffer.
"""
from __future__ import annotations

import ctypes
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
CO_PATH = os.path.join(HERE, "lib", "ffgemm_tensile_gfx950.co")

KERNARG_BYTES = 148
# kernel -> (MT0, MT1, workgroup size)
KERNELS = {
    "Cijk_Alik_Bljk_BBH_MT256x256x64_SE_K1": (256, 256, 512),
    "Cijk_Alik_Bljk_HBH_MT256x256x64_SE_K1": (256, 256, 512),
    "Cijk_Alik_Bljk_BBH_MT128x128x32_SE_K1": (128, 128, 256),
    "Cijk_Alik_Bljk_HBH_MT128x128x32_SE_K1": (128, 128, 256),
    "Cijk_Alik_Bljk_BBH_MT64x64x32_SE_K1": (64, 64, 256),
    "Cijk_Alik_Bljk_HBH_MT64x64x32_SE_K1": (64, 64, 256),
    "Cijk_Alik_Bljk_SB_MT128x128x32_SE_K1": (128, 128, 256),
    "Cijk_Alik_Bljk_SB_MT64x64x32_SE_K1": (64, 64, 256),
}


def _f16_bits(x: float) -> int:
    return struct.unpack("<H", struct.pack("<e", x))[0]


def pack_kernarg(kind: str, D: int, C: int, A: int, B: int, alpha: float, beta: float, M: int, N: int, K: int,
                 lda: int, ldb: int, ldc: int, ldd: int, batch: int = 1, bsA: int = 0, bsB: int = 0, bsC: int = 0,
                 bsD: int = 0, sizes=(0, 0, 0)) -> bytes:
    """The 148-byte argument buffer of a Tensile-layout TN kernel.  ``kind``
    is SB, HBH or BBH; HBH takes alpha and beta as fp16 in the low half of
    their 32-bit slots, the others as fp32.  sizes = (sizeC, sizeA, sizeB) in
    elements; the grid-mapping values are written as Tensile would
    (NumWorkGroups0/1) or zero (these kernels ignore them)."""
    if kind == "HBH":
        sa, sb = struct.pack("<HH", _f16_bits(alpha), 0), struct.pack("<HH", _f16_bits(beta), 0)
    else:
        sa, sb = struct.pack("<f", alpha), struct.pack("<f", beta)
    mt = 256 if kind != "SB" else 128
    buf = struct.pack("<QQQ", *sizes) + struct.pack("<QQQQ", D, C, A, B) + sa + sb
    buf += struct.pack("<8I", ldd, bsD, ldc, bsC, lda, bsA, ldb, bsB)
    buf += struct.pack("<4I", M, N, batch, K)
    buf += struct.pack("<9I", 0, (M + mt - 1) // mt, (N + mt - 1) // mt, 0, 0, 0, 0, 0, 0)
    assert len(buf) == KERNARG_BYTES
    return buf


class TensileModule:
    """hipModuleLoad of the Tensile-named code object; launch() runs one kernel
    with a packed argument buffer on a stream, as a Tensile client does."""

    HIP_LAUNCH_PARAM_BUFFER_POINTER = 1
    HIP_LAUNCH_PARAM_BUFFER_SIZE = 2
    HIP_LAUNCH_PARAM_END = 3

    def __init__(self, path: str = CO_PATH):
        if not os.path.isfile(path):
            raise RuntimeError(f"{path} not built; run python -m tensilelite_amd.build")
        # the HIP runtime already loaded into this process (same soname as the
        # one libffgemm.so links)
        self.hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.hip.hipModuleLoad.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p]
        self.hip.hipModuleGetFunction.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_char_p]
        self.hip.hipModuleLaunchKernel.argtypes = [ctypes.c_void_p] + [ctypes.c_uint] * 7 + [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipModuleUnload.argtypes = [ctypes.c_void_p]
        self.hip.hipGetErrorString.restype = ctypes.c_char_p
        self.mod = ctypes.c_void_p()
        self._check(self.hip.hipModuleLoad(ctypes.byref(self.mod), path.encode()), "hipModuleLoad")
        self._fn = {}

    def _check(self, e, what):
        if e != 0:
            raise RuntimeError(f"{what}: {self.hip.hipGetErrorString(e).decode()}")

    def function(self, name: str):
        if name not in self._fn:
            f = ctypes.c_void_p()
            self._check(self.hip.hipModuleGetFunction(ctypes.byref(f), self.mod, name.encode()),
                        f"hipModuleGetFunction({name})")
            self._fn[name] = f
        return self._fn[name]

    def launch(self, name: str, kernarg: bytes, grid, block: int, stream=None):
        buf = ctypes.create_string_buffer(kernarg, len(kernarg))
        size = ctypes.c_size_t(len(kernarg))
        extra = (ctypes.c_void_p * 5)(self.HIP_LAUNCH_PARAM_BUFFER_POINTER, ctypes.addressof(buf),
                                      self.HIP_LAUNCH_PARAM_BUFFER_SIZE, ctypes.addressof(size),
                                      self.HIP_LAUNCH_PARAM_END)
        self._check(self.hip.hipModuleLaunchKernel(self.function(name), grid[0], grid[1], grid[2], block, 1, 1, 0,
                                                   stream, None, extra), f"hipModuleLaunchKernel({name})")

    def close(self):
        if self.mod:
            self.hip.hipModuleUnload(self.mod)
            self.mod = ctypes.c_void_p()
