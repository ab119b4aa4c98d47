"""Arrow-layout columns in host memory (numpy) or device memory (torch), as the C ABI takes them.

A column is the reference's VectorBuffers (kernels/.../VectorBuffers.java:48-215): a lane type,
a row count, an optional LSB-first validity bitmap, the data buffer, int32 offsets for plain
UTF8, and a dictionary (itself a plain UTF8 column in host memory) for dictionary-encoded UTF8.
"""
import ctypes as C

import numpy as np

from . import abi

_NP_DTYPE = {abi.INT32: np.int32, abi.INT64: np.int64, abi.FLOAT64: np.float64, abi.UTF8: np.int32}


def pack_bits(bools):
    """LSB-first bitmap of a boolean array (Arrow / Bitmap.java layout), padded to 8 bytes."""
    b = np.asarray(bools, dtype=bool)
    packed = np.packbits(b, bitorder="little")
    pad = (-len(packed)) % 8
    if pad:
        packed = np.concatenate([packed, np.zeros(pad, np.uint8)])
    return packed


def unpack_bits(bitmap, n):
    return np.unpackbits(np.asarray(bitmap, dtype=np.uint8), bitorder="little", count=n).astype(bool)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data if a.size else None
    return a.data_ptr()  # torch tensor


class Column:
    """An input column; keeps every buffer it points at alive."""

    def __init__(self, type, data, validity=None, offsets=None, dictionary=None, length=None,
                 mem=abi.MEM_HOST):
        self.type = type
        self.data = data
        self.validity = validity
        self.offsets = offsets
        self.dictionary = dictionary
        self.mem = mem
        if length is None:
            if type == abi.DECIMAL128:
                length = len(data) // 2
            elif type == abi.UTF8 and dictionary is None:
                length = len(offsets) - 1
            elif type == abi.BOOL:
                raise ValueError("a BOOL column needs its length")
            else:
                length = len(data)
        self.length = int(length)
        self._dict_c = dictionary.c() if dictionary is not None else None
        self._c = abi.VrColumn(self.type, self.mem, self.length, _ptr(validity), _ptr(data),
                               _ptr(offsets),
                               C.cast(C.pointer(self._dict_c), C.c_void_p) if self._dict_c is not None else None)

    def c(self):
        return self._c

    # ---- constructors for host columns
    @staticmethod
    def ints(values, nulls=None):
        v = np.ascontiguousarray(values, dtype=np.int32)
        return Column(abi.INT32, v, None if nulls is None else pack_bits(~np.asarray(nulls, bool)))

    @staticmethod
    def longs(values, nulls=None):
        v = np.ascontiguousarray(values, dtype=np.int64)
        return Column(abi.INT64, v, None if nulls is None else pack_bits(~np.asarray(nulls, bool)))

    @staticmethod
    def doubles(values, nulls=None):
        v = np.ascontiguousarray(values, dtype=np.float64)
        return Column(abi.FLOAT64, v, None if nulls is None else pack_bits(~np.asarray(nulls, bool)))

    @staticmethod
    def bools(values, nulls=None):
        v = np.asarray(values, bool)
        return Column(abi.BOOL, pack_bits(v), None if nulls is None else pack_bits(~np.asarray(nulls, bool)),
                      length=len(v))

    @staticmethod
    def decimals(values, nulls=None):
        """DECIMAL128 from Python ints: low limb then high limb, two's complement."""
        n = len(values)
        arr = np.zeros(2 * n, dtype=np.int64)
        for i, x in enumerate(values):
            x = int(x) & ((1 << 128) - 1)
            lo, hi = x & ((1 << 64) - 1), x >> 64
            arr[2 * i] = np.int64(lo - (1 << 64) if lo >= (1 << 63) else lo)
            arr[2 * i + 1] = np.int64(hi - (1 << 64) if hi >= (1 << 63) else hi)
        return Column(abi.DECIMAL128, arr, None if nulls is None else pack_bits(~np.asarray(nulls, bool)))

    @staticmethod
    def strings(values):
        """Plain UTF8; None entries are null."""
        offs = [0]
        data = bytearray()
        nulls = []
        for s in values:
            nulls.append(s is None)
            if s is not None:
                data += s.encode("utf-8")
            offs.append(len(data))
        return Column(abi.UTF8, np.frombuffer(bytes(data) or b"\0", dtype=np.uint8).copy(),
                      pack_bits(~np.asarray(nulls, bool)) if any(nulls) else None,
                      offsets=np.asarray(offs, np.int32), length=len(values))

    @staticmethod
    def dictionary_strings(codes, dictionary, nulls=None):
        """Dictionary-encoded UTF8: int32 codes into `dictionary` (a plain UTF8 Column)."""
        v = np.ascontiguousarray(codes, dtype=np.int32)
        return Column(abi.UTF8, v, None if nulls is None else pack_bits(~np.asarray(nulls, bool)),
                      dictionary=dictionary, length=len(v))


class OutColumn:
    """A caller-allocated output column in host memory."""

    def __init__(self, type, capacity, mem=abi.MEM_HOST, validity=None, data=None):
        self.type = type
        self.capacity = int(capacity)
        self.mem = mem
        nb = max(8, (self.capacity + 7) // 8)
        self.validity = validity if validity is not None else np.zeros(nb, np.uint8)
        if data is not None:
            self.data = data
        elif type == abi.BOOL:
            self.data = np.zeros(nb, np.uint8)
        elif type == abi.DECIMAL128:
            self.data = np.zeros(2 * max(1, self.capacity), np.int64)
        else:
            self.data = np.zeros(max(1, self.capacity), _NP_DTYPE[type])
        self._c = abi.VrOutColumn(type, mem, self.capacity, _ptr(self.validity), _ptr(self.data), None)

    def c(self):
        return self._c

    def valid(self, n):
        return unpack_bits(self.validity, n)

    def values(self, n):
        """Row values as Python objects (None for null)."""
        ok = self.valid(n)
        if self.type == abi.DECIMAL128:
            d = self.data[: 2 * n].astype(np.int64)
            out = []
            for i in range(n):
                if not ok[i]:
                    out.append(None)
                    continue
                lo = int(d[2 * i]) & ((1 << 64) - 1)
                hi = int(d[2 * i + 1])
                out.append((hi << 64) | lo)
            return out
        if self.type == abi.BOOL:
            bits = unpack_bits(self.data, n)
            return [bool(bits[i]) if ok[i] else None for i in range(n)]
        d = self.data[:n]
        return [d[i].item() if ok[i] else None for i in range(n)]


def c_array(cols, struct=abi.VrColumn):
    """A contiguous ctypes array of the columns' structs."""
    arr = (struct * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c.c()
    return arr
