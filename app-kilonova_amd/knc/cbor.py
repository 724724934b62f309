"""CBOR (RFC 8949) for the small items of the reference's files.

The reference serialises its ``App`` and ``Products`` with serde through
ciborium 0.1.0 (``src/io.rs:24-35``), which writes every number in its
shortest lossless form (RFC 8949 section 4.2.2, "preferred serialization"):
integers in the smallest head, floats in the narrowest of binary16 / binary32
/ binary64 that holds the value exactly.  :func:`encode` does the same for
Python values -- ``None`` (serde's ``None``/unit), ``bool``, ``int``,
``float``, ``str``, ``bytes``, ``list``/``tuple`` (arrays) and ``dict``
(maps, in insertion order) -- and :class:`Raw` splices bytes that are
already CBOR.  :func:`decode` reads any well-formed CBOR: arrays become
lists, maps dicts (array keys become tuples), tags are dropped.

The block data of a checkpoint or products file is encoded and decoded by
the library (``knc_write_checkpoint``, ``knc_checkpoint_open``).  This
module serves the Configuration and Tasks items, the tests and
:mod:`knc.records`, for which ``decode(..., arrays=True)`` returns long
arrays of numbers as numpy arrays (``knc_cbor_read_f64``).
"""
import math
import struct

__all__ = ["Raw", "CBORError", "encode", "decode", "encode_float"]


class CBORError(ValueError):
    pass


class Raw(bytes):
    """Bytes that already are one CBOR data item."""


def _head(major, n):
    mt = major << 5
    if n < 24:
        return bytes((mt | n,))
    if n < 1 << 8:
        return bytes((mt | 24, n))
    if n < 1 << 16:
        return bytes((mt | 25,)) + n.to_bytes(2, "big")
    if n < 1 << 32:
        return bytes((mt | 26,)) + n.to_bytes(4, "big")
    if n < 1 << 64:
        return bytes((mt | 27,)) + n.to_bytes(8, "big")
    raise CBORError("integer out of range")


def encode_float(x):
    """The shortest of binary16, binary32, binary64 that holds x exactly."""
    x = float(x)
    bits = struct.pack(">d", x)
    try:
        h = struct.pack(">e", x)
        if struct.pack(">d", struct.unpack(">e", h)[0]) == bits:
            return b"\xf9" + h
    except (OverflowError, struct.error):
        pass
    try:
        f = struct.pack(">f", x)
        if struct.pack(">d", struct.unpack(">f", f)[0]) == bits:
            return b"\xfa" + f
    except (OverflowError, struct.error):
        pass
    return b"\xfb" + bits


def _encode(obj, out):
    if isinstance(obj, Raw):
        out.append(bytes(obj))
    elif obj is None:
        out.append(b"\xf6")
    elif obj is True:
        out.append(b"\xf5")
    elif obj is False:
        out.append(b"\xf4")
    elif isinstance(obj, int):
        out.append(_head(0, obj) if obj >= 0 else _head(1, -1 - obj))
    elif isinstance(obj, float):
        out.append(encode_float(obj))
    elif isinstance(obj, str):
        b = obj.encode("utf-8")
        out.append(_head(3, len(b)))
        out.append(b)
    elif isinstance(obj, (bytes, bytearray)):
        out.append(_head(2, len(obj)))
        out.append(bytes(obj))
    elif isinstance(obj, (list, tuple)):
        out.append(_head(4, len(obj)))
        for x in obj:
            _encode(x, out)
    elif isinstance(obj, dict):
        out.append(_head(5, len(obj)))
        for k, v in obj.items():
            _encode(k, out)
            _encode(v, out)
    elif hasattr(obj, "item") and not hasattr(obj, "__len__"):
        _encode(obj.item(), out)      # numpy scalars
    else:
        raise CBORError(f"cannot encode {type(obj).__name__}")


def encode(obj):
    out = []
    _encode(obj, out)
    return b"".join(out)


def _half(h):
    s, e, m = h >> 15, (h >> 10) & 0x1F, h & 0x3FF
    if e == 0:
        v = math.ldexp(m, -24)
    elif e == 31:
        v = math.inf if m == 0 else math.nan
    else:
        v = math.ldexp(m | 0x400, e - 25)
    return -v if s else v


def _native_f64(data, offset, count, arity):
    """The array at data[offset] of `count` numbers (or arity-tuples of
    numbers) through the library's decoder, as float64 of shape (count,) or
    (count, arity) and the offset past it; None where the library is absent
    or the array is not of that form."""
    try:
        import ctypes

        import numpy as np

        from . import capi
        L = capi.lib()
    except (ImportError, RuntimeError, OSError):
        return None
    fn = L.knc_cbor_read_f64
    fn.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32,
                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    out = np.empty(count * arity)
    end = ctypes.c_uint64()
    if fn(data, len(data), offset, count, arity, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(end)):
        return None
    return (out if arity == 1 else out.reshape(count, arity)), end.value


_NUMBER_HEADS = set(range(0x00, 0x1c)) | set(range(0x20, 0x3c)) | {0xf9, 0xfa, 0xfb}


class _Reader:
    def __init__(self, data, pos=0, arrays=False):
        self.d = memoryview(data)
        self.p = pos
        self.raw = bytes(data) if arrays else None

    def _arity(self):
        """1 if the next item is a number, k if it is an array of k <= 8
        numbers, else 0."""
        if self.p >= len(self.d):
            return 0
        b = self.d[self.p]
        if b in _NUMBER_HEADS:
            return 1
        if 0x81 <= b <= 0x88 and self.p + 1 < len(self.d) and self.d[self.p + 1] in _NUMBER_HEADS:
            return b - 0x80
        return 0

    def byte(self):
        if self.p >= len(self.d):
            raise CBORError("unexpected end of data")
        b = self.d[self.p]
        self.p += 1
        return b

    def take(self, n):
        if self.p + n > len(self.d):
            raise CBORError("truncated item")
        b = bytes(self.d[self.p:self.p + n])
        self.p += n
        return b

    def head(self):
        b = self.byte()
        major, minor = b >> 5, b & 31
        if minor < 24:
            return major, minor, minor
        if minor <= 27:
            return major, minor, int.from_bytes(self.take(1 << (minor - 24)), "big")
        if minor == 31:
            return major, minor, None
        raise CBORError("reserved additional information")

    def item(self, depth=0):
        if depth > 256:
            raise CBORError("nesting too deep")
        start = self.p
        major, minor, arg = self.head()
        if major == 4 and self.raw is not None and arg is not None and arg >= 64:
            arity = self._arity()
            got = _native_f64(self.raw, start, arg, arity) if arity else None
            if got is not None:
                self.p = got[1]
                return got[0]
        if arg is None and major in (0, 1, 6):
            raise CBORError("indefinite length on a non-container")
        if major == 0:
            return arg
        if major == 1:
            return -1 - arg
        if major in (2, 3):
            if arg is None:
                chunks = []
                while not self._brk():
                    m2, _, n2 = self.head()
                    if m2 != major or n2 is None:
                        raise CBORError("bad string chunk")
                    chunks.append(self.take(n2))
                raw = b"".join(chunks)
            else:
                raw = self.take(arg)
            return raw.decode("utf-8") if major == 3 else raw
        if major == 4:
            out = []
            if arg is None:
                while not self._brk():
                    out.append(self.item(depth + 1))
            else:
                for _ in range(arg):
                    out.append(self.item(depth + 1))
            return out
        if major == 5:
            out = {}
            n = 0
            while (arg is None and not self._brk()) or (arg is not None and n < arg):
                k = self.item(depth + 1)
                if isinstance(k, list):
                    k = _freeze(k)
                out[k] = self.item(depth + 1)
                n += 1
            return out
        if major == 6:
            return self.item(depth + 1)
        # major 7
        if minor == 20:
            return False
        if minor == 21:
            return True
        if minor in (22, 23):
            return None
        if minor == 25:
            return _half(arg)
        if minor == 26:
            return struct.unpack(">f", arg.to_bytes(4, "big"))[0]
        if minor == 27:
            return struct.unpack(">d", arg.to_bytes(8, "big"))[0]
        if arg is None:
            raise CBORError("unexpected break")
        return arg                      # other simple values

    def _brk(self):
        if self.p < len(self.d) and self.d[self.p] == 0xFF:
            self.p += 1
            return True
        if self.p >= len(self.d):
            raise CBORError("unexpected end of data")
        return False


def _freeze(x):
    return tuple(_freeze(v) for v in x) if isinstance(x, list) else x


def decode(data, allow_trailing=False, arrays=False):
    """Decodes one CBOR item from data (bytes-like).  With arrays=True,
    arrays of 64 or more numbers (or of equal small tuples of numbers) come
    back as float64 numpy arrays, decoded by the library when it is built."""
    r = _Reader(data, arrays=arrays)
    obj = r.item()
    if not allow_trailing and r.p != len(r.d):
        raise CBORError("trailing bytes")
    return obj
