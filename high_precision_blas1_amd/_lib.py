"""ctypes binding of ``libhpb.so`` generated from the prototypes in ``include/hpb.h``.

The header is the single source of truth: every ``int hpb_*(...)`` prototype is
parsed and bound with matching argument types, so the Python view cannot drift
from the C ABI.  There is no fallback: if the library is missing the import of
:func:`lib` raises.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HEADER = os.path.join(ROOT, "include", "hpb.h")
LIBPATH = os.path.join(PKG, "libhpb.so")

_SCALARS = {
    "size_t": ctypes.c_size_t,
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "uint64_t": ctypes.c_uint64,
    "uint32_t": ctypes.c_uint32,
    "int64_t": ctypes.c_int64,
    "hpb_handle_t": ctypes.c_void_p,
    "hpb_batch_t": ctypes.c_void_p,
    "hpb_trsv_t": ctypes.c_void_p,
}

_PROTO = re.compile(r"^(int|const char\*|void)\s+(hpb_\w+)\s*\(([^;]*?)\)\s*;", re.M | re.S)


class HPBError(RuntimeError):
    """Raised when a C ABI call returns a non-zero status."""

    def __init__(self, name, status, msg):
        super().__init__("%s returned %d: %s" % (name, status, msg))
        self.status = status


def _ctype(decl):
    decl = decl.strip()
    decl = re.sub(r"\s+", " ", decl)
    if decl == "void":
        return None
    # drop the parameter name
    m = re.match(r"^(.*?)([A-Za-z_]\w*)$", decl)
    t = m.group(1).strip() if m else decl
    if "*" in t:
        return ctypes.c_void_p
    t = t.replace("const ", "").strip()
    if t not in _SCALARS:
        raise TypeError("unmapped C type %r" % t)
    return _SCALARS[t]


def parse_header(path=HEADER):
    """Return {name: (restype_str, [ctypes argtypes])} for every prototype."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for ret, name, args in _PROTO.findall(src):
        args = args.strip()
        types = [] if args in ("", "void") else [_ctype(a) for a in args.split(",")]
        out[name] = (ret, types)
    return out


_LIB = None
_PROTOS = None


def lib():
    """Load libhpb.so (built by ``high_precision_blas1_amd/build.py``)."""
    global _LIB, _PROTOS
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libhpb.so is not built (run __graft_entry__.build() or "
                           "python high_precision_blas1_amd/build.py); there is no fallback")
    L = ctypes.CDLL(LIBPATH)
    _PROTOS = parse_header()
    for name, (ret, types) in _PROTOS.items():
        try:
            f = getattr(L, name)
        except AttributeError:
            # declared but not exported: a library older than the header; calling it
            # raises, and test_capi's export check fails
            _MISSING.append(name)
            continue
        f.argtypes = types
        f.restype = ctypes.c_char_p if ret == "const char*" else (None if ret == "void" else ctypes.c_int)
    _LIB = L
    return L


_MISSING = []


def missing_symbols():
    """Names declared in include/hpb.h that the loaded libhpb.so does not export."""
    lib()
    return list(_MISSING)


def prototypes():
    lib()
    return dict(_PROTOS)


def call(name, *args):
    """Call ``name`` and raise :class:`HPBError` on a non-zero status."""
    L = lib()
    st = getattr(L, name)(*args)
    if st != 0:
        msg = L.hpb_last_error()
        raise HPBError(name, st, msg.decode() if msg else "")
    return st
