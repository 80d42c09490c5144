"""ctypes binding of include/oxide_rocm.h.

The library is loaded from this package directory (built in-tree by
``oxidelake_amd.build``). There is no fallback: when the library is missing
every entry point raises, so nothing can silently run on the CPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from .plan import Expr

LIB_PATH = Path(__file__).resolve().parent / "liboxide_rocm.so"

OK, EUNSUPPORTED, EDEVICE, EEXECUTION, EPLAN, EFORMAT, EALLOCATION = range(7)


class Column(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("dim", C.c_int32),
        ("len", C.c_uint64),
        ("values", C.c_void_p),
        ("validity", C.c_void_p),
    ]


class ArrowSchema(C.Structure):
    pass


ArrowSchema._fields_ = [
    ("format", C.c_char_p),
    ("name", C.c_char_p),
    ("metadata", C.c_char_p),
    ("flags", C.c_int64),
    ("n_children", C.c_int64),
    ("children", C.POINTER(C.POINTER(ArrowSchema))),
    ("dictionary", C.POINTER(ArrowSchema)),
    ("release", C.c_void_p),
    ("private_data", C.c_void_p),
]


class ArrowArray(C.Structure):
    pass


ArrowArray._fields_ = [
    ("length", C.c_int64),
    ("null_count", C.c_int64),
    ("offset", C.c_int64),
    ("n_buffers", C.c_int64),
    ("n_children", C.c_int64),
    ("buffers", C.POINTER(C.c_void_p)),
    ("children", C.POINTER(C.POINTER(ArrowArray))),
    ("dictionary", C.POINTER(ArrowArray)),
    ("release", C.c_void_p),
    ("private_data", C.c_void_p),
]


class GenSpec(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("dist", C.c_int32),
        ("dim", C.c_int32),
        ("pad", C.c_int32),
        ("a", C.c_double),
        ("b", C.c_double),
        ("null_frac", C.c_double),
        ("seed", C.c_uint64),
    ]


class Agg(C.Structure):
    _fields_ = [
        ("func", C.c_int32),
        ("pad", C.c_int32),
        ("input", Expr),
        ("name", C.c_char_p),
    ]


class AggSpec(C.Structure):
    _fields_ = [
        ("group_by", C.c_int32),
        ("naggs", C.c_int32),
        ("aggs", C.POINTER(Agg)),
        ("filter", C.POINTER(Expr)),
        ("groups_hint", C.c_uint64),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I32 = C.c_int32
U64 = C.c_uint64

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "oxr_last_error": (C.c_char_p, []),
    "oxr_last_feature": (C.c_char_p, []),
    "oxr_probe": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
    "oxr_open": (C.c_int, [C.c_int, PP]),
    "oxr_mem_info": (C.c_int, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(C.c_int)]),
    "oxr_arena_info": (C.c_int, [P, C.POINTER(U64), C.POINTER(U64)]),
    "oxr_alloc_device": (C.c_int, [P, C.c_size_t, PP]),
    "oxr_alloc_pinned": (C.c_int, [P, C.c_size_t, PP]),
    "oxr_buffer_ptr": (P, [P]),
    "oxr_buffer_len": (C.c_size_t, [P]),
    "oxr_buffer_free": (None, [P]),
    "oxr_stream_create": (C.c_int, [P, PP]),
    "oxr_stream_wrap": (C.c_int, [P, P, PP]),
    "oxr_stream_destroy": (None, [P]),
    "oxr_stream_sync": (C.c_int, [P]),
    "oxr_copy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "oxr_copy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "oxr_upload": (C.c_int, [P, C.POINTER(ArrowSchema), C.POINTER(ArrowArray), PP]),
    "oxr_download": (C.c_int, [P, P, C.POINTER(ArrowSchema), C.POINTER(ArrowArray)]),
    "oxr_batch_wrap": (C.c_int, [P, C.POINTER(Column), C.POINTER(C.c_char_p), I32, U64, PP]),
    "oxr_batch_rows": (U64, [P]),
    "oxr_batch_ncols": (I32, [P]),
    "oxr_batch_column": (C.c_int, [P, I32, C.POINTER(Column)]),
    "oxr_batch_name": (C.c_char_p, [P, I32]),
    "oxr_batch_free": (None, [P]),
    "oxr_batch_free_many": (None, [PP, U64]),
    "oxr_supports_predicate": (C.c_int, [C.POINTER(Expr)]),
    "oxr_supports_hash_join": (C.c_int, []),
    "oxr_supports_aggregate": (C.c_int, []),
    "oxr_filter_project": (
        C.c_int,
        [P, P, C.POINTER(Expr), C.POINTER(Expr), I32, C.POINTER(C.c_char_p), PP],
    ),
    "oxr_filter_kernel_source": (
        C.c_int,
        [C.POINTER(Column), I32, C.POINTER(Expr), C.POINTER(Expr), I32, I32, C.c_char_p,
         C.c_size_t, C.POINTER(C.c_size_t)],
    ),
    "oxr_filter_project_batches": (
        C.c_int,
        [P, P, U64, C.POINTER(Expr), C.POINTER(Expr), I32, C.POINTER(C.c_char_p), PP, U64,
         C.POINTER(U64)],
    ),
    "oxr_filter_project_noalloc": (
        C.c_int,
        [P, P, C.POINTER(Expr), C.POINTER(Expr), I32, C.POINTER(Column), C.POINTER(U64)],
    ),
    "oxr_join_build": (C.c_int, [P, P, I32, PP]),
    "oxr_join_table_info": (C.c_int, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(C.c_int)]),
    "oxr_join_table_free": (None, [P]),
    "oxr_join_probe": (C.c_int, [P, P, P, I32, PP, C.POINTER(U64)]),
    "oxr_join_matches_free": (None, [P]),
    "oxr_join_gather": (C.c_int, [P, P, P, P, P, U64, U64, C.POINTER(I32), I32, PP]),
    "oxr_aggregate": (C.c_int, [P, P, C.POINTER(AggSpec), PP]),
    "oxr_agg_create": (C.c_int, [P, P, C.POINTER(AggSpec), PP]),
    "oxr_agg_update": (C.c_int, [P, P, P]),
    "oxr_agg_update_join": (C.c_int, [P, P, P, P, I32, P, I32]),
    "oxr_agg_finish": (C.c_int, [P, P, PP]),
    "oxr_agg_free": (None, [P]),
    "oxr_aggregate_kernel_source": (
        C.c_int,
        [C.POINTER(Column), I32, C.POINTER(AggSpec), I32, I32, C.c_char_p, C.c_size_t,
         C.POINTER(C.c_size_t)],
    ),
    "oxr_agg_batch_prepare": (C.c_int, [P, P, C.POINTER(AggSpec), C.POINTER(I32), PP]),
    "oxr_agg_batch_columns": (C.c_int, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(C.c_char_p)]),
    "oxr_agg_batch_run": (
        C.c_int,
        [P, P, C.POINTER(AggSpec), I32, C.POINTER(C.c_int64), C.POINTER(C.c_uint8), C.POINTER(I32)],
    ),
    "oxr_agg_batch_free": (None, [P]),
    "oxr_agg_batch_kernel_source": (
        C.c_int,
        [C.POINTER(Column), I32, C.POINTER(AggSpec), I32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)],
    ),
    "oxr_vector_distance": (
        C.c_int,
        [P, P, I32, C.POINTER(C.c_float), I32, I32, C.c_char_p, PP],
    ),
    "oxr_vector_topk": (
        C.c_int,
        [P, P, I32, C.POINTER(C.c_float), I32, I32, I32, I32, C.POINTER(C.c_int64),
         C.POINTER(C.c_float)],
    ),
    "oxr_vector_index_create": (C.c_int, [P, P, I32, I32, I32, PP]),
    "oxr_vector_index_search": (
        C.c_int,
        [P, P, C.POINTER(C.c_float), I32, I32, I32, C.POINTER(C.c_int64), C.POINTER(C.c_float)],
    ),
    "oxr_vector_index_free": (None, [P]),
    "oxr_generate": (C.c_int, [P, C.POINTER(GenSpec), C.POINTER(C.c_char_p), I32, U64, PP]),
    "oxr_tune_set": (C.c_int, [C.c_char_p, C.c_int64]),
    "oxr_tune_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "oxr_profile_enable": (None, [C.c_int]),
    "oxr_profile_reset": (None, []),
    "oxr_profile_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(U64)]),
}


class _Lib:
    def __init__(self) -> None:
        self._lib = None

    def load(self) -> C.CDLL:
        if self._lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m oxidelake_amd.build` "
                    "(the ROCm backend has no CPU fallback)"
                )
            lib = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            self._lib = lib
        return self._lib

    def __getattr__(self, name: str):
        return getattr(self.load(), name)


lib = _Lib()


def exported_symbols() -> list[str]:
    """Names the C ABI declares (every `oxr_*` function in the header)."""
    import re

    header = Path(__file__).resolve().parents[1] / "include" / "oxide_rocm.h"
    text = header.read_text()
    return sorted(set(re.findall(r"\b(oxr_[a-z0-9_]+)\s*\(", text)))
