"""The C-ABI library loads and exports every symbol include/oxide_rocm.h declares
(no device calls: this runs without a GPU)."""
import ctypes as C

from oxidelake_amd._capi import LIB_PATH, SIGNATURES, exported_symbols


def test_library_is_built():
    assert LIB_PATH.exists(), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(LIB_PATH))
    missing = [s for s in exported_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_bindings_cover_the_header():
    assert set(exported_symbols()) <= set(SIGNATURES), set(exported_symbols()) - set(SIGNATURES)


def test_probe_without_a_device_is_a_typed_decline():
    from oxidelake_amd._capi import lib

    why = C.create_string_buffer(256)
    rc = lib.oxr_probe(0, why, 256)
    if rc != 0:  # no GPU in this container: a reason, never a crash
        assert rc == 1
        assert why.value


def test_supports_predicate_is_host_only():
    from oxidelake_amd.backend import RocmBackend  # noqa: F401  (import only)
    from oxidelake_amd._capi import lib
    from oxidelake_amd import plan

    C_ = plan.Comparison
    p = plan.Encoded(plan.and_(plan.cmp(plan.col(0), C_.GtEq, plan.Lit.int64(2)),
                               plan.cmp(plan.col(1), C_.Lt, plan.Lit.float64(4.0))))
    assert lib.oxr_supports_predicate(p.ref()) == 1
    bad = plan.Encoded(plan.and_(plan.col(0), plan.col(1)))  # AND of non-booleans
    assert lib.oxr_supports_predicate(bad.ref()) == 0


def test_tune_set_refuses_unknown_knobs_and_values():
    # a shape the kernels do not take is refused, never replaced by another:
    # a benchmark line reports only shapes that ran
    import pytest

    from oxidelake_amd import autotune
    from oxidelake_amd.backend import PlanError

    for knob, bad in [("filter.rows", 0), ("filter.rows", 33), ("agg.rows", 64), ("filter.v", 3),
                      ("agg.spill_pct", 0), ("vector.wide", 4), ("agg.jit", 2), ("no.such_knob", 1)]:
        with pytest.raises(PlanError):
            autotune.set_knob(knob, bad)
    with pytest.raises(PlanError):
        autotune.set_knob("no.such_knob", -1)
    try:
        for rows in (1, 3, 16, 20, 32):
            autotune.set_knob("filter.rows", rows)
            assert autotune.get_knob("filter.rows") == rows
    finally:
        autotune.set_knob("filter.rows", -1)
    assert autotune.get_knob("filter.rows") == -1


def test_capability_hooks_are_host_only():
    # backend.rs:189-197: asked before any upload, so they touch no device
    from oxidelake_amd._capi import lib

    assert lib.oxr_supports_hash_join() == 1
    assert lib.oxr_supports_aggregate() == 1
