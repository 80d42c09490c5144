"""GPU: oxr_filter_project_noalloc, the entry point that allocates nothing,
checked through the C ABI against oxr_filter_project bit for bit (values,
NULLs, row count), on two calls in a row into the same caller buffers (the
second over a smaller input), on the generated kernel (loaded by an earlier
oxr_filter_project of the shape) and on the interpreter kernel (a shape
never loaded), and its refusals before anything runs."""
import numpy as np
import pytest

from oxidelake_amd import autotune, datagen, plan
from oxidelake_amd._capi import lib
from oxidelake_amd.backend import PlanError
from oxidelake_amd.plan import Comparison as Cmp
from oxidelake_amd.plan import Lit, col
from tests.test_filter_gpu import PREDICATES, PROJECTIONS
from tests.util import assert_batches_equal

pytestmark = pytest.mark.gpu

WIDTH = {plan.INT32: 4, plan.FLOAT32: 4, plan.INT64: 8, plan.FLOAT64: 8}


class _Ptr:
    """A device address with a shape, for RocmBackend.wrap."""

    def __init__(self, ptr: int, shape):
        self.ptr, self.shape = ptr, shape

    def data_ptr(self) -> int:
        return self.ptr


def out_buffers(backend, like, cap):
    """Device buffers with room for `cap` rows of each column of `like` (a
    result of oxr_filter_project: it tells each projection's type, dim and
    whether it can be NULL)."""
    bufs = []
    for i in range(like.num_columns):
        c = like.column(i)
        width = 4 * c.dim if c.type == plan.F32_VECTOR else WIDTH[c.type]
        bufs.append((backend.alloc_device(max(16, cap * width)), backend.alloc_device(max(4, (cap + 31) // 32 * 4)),
                     c.type, c.dim, bool(c.validity)))
    return bufs


def run_noalloc(backend, stream, dev, pred, proj, bufs, cap):
    outs = [(v.ptr, val.ptr, t, d, cap) for v, val, t, d, _ in bufs]
    rows = backend.filter_project_noalloc(stream, dev, pred, proj, outs)
    view = backend.wrap([(_Ptr(v.ptr, (rows, d) if t == plan.F32_VECTOR else (rows,)),
                          _Ptr(val.ptr, (rows,)) if nullable else None, t, d)
                         for v, val, t, d, nullable in bufs], [f"p{i}" for i in range(len(bufs))], rows=rows)
    return backend.download(stream, view)


def kernel_launches(name):
    import ctypes as C

    ms, n = C.c_double(), C.c_uint64()
    lib.oxr_profile_get(name.encode(), C.byref(ms), C.byref(n))
    return int(n.value)


# the config-2 query (random_batch's first four columns have its types), an
# arithmetic mix, and a vector projection
CASES = [datagen.filter_query(50), (PREDICATES[12], PROJECTIONS[3]), (PREDICATES[0], PROJECTIONS[1])]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_two_calls_match_filter_project_on_the_generated_kernel(backend, stream, case):
    pred, proj = CASES[case]
    rng = np.random.default_rng(100 + case)
    big, small = datagen.random_batch(rng, 50_000), datagen.random_batch(rng, 20_001)
    bufs = None
    for batch in (big, small):  # the second call's input is smaller
        dev = backend.upload(stream, batch)
        ref = backend.filter_project(stream, dev, pred, proj)  # also loads the shape's kernel
        want = backend.download(stream, ref)
        if bufs is None:
            bufs = out_buffers(backend, ref, big.num_rows)
        lib.oxr_profile_reset()
        lib.oxr_profile_enable(1)
        try:
            got = run_noalloc(backend, stream, dev, pred, proj, bufs, big.num_rows)
            ran = kernel_launches("filter_project"), kernel_launches("filter_project_vm")
        finally:
            lib.oxr_profile_enable(0)
            lib.oxr_profile_reset()
        assert ran == (1, 0), ran
        assert_batches_equal(got, want, names=False)
    for b in bufs:
        b[0].free()
        b[1].free()


def test_a_shape_never_loaded_runs_the_interpreter_kernel(backend, stream):
    rng = np.random.default_rng(7)
    # a literal no other test uses: this shape's generated kernel was never compiled
    pred = plan.or_(plan.cmp(col(1), Cmp.Lt, Lit.int32(-3)),
                    plan.cmp(plan.mod(col(0), Lit.int64(987_654_323)), Cmp.Eq, Lit.int64(5)))
    proj = [col(0), plan.mul(col(2), col(3)), col(4)]
    bufs = None
    for n in (40_000, 3_333):
        batch = datagen.random_batch(rng, n)
        dev = backend.upload(stream, batch)
        lib.oxr_profile_reset()
        lib.oxr_profile_enable(1)
        try:
            if bufs is None:
                # the result types, from a batch the interpreter filters
                autotune.set_knob("filter.jit", 0)
                try:
                    like = backend.filter_project(stream, dev, pred, proj)
                finally:
                    autotune.set_knob("filter.jit", -1)
                bufs = out_buffers(backend, like, 40_000)
                lib.oxr_profile_reset()
            got = run_noalloc(backend, stream, dev, pred, proj, bufs, 40_000)
            ran = kernel_launches("filter_project"), kernel_launches("filter_project_vm")
        finally:
            lib.oxr_profile_enable(0)
            lib.oxr_profile_reset()
        assert ran == (0, 1), ran
        # oxr_filter_project on the interpreter kernel too, so that nothing
        # loads this shape's generated kernel before the next call
        autotune.set_knob("filter.jit", 0)
        try:
            want = backend.download(stream, backend.filter_project(stream, dev, pred, proj))
        finally:
            autotune.set_knob("filter.jit", -1)
        assert_batches_equal(got, want, names=False)
    for b in bufs:
        b[0].free()
        b[1].free()


def test_refusals_touch_nothing(backend, stream):
    rng = np.random.default_rng(9)
    batch = datagen.random_batch(rng, 10_000)
    dev = backend.upload(stream, batch)
    pred, proj = CASES[0]
    like = backend.filter_project(stream, dev, pred, proj)
    bufs = out_buffers(backend, like, 10_000)
    with pytest.raises(PlanError):  # room for fewer rows than the input has
        run_noalloc(backend, stream, dev, pred, proj, bufs, 9_999)
    wrong = [(v, val, plan.FLOAT64 if t == plan.INT64 else plan.INT64, d, nl) for v, val, t, d, nl in bufs]
    with pytest.raises(PlanError):  # an output of another type than the projection's
        run_noalloc(backend, stream, dev, pred, proj, wrong, 10_000)
    # more tiles than the scratch holds: an interpreter-kernel shape (512-row
    # tiles, 262,127 of them at most) over 140M rows, refused before the
    # outputs (here large enough anyway) are touched
    n = 140_000_000
    big = datagen.table(backend, stream, [datagen.GenColumn("x", plan.INT64, datagen.GEN_SEQUENCE)], n)
    bpred = plan.cmp(plan.mod(col(0), Lit.int64(987_654_329)), Cmp.Eq, Lit.int64(1))
    vals, valid = backend.alloc_device(n * 8), backend.alloc_device((n + 31) // 32 * 4)
    with pytest.raises(PlanError, match="scratch"):
        backend.filter_project_noalloc(stream, big, bpred, [col(0)], [(vals.ptr, valid.ptr, plan.INT64, 0, n)])
    for b in [vals, valid] + [x for b in bufs for x in b[:2]]:
        b.free()
