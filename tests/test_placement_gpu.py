"""GPU: the configs' query shapes lowered for the rocm target run on the
device through the operator (and the fused join + aggregate through
oxr_agg_update_join) and return the plain plan's rows (tests/plan_exec.py
runs the plain plans on pyarrow)."""
import pytest

from oxidelake_amd import placement as P
from tests import plan_exec
from tests.test_placement import config1, config1_tables, config2, config2_tables, config3, config3_tables

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("build,data,node", [
    (config1, lambda: config1_tables(200_000, 31), P.GpuAggregateExec),
    (config2, lambda: config2_tables(300_000, 32), P.GpuFilterExec),
    (config3, lambda: config3_tables(400_000, 20_000, 1_000, 33), P.GpuJoinAggregateExec),
], ids=["config1", "config2", "config3"])
def test_lowered_config_shapes_match_the_plain_plan(backend, build, data, node):
    tables = data()
    rule = P.HardwarePlacementRule("rocm")
    lowered = rule.optimize(build())
    assert rule.notes == [] and isinstance(lowered, node), P.explain(lowered)
    got = plan_exec.run(lowered, tables, backend)
    want = plan_exec.run(build(), tables)
    assert plan_exec.same_rows(got, want), P.explain(lowered)
