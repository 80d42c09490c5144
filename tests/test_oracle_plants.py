"""The oracle's pinning tests (tests/test_oracle_pinning.py) must fail when
a plausible mistake is planted in the oracle: each case copies the oracle's
C source, changes one rule, builds the copy with the oracle's own flags and
runs the pinning tests against it in a child process. CPU only; nothing
here touches the product library.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# (name, [(text in oracle/oxide_oracle.c, replacement)])
PLANTS = [
    # Float32 comparisons without DataFusion's -0.0 -> +0.0 normalisation
    ("float32_order_keeps_negative_zero",
     [("total_order_key32(positive_zero32(x)), ky = total_order_key32(positive_zero32(y))",
       "total_order_key32(x), ky = total_order_key32(y)")]),
    # IEEE remainder (round to nearest) instead of fmod's truncation
    ("float_mod_as_ieee_remainder",
     [("fmod(xf, yf)", "remainder(xf, yf)"), ("fmodf(xs, ys)", "remainderf(xs, ys)")]),
    # a positive default NaN (the GPU's) for invalid Float64 operations
    ("positive_default_nan",
     [("            res.f[r] = z;",
       "            res.f[r] = (isnan(z) && !isnan(xf) && !isnan(yf)) ? fabs(z) : z;")]),
    # Float32 MIN/MAX with zeros normalised
    ("float32_min_max_normalise_zero",
     [("int32_t kv = total_order_key32(vs), kh = total_order_key32(c->f32[g]);",
       "int32_t kv = total_order_key32(positive_zero32(vs)), kh = total_order_key32(positive_zero32(c->f32[g]));")]),
    # SUM over integers wrapping at 32 bits instead of 64
    ("int_sum_wraps_at_32_bits",
     [("c->i[g] = c->has[g] ? (int64_t)((uint64_t)c->i[g] + (uint64_t)vi) : vi;",
       "c->i[g] = c->has[g] ? (int64_t)(int32_t)(uint32_t)((uint64_t)c->i[g] + (uint64_t)vi) : vi;")]),
]


def build_planted(tmp: Path, edits) -> Path:
    dst = tmp / "oracle"
    dst.mkdir()
    for f in ("__init__.py", "oracle.py", "oxide_oracle.c", "Makefile"):
        shutil.copy(ROOT / "oracle" / f, dst / f)
    src = (dst / "oxide_oracle.c").read_text()
    for old, new in edits:
        assert src.count(old) == 1, f"the plant's anchor is not unique in oxide_oracle.c: {old}"
        src = src.replace(old, new)
    (dst / "oxide_oracle.c").write_text(src)
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                    f"-I{ROOT / 'include'}", "-shared", "-o", str(dst / "liboxide_oracle.so"),
                    str(dst / "oxide_oracle.c"), "-lm"], check=True)
    return tmp


def run_pinning(tmp: Path) -> subprocess.CompletedProcess:
    env = dict(os.environ, PYTHONPATH=f"{tmp}{os.pathsep}{ROOT}")
    return subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_pinning.py"), "-q",
                           "-p", "no:cacheprovider", f"--rootdir={tmp}"],
                          cwd=tmp, env=env, capture_output=True, text=True, timeout=300)


def test_the_unplanted_copy_passes(tmp_path):
    # the harness itself: an exact copy of the oracle passes the pinning tests
    r = run_pinning(build_planted(tmp_path, []))
    assert r.returncode == 0, r.stdout[-2000:]


@pytest.mark.parametrize("name,edits", PLANTS, ids=[p[0] for p in PLANTS])
def test_a_planted_mistake_fails_the_pinning_tests(tmp_path, name, edits):
    r = run_pinning(build_planted(tmp_path, edits))
    assert r.returncode == 1 and " failed" in r.stdout, r.stdout[-2000:]
