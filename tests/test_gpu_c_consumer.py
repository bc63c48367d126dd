"""The C consumer (tests/c/consumer.c) built and run on the GPU: the
reference C consumer's product (crates/tprims-capi/tests/c/standard_consumer.c:
[22, 28, 49, 64] from host arrays) through the TAPP C ABI on a created
executor and on executor 0, plus the scatter extensions from C."""
import subprocess

import pytest

from test_capi_cpu import _build_c_consumer


@pytest.mark.gpu
def test_c_consumer_runs(gpu, tmp_path):
    exe = _build_c_consumer(tmp_path)
    print("running", exe, flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", (out.returncode, out.stdout, out.stderr)


@pytest.mark.gpu
def test_fixed_cost_program_runs(gpu, tmp_path):
    """tests/c/fixed_cost.c on the GPU: 2x2x2 products through executor 0, a
    created executor and the same executor asynchronously, with the result
    checked in the program ([23, 34, 31, 46])."""
    import json

    from test_capi_cpu import build_fixed_cost

    exe = build_fixed_cost(tmp_path)
    print("running", exe, flush=True)
    out = subprocess.run([exe, "500"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["correct"] is True
    print(rec)
