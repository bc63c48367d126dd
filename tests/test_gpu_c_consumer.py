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
