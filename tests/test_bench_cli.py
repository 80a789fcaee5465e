"""bench.py on the CPU: the one-GPU contract of its command line, and every
workload's synthetic columns decode back to the generated values through the
oracle (small row counts), so a bench line measures valid segments."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench as B  # noqa: E402
from oracle import oracle as O  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("name", ["int_date_mix_64M", "bp_int32_8M", "alp_double_32M", "copy_256M"])
def test_fixed_workloads_decode_to_their_values(name):
    for cname, values, enc in B.build_workload(name, rows=300_000):
        got = O.decode_encoded(enc, threads=4)
        np.testing.assert_array_equal(got, np.ascontiguousarray(values).view(np.uint8), err_msg=cname)


def test_string_workload_decodes_to_its_values():
    (cname, (chars, offs), enc), = B.build_workload("fsst_varchar_16M", rows=100_000)
    got_offs, got_chars = O.decode_strings(enc, threads=4)
    np.testing.assert_array_equal(got_offs.astype(np.int64), np.asarray(offs, dtype=np.int64))
    np.testing.assert_array_equal(got_chars[:int(offs[-1])], chars[:int(offs[-1])])
