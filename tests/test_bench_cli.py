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


def _served_leg():
    main = {"decoded_gbs": 600.0, "achieved_gbs": 800.0, "span_ms": 300.0, "requests": 20000,
            "decoded_bytes": 180 * 10 ** 9, "compressed_bytes": 60 * 10 ** 9, "p50_ms": 10.0, "p99_ms": 30.0,
            "verified": 1250, "verified_ok": 1250, "mismatched_requests": 0,
            "server_stats": {"gpu_union_ms": 250.0, "launches": 100}}
    sr = {"persistent": main, "passes": 2, "rate_req_per_s": 30000.0, "resident_gb": 32.0, "frac_hbm": 0.1,
          "baselines": {"one_launch_per_request": {"p50_ms": 200.0, "p99_ms": 400.0}}}
    cpu = {"value": 5.0, "unit": "GB/s", "cores": 16, "kind": "port", "sample": "served stream",
           "per_config": {"served_mixed_10k": {"gbs_all": 5.0, "gbs_1": 0.5}}}
    return {"served": sr, "cpu": cpu, "tuning": {}}


def test_headline_line_takes_the_served_leg_as_value_and_merges_the_other_leg():
    args = B.argparse.Namespace(warmup=5)
    x = {"workloads": {"int_date_mix_64M": {"kernels": {"RLE": {"us": 90.0, "alg_bytes": 10, "gbs": 5000.0}}}},
         "staged_from_file": {"ms_per_step": 4.0},
         "cpu": {"per_config": {"bp_int32_8M": {"gbs_all": 9.0, "gbs_1": 1.0}}}}
    line = B.headline_line(args, _served_leg(), x)
    assert line["value"] == 600.0 and line["unit"] == "GB/s" and line["n_gpus"] == 1
    assert line["config"]["workload"] == "served_mixed_10k" and line["steps"] == 2
    assert line["roofline"]["achieved"] == round(240e9 / 0.25 / 1e9, 1)
    assert line["roofline"]["frac"] == round(960.0 / 8000.0, 4)
    assert line["roofline_int_date_mix_64M"]["kernel"] == "RLE"
    assert set(line["cpu_baseline"]["per_config"]) == {"served_mixed_10k", "bp_int32_8M"}
    assert line["cpu_baseline"]["value"] == 5.0 and line["legs"]["extra"] == "ok"
    B.json.dumps(line)


def test_headline_line_reports_a_failed_other_leg():
    line = B.headline_line(B.argparse.Namespace(warmup=5), _served_leg(), {"rc": -11})
    assert line["value"] == 600.0 and line["workloads"] is None
    assert line["legs"]["extra"] == "exit status -11"
