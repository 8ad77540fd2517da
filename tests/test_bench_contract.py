"""The benchmark's host logic, without a GPU: the counters file is used only when it
was measured on this tree's kernel sources, `--gpus N > 1` is refused, and the CPU
baseline's core count follows the affinity mask and the cgroup quota."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)

import bench  # noqa: E402


def test_committed_counters_match_the_sources():
    c = bench.read_counters(os.path.join(REPO, "profiles", "counters_nsw6_169000x96_fp64.txt"))
    assert c is not None
    assert c["kernel_sources_sha16"] == bench.kernel_sources_sha16()
    # the per-kernel bytes add up to the step's
    parts = sum(v for k, v in c.items() if k.endswith("_read_bytes") or k.endswith("_write_bytes"))
    assert abs(parts - c["bytes_per_step"]) <= 1e-6 * c["bytes_per_step"] + 16
    assert c["k_rates_valu_cycles"] > 0 and c["k_rates_fp64_flop"] > 0


def test_digest_follows_every_counted_source(tmp_path, monkeypatch):
    ref = bench.kernel_sources_sha16()
    for rel in bench.COUNTED_SOURCES:
        assert os.path.exists(os.path.join(REPO, rel)), rel
    # the same files under another root with one byte changed give another digest
    root = tmp_path / "tree"
    for rel in bench.COUNTED_SOURCES:
        dst = root / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        dst.write_bytes(open(os.path.join(REPO, rel), "rb").read())
    monkeypatch.setattr(bench, "REPO", str(root))
    assert bench.kernel_sources_sha16() == ref
    p = root / bench.COUNTED_SOURCES[0]
    p.write_bytes(p.read_bytes() + b"\n")
    assert bench.kernel_sources_sha16() != ref


def test_read_counters_missing_file(tmp_path):
    assert bench.read_counters(str(tmp_path / "none.txt")) is None


def test_host_cores_within_the_affinity_mask():
    cores, total = bench.host_cores()
    assert 1 <= cores <= len(os.sched_getaffinity(0))
    assert cores <= total


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]
