"""The benchmark's contract on the host side (no GPU): it refuses more than one GPU, names the
adjoint pair kernel a trajectory runs (lanes per molecule, list entry width), and takes the HBM
traffic of that kernel from the committed counter run only when the run was taken on the current
kernel sources."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    # bench.py puts rev-sim_amd at the front of sys.path when imported; keep this process's order
    saved = list(sys.path)
    try:
        spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
    finally:
        sys.path[:] = saved
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_adjoint_kernel_names():
    b = _bench()
    assert b.adj_kernel(False, 0, 32768) == "k_pair<P32, 16, true, false, false, true>"
    assert b.adj_kernel(True, 0, 32768) == "k_pair<P64, 16, true, false, false, true>"
    assert b.adj_kernel(False, 8, 32768) == "k_pair<P32, 8, true, false, false, true>"
    # beyond 65,536 molecules the list holds 32-bit entries
    assert b.adj_kernel(False, 0, 65537) == "k_pair<P32, 16, true, false, false, false>"


def test_traffic_needs_matching_sources(tmp_path, monkeypatch):
    b = _bench()
    kernel = "k_pair<P32, 16, true, false, false, true>"
    rec = {"command": "c", "scale": {}, "sources_sha256": "ab" * 32,
           "kernels": {kernel: {"fetch_bytes": 3.0e7, "write_bytes": 1.0e7}}}
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps(rec))
    monkeypatch.setattr(b, "TRAFFIC_FILE", str(f))
    t, src = b.traffic_from_profile("ab" * 32, kernel)
    assert t == 4.0e7 and src["sources_match"]
    t, src = b.traffic_from_profile("cd" * 32, kernel)
    assert t is None and not src["sources_match"]
    t, _ = b.traffic_from_profile("ab" * 32, "k_other")
    assert t is None


def test_committed_traffic_names_the_adjoint_kernel():
    """The committed counter run covers the kernel the default workload's roofline reads."""
    b = _bench()
    with open(b.TRAFFIC_FILE) as f:
        rec = json.load(f)
    for fp64 in (False, True):
        k = rec["kernels"][b.adj_kernel(fp64, 0, 32768)]
        assert k["fetch_bytes"] > 0 and k["write_bytes"] > 0
    assert len(rec["sources_sha256"]) == 64
