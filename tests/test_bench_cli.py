"""The benchmark's command-line contract, checked without a GPU: more than one
GPU is refused with a message and exit status 2 (one GPU is what this build
measures), the options parse, and the committed PMC summaries that the bench
line reads for `roofline.traffic` are found for the dominant kernel."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_options_parse():
    sys.path.insert(0, ROOT)
    import bench
    argv = sys.argv
    try:
        sys.argv = ["bench.py", "--steps", "2", "--warmup", "1", "--eager", "--chunk", "64",
                    "--workload", "adjoint", "--no-cpu"]
        a = bench.parse()
    finally:
        sys.argv = argv
    assert (a.steps, a.warmup, a.eager, a.chunk, a.workload, a.no_cpu) == (2, 1, True, 64,
                                                                             "adjoint", True)


def test_committed_traffic_summaries_are_read():
    sys.path.insert(0, ROOT)
    import bench
    for workload in ("rgf_energies", "rgf_forward_adjoint", "rgf_single_dominant_large"):
        t = bench.measured_traffic(workload, "gemm_kernel<double>")
        assert t is not None and t > 0, workload
