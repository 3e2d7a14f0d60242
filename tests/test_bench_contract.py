"""bench.py's contract on the CPU: it runs on one GPU only and says so for
any other count, and it maps rocprofv3's kernel names to the bare function
names its per-kernel tables use."""
import os
import subprocess
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_kernel_names_reduce_to_the_function_name():
    import bench

    assert bench._kernel_key("void (anonymous namespace)::k_radix_onesweep<1024, 8, 12>"
                             "(CgsRadixArgs)") == "k_radix_onesweep"
    assert bench._kernel_key("k_composite_q<2, 2, 4>(CgsCompositeArgs)") == "k_composite_q"
    assert bench._kernel_key("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
