"""The request server's admission core (tight_binding_amd/csrc/admission.h:
the device-memory arena, admission against the HBM budget, result rows, the
waiting-time rule, lazy slabs and a device that refuses an allocation) run
on the CPU by tests/cpp/admission.cpp over a fake device memory.  The server
(server.hip) drives the same code with hipMalloc / hipFree."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = ["whole_budget", "result_rows", "bypass_order", "best_fit_and_merge", "lazy_slabs", "retire_idle_slab",
         "device_refusal"]


def test_admission_core_on_a_fake_device(tmp_path):
    exe = os.path.join(str(tmp_path), "admission")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                           "-I" + os.path.join(ROOT, "tight_binding_amd", "csrc"), "-o", exe,
                           os.path.join(ROOT, "tests", "cpp", "admission.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    for c in CASES:
        assert "ok " + c in lines, (c, out.stdout)
    assert "all ok" in lines
