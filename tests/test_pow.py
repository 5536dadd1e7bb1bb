"""pow_pos (acme-ecp_amd/csrc/crm_pow.hpp), the FP64 power of the precipitation terminal
velocity (microphysics.F90:453-486), built for the host and checked against the C library's
correctly rounded pow: within 2 ulp over seeded samples and every power of two, and pow's
values at 0, +inf, negative and NaN bases."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pow_pos_against_libm(tmp_path):
    exe = tmp_path / "pow_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "acme-ecp_amd", "csrc"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "native", "pow_check.cpp")], check=True)
    out = subprocess.run([str(exe), "2000000"], check=True, capture_output=True, text=True).stdout
    vals = dict(line.split(None, 1) for line in out.strip().splitlines())
    print(out)
    assert float(vals["worst_ulp"]) <= 2.0
    assert float.fromhex(vals["zero"]) == 0.0
    assert float.fromhex(vals["inf"]) == float("inf")
    assert vals["neg"].strip().lstrip("-") == "nan" and vals["nan"].strip().lstrip("-") == "nan"
    assert float.fromhex(vals["one"]) == 1.0
