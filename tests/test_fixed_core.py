"""The Q16.16 core: device header (host build) == oracle == exact closed forms."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fixed_core_against_closed_forms(tmp_path):
    exe = tmp_path / "fixed_core_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(HERE, "fixed_core_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
