"""The portable logarithm's table index (ptx_common.h): the guess-and-settle
index equals a bisection of the table for guesses within one of it, around
every table boundary and at random points, and plog equals its bisection
form bit for bit (tests/plog_check.cpp, built with g++ here)."""
import os
import subprocess

from conftest import REPO


def test_plog_index_matches_bisection(tmp_path):
    exe = str(tmp_path / "plog_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(REPO, "pt-icm_amd", "csrc"),
                    "-o", exe, os.path.join(REPO, "tests", "plog_check.cpp")], check=True, timeout=120)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
