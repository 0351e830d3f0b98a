"""CPU check of the working-storage layout arithmetic (blochlib_amd/csrc/blfid_layout.h): the
batch of crystallites chosen for a workspace fits it exactly and is the largest that does, for
byte counts on the edges of every batch-size threshold (the region sizes are not linear in the
batch: each region is rounded up to 256 bytes)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_choice_fits_and_is_largest(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "layout_check.cpp")
    exe = str(tmp_path / "layout_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", src, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK"), r.stdout
