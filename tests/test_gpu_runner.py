"""The C++ runner (nns_amd/nbx_runner), the counterpart of the reference's
./bin/runner (3D/CUDA/CUDA_NBody/src/main.cu:19-190): it takes the reference's
options (-n -i -b -g -R -r -v), prints one [TIME ] line per step with -v and
the per-stage totals (BarnesHut.cu:209-298), and exits 0.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nns_amd", "nbx_runner")


@pytest.mark.parametrize("extra", [[], ["--locked-build", "--model", "plummer"], ["--copy", "--model", "star_disk"]])
def test_runner_reference_options_and_time_lines(extra):
    assert os.access(RUNNER, os.X_OK), "nns_amd/nbx_runner is not built (python -m nns_amd.build --runner)"
    cmd = [RUNNER, "-n", "8192", "-i", "3", "-b", "128", "-g", "512", "-R", "5", "-r", "-v", *extra]
    print("case runner", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    steps = re.findall(r"^\[TIME \] step (\d+): reset .* all ([0-9.]+) ms$", r.stdout, re.M)
    assert [int(s) for s, _ in steps] == [0, 1, 2]
    assert all(float(t) > 0 for _, t in steps)
    for stage in ("reset arrays", "compute bounding boxes", "build tree", "compute COM", "sort",
                  "compute forces", "update bodies", "TOTAL"):
        assert f"[TIME ] Time to {stage}:" in r.stdout
    assert "[INFO ] Number of particles: 8192" in r.stdout
    copy = re.search(r"Time to copy from device to host: ([0-9.]+) ms", r.stdout)
    assert copy and (float(copy.group(1)) > 0) == ("--copy" in extra)  # the reference's per-step D2H copy

