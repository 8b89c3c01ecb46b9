"""The C++ runner (nns_amd/nbx_runner), the counterpart of the reference's
./bin/runner (3D/CUDA/CUDA_NBody/src/main.cu:19-190): it takes the reference's
options (-n -i -b -g -R -r -v), prints the reference's log lines -- per stage
and per step as src/BarnesHut.cu:210-298 writes them, the totals of
main.cu:152-181 -- filtered by the -v count as main.cu:63-78 sets the level,
and exits 0.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nns_amd", "nbx_runner")

STEP_LINES = ("Reset arrays", "Bounding box", "Building tree", "Center of mass", "Sort particles", "Compute forces",
              "Update particles", "Copying to host")
TOTALS = ("reset arrays", "compute bounding boxes", "build tree", "compute COM", "sort", "compute forces",
          "update bodies", "copy from device to host")
NUM = r"([0-9.e+-]+)"


def _run(args):
    assert os.access(RUNNER, os.X_OK), "nns_amd/nbx_runner is not built (python -m nns_amd.build --runner)"
    cmd = [RUNNER, *args]
    print("case runner", " ".join(args), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("extra", [[], ["--locked-build", "--model", "plummer"], ["--copy", "--model", "star_disk"],
                                   ["--model", "star_disk", "--suns", "1"]])
def test_runner_reference_options_and_log_lines(extra):
    out = _run(["-n", "8192", "-i", "3", "-b", "128", "-g", "512", "-R", "5", "-r", "-v", *extra])
    for s in range(3):
        assert re.search(rf"^\[TIME \] Elapsed time for step {s} : {NUM} ms$", out, re.M), s
    for name in STEP_LINES:
        got = re.findall(rf"^\[TIME \] \t{name}: {NUM} ms$", out, re.M)
        assert len(got) == 3, name
        if name != "Copying to host":
            assert all(float(v) > 0 for v in got), name
    copy = [float(v) for v in re.findall(rf"^\[TIME \] \tCopying to host: {NUM} ms$", out, re.M)]
    assert all((v > 0) == ("--copy" in extra) for v in copy)  # the reference's per-step D2H copy
    for name in TOTALS:
        assert re.search(rf"^\[TIME \] Time to {name}: {NUM}ms\n\[TIME \] \tper step: {NUM}ms$", out, re.M), name
    tot = re.search(rf"^\[TIME \] TOTAL TIME: {NUM}ms\n\[TIME \] \tper step: {NUM}ms\n"
                    rf"\[TIME \] TOTAL TIME \(without copying\): {NUM}ms\n\[TIME \] \tper step: {NUM}ms$", out, re.M)
    assert tot and float(tot.group(1)) >= float(tot.group(3)) > 0
    assert "[INFO ] ----------------------FINISHED----------------------" in out
    assert "[INFO ] Number of particles: 8192" in out and "[INFO ] Number of iterations: 3" in out
    if "--suns" in extra:
        assert "[INFO ] Number of suns (outside the tree): 1" in out


@pytest.mark.parametrize("count,shown", [(0, "TIME"), (1, "INFO"), (2, "ERROR"), (3, "WARN"), (4, "DEBUG")])
def test_runner_verbosity_is_a_repeat_count(count, shown):
    """-v given k times sets the level as main.cu:63-78: 0 TIME, 1 INFO,
    2 ERROR, 3 WARN, 4 DEBUG; a line prints when its level is at or above
    it (DEBUG < WARN < ERROR < INFO < TIME)."""
    flags = [] if count == 0 else (["-" + "v" * count] if count < 4 else ["-v"] * count)
    out = _run(["-n", "4096", "-i", "1", *flags])
    order = ["DEBUG", "WARN", "ERROR", "INFO", "TIME"]
    for lvl in order:
        label = f"[{lvl:<5}] {lvl} output"
        assert (label in out) == (order.index(lvl) >= order.index(shown)), (lvl, shown)
    assert ("Number of particles" in out) == (order.index("INFO") >= order.index(shown))
    assert re.search(r"^\[TIME \] Elapsed time for step 0 : ", out, re.M)
