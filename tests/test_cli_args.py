"""The two command lines refuse bad arguments before they touch a device:
ACHRCuda <model.mps> <warmup.csv> <nFiles> <pointsPerFile> <stepsPerPoint>
(ACHRCuda.cu:552-554, read there with atoi and not checked) and
createWarmupPts <model.mps> <nPts> [-1] (createWarmupPtsGLPK.c:204-230).
Usage errors exit with status 2 and name the argument; a missing model file
exits with status 1 and says which file."""
import os
import subprocess

import pytest

import achr


def run(args, tmp_path):
    assert os.access(args[0], os.X_OK), f"{args[0]} is not built"
    return subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,code,msg", [
    ([], 2, "Usage"),
    (["m.mps", "w.csv", "1", "5"], 2, "Usage"),
    (["m.mps", "w.csv", "x", "5", "5"], 2, "bad nFiles 'x'"),
    (["m.mps", "w.csv", "1", "0", "5"], 2, "bad pointsPerFile '0'"),
    (["m.mps", "w.csv", "1", "5", "-1"], 2, "bad stepsPerPoint '-1'"),
    (["m.mps", "w.csv", "1", "5", "5", "--lossy"], 2, "Usage"),
    (["m.mps", "w.csv", "1", "5", "5"], 1, "cannot open MPS file m.mps"),
])
def test_achrcuda_arguments(tmp_path, args, code, msg):
    r = run([achr.CLI_PATH] + args, tmp_path)
    assert r.returncode == code and msg in r.stderr, (r.returncode, r.stderr)
    assert not list(tmp_path.iterdir())  # no File<k> written


@pytest.mark.parametrize("args,code,msg", [
    ([], 2, "Usage"),
    (["m.mps", "0"], 2, "bad nPts '0'"),
    (["m.mps", "ten"], 2, "bad nPts 'ten'"),
    (["m.mps", "10", "-2"], 2, "third argument must be -1"),
    (["m.mps", "10"], 1, "cannot open MPS file m.mps"),
])
def test_create_warmup_pts_arguments(tmp_path, args, code, msg):
    r = run([achr.WARMUP_CLI_PATH] + args, tmp_path)
    assert r.returncode == code and msg in r.stderr, (r.returncode, r.stderr)
    assert not list(tmp_path.iterdir())
