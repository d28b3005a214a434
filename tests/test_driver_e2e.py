"""The reference's run shape end to end, through the two binaries.

A mesh file is written in the reference's input format (`minifem gen`), the oracle's
timing binary runs the reference's COLORING iteration on it and prints the norms of
values and prec (src/FEM.cc:48-56), those norms become the checkings file
(src/IO.cc:42-58), and the `minifem` driver (src/main.cc:96-388, src/FEM.cc:139-285)
reads the same file, runs the iterations on the GPU with either schedule and compares
its norms with the checkings (src/FEM.cc:59-98). For FP64 the difference it writes to
numerical_results_0 must be exactly 0: the GPU values and prec are bit-identical to the
oracle's COLORING order, and both norms are summed in the reference's serial order.
"""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MINIFEM = os.path.join(ROOT, "mini-fem_amd", "bin", "minifem")
ORACLE = os.path.join(ROOT, "oracle", "bin", "minifem_oracle")


def run(cmd, cwd=None, check=True):
    print("$", " ".join(str(c) for c in cmd), flush=True)
    p = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True,
                       timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"{cmd[0]} exited {p.returncode}\n{p.stdout}\n{p.stderr}")
    return p


def differences(path):
    """The two relative differences (matrix, prec) a driver run wrote."""
    txt = open(path).read()
    d = [float(x) for x in re.findall(r"difference : (\S+)", txt)]
    assert len(d) == 2, txt
    return d


def binaries():
    for b in (MINIFEM, ORACLE):
        if not os.path.exists(b):
            pytest.fail(f"{b} is not built (run __graft_entry__.build())")


def test_driver_rejects_unknown_schedule_and_precision(tmp_path):
    """Argument errors end the driver before any GPU call, with a message."""
    binaries()
    for opt in (["--schedule", "tiledd"], ["--precision", "fp16"], ["--bogus"],
                ["--solve", "abc"], ["--solve", "0"]):
        p = run([MINIFEM, "CUBE", "ela", "1", *opt, "--data", tmp_path], check=False)
        assert p.returncode != 0
        assert "Error:" in p.stderr, p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["ela", "lap"])
def test_driver_end_to_end_matches_oracle(tmp_path, mf, op):
    binaries()
    data = tmp_path / "data"
    run([MINIFEM, "gen", "CUBE12", op, "12", "0.2", "1", "7", "--data", data])
    mesh_file = data / "CUBE12" / "inputs" / f"{op}_1_0"
    assert mesh_file.exists()
    orc = json.loads(run([ORACLE, mesh_file, op, "3", "coloring"]).stdout.strip().splitlines()[-1])
    ck = data / "CUBE12" / "checkings"
    ck.mkdir()
    mf.store_ref_assembly(ck / f"{op}_1_0", orc["values_norm"], orc["prec_norm"])

    for schedule in ("tiled", "colored"):
        work = tmp_path / f"run_{schedule}"
        work.mkdir()
        # --solve: the solver the reference's main loop names (src/main.cc:353), on the
        # last iteration's operator and prec; the numerical check is unchanged by it
        out = run([MINIFEM, "CUBE12", op, "3", "--schedule", schedule, "--solve", "1e-9",
                   "--data", data], cwd=work).stdout
        assert f"Schedule               :  {schedule}, fp64" in out
        if schedule == "tiled":
            assert "Building the tiled plan" in out
        assert differences(work / "numerical_results_0") == [0.0, 0.0], (schedule, out)
        it, rel = re.search(r"done .*?, (\d+) iterations, relative residual (\S+)\)", out).groups()
        assert 0 < int(it) < 100000 and float(rel) <= 1e-9, out

    # the driver's own checkings (--store-check) hold the oracle's norms exactly
    data2 = tmp_path / "data2"
    run([MINIFEM, "gen", "CUBE12", op, "12", "0.2", "1", "7", "--data", data2])
    work = tmp_path / "run_store"
    work.mkdir()
    run([MINIFEM, "CUBE12", op, "3", "--store-check", "--data", data2], cwd=work)
    m, p = mf.read_ref_assembly(data2 / "CUBE12" / "checkings" / f"{op}_1_0")
    assert (m, p) == (orc["values_norm"], orc["prec_norm"])

    # mixed precision: the norms within the tolerance stated for that mode
    work = tmp_path / "run_mixed"
    work.mkdir()
    run([MINIFEM, "CUBE12", op, "3", "--precision", "mixed", "--data", data], cwd=work)
    dm, dp = differences(work / "numerical_results_0")
    assert dm <= 1e-6 and dp <= 1e-6, (dm, dp)
