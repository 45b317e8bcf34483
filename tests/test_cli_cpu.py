"""The red-compatible command-line driver's option, parameter and start-file
handling (src/red.cuda/options.cpp:56-175, parameter.cpp:36-252): every error
here is raised before the driver touches a GPU."""
import os
import subprocess

import pytest

import clifiles as CF
import redamd as R

RED = os.path.join(os.path.dirname(os.path.abspath(R.__file__)), "red")
pytestmark = pytest.mark.skipif(not os.path.exists(RED), reason="red.cuda_amd/red is not built")


def red(*args, cwd=None):
    return subprocess.run([RED, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_usage():
    res = red("-h")
    assert res.returncode == 0
    for sw in ("-gpu", "-idir", "-odir", "-i", "-p", "-ga", "-id_dev", "-ef", "-pts", "-pe", "-prec"):
        assert sw in res.stdout


@pytest.mark.parametrize("args,msg", [
    (["-cpu"], "CPU path"),
    (["-n_chg", "4"], "CPU path"),
    (["-gf", "x"], "FARGO"),
    (["-bogus"], "Invalid switch"),
    (["-prec", "fp16"], "-prec"),
    (["-id_dev", "x"], "Invalid number"),
    (["-idir"], "Missing value"),
    (["-gpu", "-i", "start_files.txt"], "-p <file>"),
    (["-gpu", "-p", "parameters.txt"], "-i <file>"),
])
def test_refused_command_lines(args, msg):
    res = red(*args)
    assert res.returncode != 0
    assert msg in res.stderr, res.stderr


def _dir(tmp_path, params, start=None):
    d = str(tmp_path)
    r, v, p, b, nbt = R.make_disk(4, 1)
    CF.write_run(d, r, v, p, b, nbt, 0.0, 0.01, params)
    if start is not None:
        with open(os.path.join(d, "start_files.txt"), "w") as f:
            f.write(start)
    return d


@pytest.mark.parametrize("params,msg", [
    ({"lenght": "1.0"}, "Invalid parameter: lenght"),
    ({"length": "one"}, "Invalid number at: length"),
    ({"integrator": "rkf5"}, "Invalid integrator type"),
    ({"integrator": "leapfrog"}, "Invalid integrator type"),
    ({"odr": "hdf5"}, "Invalid data type"),
    ({"error check for tp": "yes"}, "Invalid value at: error check for tp"),
])
def test_bad_parameter_files(tmp_path, params, msg):
    d = _dir(tmp_path, params)
    res = red("-gpu", "-idir", d, "-p", "parameters.txt", "-i", "start_files.txt")
    assert res.returncode != 0
    assert msg in res.stderr, res.stderr
    # the error also goes to the log file, after the command line
    log = open(os.path.join(d, "log.txt")).read() if os.path.exists(os.path.join(d, "log.txt")) else ""
    assert log == "" or "Error" in log


def test_key_without_value(tmp_path):
    d = _dir(tmp_path, {"length": "1.0"})
    with open(os.path.join(d, "parameters.txt"), "a") as f:
        f.write("tolerance\n")
    res = red("-gpu", "-idir", d, "-p", "parameters.txt", "-i", "start_files.txt")
    assert res.returncode != 0 and "Invalid key/value pair" in res.stderr


@pytest.mark.parametrize("start,msg", [
    ("run.info.txt\n", "must name the data-info and the data file"),
    ("run.info.txt\nrun.csv\n", "either 'txt' or 'dat'"),
    ("missing.info.txt\nrun.txt\n", "Cannot open"),
])
def test_bad_start_files(tmp_path, start, msg):
    d = _dir(tmp_path, {"length": "1.0"}, start=start)
    res = red("-gpu", "-idir", d, "-p", "parameters.txt", "-i", "start_files.txt")
    assert res.returncode != 0
    assert msg in res.stderr, res.stderr


def test_truncated_data_file(tmp_path):
    d = _dir(tmp_path, {"length": "1.0"})
    path = os.path.join(d, "run.txt")
    lines = open(path).read().splitlines()
    with open(path, "w") as f:
        f.write("\n".join(lines[:-1]) + "\n")
    res = red("-gpu", "-idir", d, "-p", "parameters.txt", "-i", "start_files.txt")
    assert res.returncode != 0 and "Cannot read body 4" in res.stderr, res.stderr


def test_units_of_the_parameter_file():
    # parameter.cpp: years -> days (365.25) -> k days (Gauss' constant)
    assert CF.internal_time("1.0") == 365.25 * 1.720209895e-2
    assert CF.K_GAUSS == R.K_GAUSS
