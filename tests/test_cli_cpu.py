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
    for sw in ("-gpu", "-idir", "-odir", "-i", "-p", "-ga", "-gf", "-id_dev", "-ef", "-pts", "-pe", "-prec"):
        assert sw in res.stdout


@pytest.mark.parametrize("args,msg", [
    (["-cpu"], "CPU path"),
    (["-n_chg", "4"], "CPU path"),
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


def _fargo_dir(tmp_path, **kw):
    """A valid run directory with a FARGO run beside it (fargo.par, used_rad.dat, frame 0)."""
    from systems import fargo_grid
    d = _dir(tmp_path, {"length": "1.0"})
    grid = fargo_grid(6, 10, 3)
    CF.write_fargo(d, *grid, **kw)
    return d, grid


def _red_fargo(d):
    return red("-gpu", "-idir", d, "-p", "parameters.txt", "-i", "start_files.txt", "-gf", "fargo.par")


FARGO_ERRORS = [
    # (what to break, the reference's message)
    ("missing_par", "could not opened"),
    ("bad_key", "Invalid parameter :bogus."),
    ("bad_number", "Invalid number at: nrad."),
    ("no_value", "Invalid key/value pair: Nsec."),
    ("no_nrad", "does not set Nrad and Nsec"),
    ("bad_radius", "Invalid number (0.4x) in file"),
    ("few_radii", "radii; Nrad + 1 = 7 expected"),
    ("short_frame", "gasvrad0.dat' has different number of data than expected!"),
    ("missing_frame", "gasvtheta0.dat' could not opened!"),
]


@pytest.mark.parametrize("case,msg", FARGO_ERRORS)
def test_fargo_file_errors(tmp_path, case, msg):
    """fargo_gas_disk's checks (fargo_gas_disk.cu:15-42, 190-250, 290-555;
    file_util.cpp:101-153), raised before the GPU is touched."""
    d, grid = _fargo_dir(tmp_path)
    par = os.path.join(d, "fargo.par")
    text = open(par).read()
    if case == "missing_par":
        os.remove(par)
    elif case == "bad_key":
        open(par, "w").write(text + "Bogus 1\n")
    elif case == "bad_number":
        open(par, "w").write(text.replace("Nrad            6", "Nrad            six"))
    elif case == "no_value":
        open(par, "w").write(text + "Nsec\n")
    elif case == "no_nrad":
        open(par, "w").write(text.replace("Nrad            6", "Ntot 5"))
    elif case == "bad_radius":
        ur = open(os.path.join(d, "used_rad.dat")).read().splitlines()
        ur[0] = "0.4x"
        open(os.path.join(d, "used_rad.dat"), "w").write("\n".join(ur) + "\n")
    elif case == "few_radii":
        ur = open(os.path.join(d, "used_rad.dat")).read().splitlines()
        open(os.path.join(d, "used_rad.dat"), "w").write("\n".join(ur[:-1]) + "\n")
    elif case == "short_frame":
        b = open(os.path.join(d, "gasvrad0.dat"), "rb").read()
        open(os.path.join(d, "gasvrad0.dat"), "wb").write(b[:-8])
    elif case == "missing_frame":
        os.remove(os.path.join(d, "gasvtheta0.dat"))
    res = _red_fargo(d)
    assert res.returncode != 0
    assert msg in res.stderr, res.stderr


@pytest.mark.parametrize("crlf", [False, True])
def test_fargo_files_are_read(tmp_path, crlf):
    """A FARGO directory written as FARGO writes it -- comment lines, trailing
    words after each value, CRLF line ends or not -- passes every file check:
    whatever the run then reports, it is none of fargo_gas_disk's errors."""
    d, _ = _fargo_dir(tmp_path, crlf=crlf)
    res = _red_fargo(d)
    for _, msg in FARGO_ERRORS:
        assert msg not in res.stderr, res.stderr
    assert "Invalid" not in res.stderr and "could not opened" not in res.stderr, res.stderr
