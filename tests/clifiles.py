"""Writers and readers of red's data, data-info, parameter and start files
(src/red.cuda.utility/file_util.cpp:325-391, 562-647), used by the
command-line driver tests."""
import os

import numpy as np

K_GAUSS = 1.720209895e-2
YEAR = 365.25


def write_run(directory, r, v, p, b, nbt, t0, dt, params, binary=False, stem="run"):
    """Writes stem(.info).txt|dat, parameters.txt and start_files.txt into directory."""
    ext = "dat" if binary else "txt"
    data_fn, info_fn = f"{stem}.{ext}", f"{stem}.info.{ext}"
    if binary:
        with open(os.path.join(directory, info_fn), "wb") as f:
            f.write(np.array([t0, dt], "<f8").tobytes())
            f.write(np.array([0], "<u4").tobytes())
            f.write(np.asarray(nbt, "<u4").tobytes())
        with open(os.path.join(directory, data_fn), "wb") as f:
            for i in range(len(r)):
                f.write(f"b{i}".encode().ljust(30, b"\0"))
                f.write(b[i:i + 1].tobytes())
                f.write(p[i:i + 1].tobytes())
                f.write(np.array([r["x"][i], r["y"][i], r["z"][i]], "<f8").tobytes())
                f.write(np.array([v["x"][i], v["y"][i], v["z"][i]], "<f8").tobytes())
    else:
        with open(os.path.join(directory, info_fn), "w") as f:
            f.write(f"{t0:.16e} {dt:.16e} 0 " + " ".join(str(int(n)) for n in nbt) + "\n")
        with open(os.path.join(directory, data_fn), "w") as f:
            for i in range(len(r)):
                f.write(f"b{i} {int(b['id'][i])} {int(b['body_type'][i])} {int(b['mig_type'][i])} "
                        f"{b['mig_stop_at'][i]:.16e} {p['mass'][i]:.16e} {p['radius'][i]:.16e} "
                        f"{p['density'][i]:.16e} {p['cd'][i]:.16e} {r['x'][i]:.16e} {r['y'][i]:.16e} "
                        f"{r['z'][i]:.16e} {v['x'][i]:.16e} {v['y'][i]:.16e} {v['z'][i]:.16e}\n")
    with open(os.path.join(directory, "parameters.txt"), "w") as f:
        for k, val in params.items():
            f.write(f"{k} = {val}\n")
    with open(os.path.join(directory, "start_files.txt"), "w") as f:
        f.write(f"{info_fn}\n{data_fn}\n")


def read_data(path, body_md_dtype):
    """(names, ids, positions [n,3], velocities [n,3], masses) of a data file."""
    if path.endswith(".dat"):
        raw = open(path, "rb").read()
        rec = 30 + body_md_dtype.itemsize + 32 + 48
        names, ids, rr, vv, mm = [], [], [], [], []
        for q in range(len(raw) // rec):
            chunk = raw[q * rec:(q + 1) * rec]
            names.append(chunk[:30].rstrip(b"\0").decode())
            bm = np.frombuffer(chunk[30:30 + body_md_dtype.itemsize], body_md_dtype)[0]
            ids.append(int(bm["id"]))
            off = 30 + body_md_dtype.itemsize
            mm.append(np.frombuffer(chunk[off:off + 8], "<f8")[0])
            rr.append(np.frombuffer(chunk[off + 32:off + 56], "<f8"))
            vv.append(np.frombuffer(chunk[off + 56:off + 80], "<f8"))
        return names, np.array(ids), np.array(rr), np.array(vv), np.array(mm)
    names, ids, rr, vv, mm = [], [], [], [], []
    for line in open(path):
        f = line.split()
        names.append(f[0])
        ids.append(int(f[1]))
        mm.append(float(f[5]))
        rr.append([float(x) for x in f[9:12]])
        vv.append([float(x) for x in f[12:15]])
    return names, np.array(ids), np.array(rr), np.array(vv), np.array(mm)


def internal_time(years_text):
    """parameter::set_param + transform_time: years -> day -> k day, in that order."""
    return float(years_text) * YEAR * K_GAUSS


def read_run(directory, var4, param, body_md):
    """(r, v, p, b, n_by_type, t, dt) of the text run named by start_files.txt."""
    info_fn, data_fn = open(os.path.join(directory, "start_files.txt")).read().split()[:2]
    info = open(os.path.join(directory, info_fn)).read().split()
    t, dt, nbt = float(info[0]), float(info[1]), np.array([int(x) for x in info[3:10]], np.uint32)
    rows = [line.split() for line in open(os.path.join(directory, data_fn))]
    n = len(rows)
    r, v, p, b = np.zeros(n, var4), np.zeros(n, var4), np.zeros(n, param), np.zeros(n, body_md)
    for i, f in enumerate(rows):
        b["id"][i], b["body_type"][i], b["mig_type"][i] = int(f[1]), int(f[2]), int(f[3])
        b["mig_stop_at"][i] = float(f[4])
        p["mass"][i], p["radius"][i], p["density"][i], p["cd"][i] = (float(x) for x in f[5:9])
        r["x"][i], r["y"][i], r["z"][i] = (float(x) for x in f[9:12])
        v["x"][i], v["y"][i], v["z"][i] = (float(x) for x in f[12:15])
    return r, v, p, b, nbt, t, dt


FARGO_PAR = """### Disk parameters
AspectRatio     0.05            Thickness over Radius in the disc
Sigma0          6.3661977237e-4 Surface Density at r=1
AlphaViscosity  0.001           Uniform kinematic viscosity
SigmaSlope      0.5             Slope of surface density profile.
FlaringIndex    0.0
ExcludeHill     YES

### Planet parameters
PlanetConfig    in/Jup.cfg
ThicknessSmoothing 0.6          Smoothing parameters in disk thickness

### Numerical method parameters
Transport       FARGO
InnerBoundary   OPEN            choose : OPEN or RIGID or NONREFLECTING
Disk            YES
OmegaFrame      0
Frame           COROTATING
IndirectTerm    YES

### Mesh parameters
Nrad            {n_rad}             Radial number of zones
Nsec            {n_sec}             Azimuthal number of zones (sectors)
Rmin            {r_min}             Inner boundary radius
Rmax            {r_max}             Outer boundary radius
RadialSpacing   LOGARITHMIC     Zone interfaces evenly spaced

### Output control parameters
Ntot            1000            Total number of time steps
Ninterm         20              Time steps between outputs
DT              0.314159265359  Time step length. 2PI = 1 orbit
OutputDir       out
"""


def write_fargo(directory, used_rad, dens, vrad, vtheta, par="fargo.par", crlf=False, par_text=None):
    """A FARGO output directory as the reference reads it (fargo_gas_disk.cu:15-250):
    the parameter file, used_rad.dat (n_rad + 1 ring radii, one per line) and frame 0
    (gasdens0.dat, gasvrad0.dat, gasvtheta0.dat: n_rad * n_sec doubles each)."""
    n_rad, n_sec = dens.shape
    text = par_text if par_text is not None else FARGO_PAR.format(n_rad=n_rad, n_sec=n_sec, r_min=used_rad[0],
                                                                  r_max=used_rad[-1])
    nl = "\r\n" if crlf else "\n"
    with open(os.path.join(directory, par), "w", newline="") as f:
        f.write(text.replace("\n", nl))
    with open(os.path.join(directory, "used_rad.dat"), "w", newline="") as f:
        for x in used_rad:
            f.write(f"{x:.17g}{nl}")
    for name, a in (("gasdens0.dat", dens), ("gasvrad0.dat", vrad), ("gasvtheta0.dat", vtheta)):
        with open(os.path.join(directory, name), "wb") as f:
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
