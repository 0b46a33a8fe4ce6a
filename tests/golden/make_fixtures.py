"""Writes tests/golden/*.npz from the reference's own test inputs and golden
outputs (tests/01..04 inputs, tests/99_ref goldens).  Run in a container
where the reference is mounted:

    python tests/golden/make_fixtures.py [/root/reference]

Only data is copied: interaction tables, species counts, golden
configurations, energy trajectories, radial densities and diagnostics.
"""
import os
import sys

import numpy as np
from scipy.io import netcdf_file

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def vij(path, S, R):
    vals = np.array(open(path).read().split(), dtype=np.float64)
    return vals[: S * S * R].reshape((S, S, R), order="F")


def nc_config(path):
    with netcdf_file(path, "r", mmap=False) as nc:
        a = nc.variables["configuration"][:].copy()
    # stored (z, y, x, b) -> Fortran (x, y, z)
    return np.asfortranarray(a[:, :, :, 0].transpose(2, 1, 0).astype(np.int8))


def nc_rho(path):
    with netcdf_file(path, "r", mmap=False) as nc:
        rho = nc.variables["rho data"][:].copy()   # (T, r, j, i)
        r = nc.variables["r data"][:].copy()
        U = nc.variables["U data"][:].copy()
    return np.asfortranarray(np.transpose(rho, (3, 2, 1, 0))), r, U


def traj(path):
    rows = [l.split() for l in open(path).read().splitlines()[1:] if l.strip()]
    return np.array([int(r[0]) for r in rows]), np.array([float(r[1]) for r in rows])


def diag(path):
    row = open(path).read().splitlines()[1].split()
    return np.array([float(x) for x in row])


t = os.path.join(REF, "tests")
g = os.path.join(t, "99_ref")

# 01: serial Metropolis, fcc 4^3, 5 species (51,51,51,51,52), 6 shells
d = os.path.join(g, "01_serial_metropolis")
steps, e = traj(os.path.join(d, "proc_0000_energy_trajectory_at_T_0300.0.dat"))
rho, r, U = nc_rho(os.path.join(d, "proc_0000_rho_of_T.nc"))
np.savez_compressed(
    os.path.join(OUT, "ref01_serial_metropolis.npz"),
    lattice=1, n=np.array([4, 4, 4]), species_numbers=np.array([51, 51, 51, 51, 52]),
    v_ex=vij(os.path.join(t, "01_serial_metropolis", "fcc_epi.vij"), 5, 6),
    T=300.0, n_mc_steps=256, n_sample_steps=1, wc_range=3, seed=110179,
    initial_config=nc_config(os.path.join(d, "proc_0000_initial_config_at_0300.0.nc")),
    final_config=nc_config(os.path.join(d, "proc_0000_final_config_at_0300.0.nc")),
    traj_steps=steps, traj_energy=e, rho=rho, shells=r, U=U)

# 02: four independent chains (ranks 0..3), bcc 4^3, concentrations 0.25
d = os.path.join(g, "02_parallel_metropolis")
rec = dict(lattice=2, n=np.array([4, 4, 4]), concentrations=np.full(4, 0.25),
           v_ex=vij(os.path.join(t, "02_parallel_metropolis", "bcc_epi.vij"), 4, 6),
           T=300.0, n_mc_steps=128, n_sample_steps=1, wc_range=2)
for rank in range(4):
    p = os.path.join(d, f"proc_{rank:04d}_")
    rec[f"seed{rank}"] = 110179 + 11 * rank
    rec[f"initial_config{rank}"] = nc_config(p + "initial_config_at_0300.0.nc")
    rec[f"final_config{rank}"] = nc_config(p + "final_config_at_0300.0.nc")
    rec[f"traj_energy{rank}"] = traj(p + "energy_trajectory_at_T_0300.0.dat")[1]
    rec[f"diagnostics{rank}"] = diag(p + "energy_diagnostics.dat")
    rec[f"rho{rank}"] = nc_rho(p + "rho_of_T.nc")[0]
rec["av_diagnostics"] = diag(os.path.join(d, "av_energy_diagnostics.dat"))
np.savez_compressed(os.path.join(OUT, "ref02_parallel_metropolis.npz"), **rec)

# 03: nested sampling, fcc 3^3, 5 species (21,21,21,21,24), 4 shells
d = os.path.join(g, "03_serial_nested_sampling")
lines = open(os.path.join(d, "fcc_al_1.00_crfeconi_K100.energies")).read().splitlines()
np.savez_compressed(
    os.path.join(OUT, "ref03_nested_sampling.npz"),
    lattice=1, n=np.array([3, 3, 3]), species_numbers=np.array([21, 21, 21, 21, 24]),
    v_ex=vij(os.path.join(t, "03_serial_nested_sampling", "fcc_al_1.00_crfeconi.vij"), 5, 4),
    n_walkers=100, n_steps=500, n_iter=1000, seed=110179,
    header=np.array(lines[0].split()),
    e_limits=np.array([float(l.split()[1]) for l in lines[1:]]))

# 04: Wang-Landau (time-seeded reference run: statistical golden)
d = os.path.join(g, "04_parallel_wang-landau")
with netcdf_file(os.path.join(d, "wl_dos.nc"), "r", mmap=False) as nc:
    lng = nc.variables["grid data"][:].copy()
np.savez_compressed(
    os.path.join(OUT, "ref04_wang_landau.npz"),
    lattice=2, n=np.array([4, 4, 4]), concentrations=np.full(4, 0.25),
    v_ex=vij(os.path.join(t, "04_parallel_wang-landau", "bcc_epi.vij"), 4, 6),
    bins=512, num_windows=4, bin_overlap=0.25, energy_min=-96.0, energy_max=0.0,
    wl_f=0.05, tolerance=5e-5, flatness=0.90, mc_sweeps=100, lng=lng)
print("fixtures written to", OUT)


# ---------------------------------------------------------------------------
# Output files of the reference's tests, kept as they are (data), for the
# format tests of brawl_amd/io.py.
# ---------------------------------------------------------------------------
GOLDEN_FILES = [
    ("01_serial_metropolis", "proc_0000_energy_trajectory_at_T_0300.0.dat"),
    ("01_serial_metropolis", "proc_0000_asro_trajectory_at_T_0300.0.dat"),
    ("01_serial_metropolis", "proc_0000_final_config_at_0300.0.nc"),
    ("01_serial_metropolis", "proc_0000_rho_of_T.nc"),
    ("02_parallel_metropolis", "proc_0000_energy_diagnostics.dat"),
    ("02_parallel_metropolis", "av_energy_diagnostics.dat"),
    ("03_serial_nested_sampling", "fcc_al_1.00_crfeconi_K100.energies"),
]


def copy_golden_files():
    import shutil
    dst = os.path.join(OUT, "files")
    os.makedirs(dst, exist_ok=True)
    for test, name in GOLDEN_FILES:
        shutil.copyfile(os.path.join(REF, "tests", "99_ref", test, name),
                        os.path.join(dst, f"{test[:2]}_{name}"))


if __name__ == "__main__" and os.path.isdir(os.path.join(REF, "tests", "99_ref")):
    copy_golden_files()
