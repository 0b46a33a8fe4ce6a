"""The reference's output files, written from this library's results.

Text files follow the reference's explicit Fortran formats
(metropolis_output.f90:25-135):

    energy trajectory   ' # step_number E', then (I13,x,f20.10)
    ASRO trajectory     ' # step_number ASRO', then (I13,x) + (f8.5,x) per value
    diagnostics         ' # T E C acceptance_rate', then (F8.1,2X,F24.15,2X,F24.15,2X,F6.4)

NetCDF-classic files carry the reference's variables, dimensions and global
attributes (netcdf_io.f90:150-258, 495-598):

    configuration       'configuration' int16 (z, y, x, b) + lattice attributes
    radial densities    'rho data' (T, r, j, i), 'r data', 'T data', 'U data'

The nested-sampling energies file (nested_sampling.f90:84, 121) is written
list-directed in the reference, so its spacing depends on the Fortran
compiler; it is written here in gfortran's list-directed layout and is to
be compared numerically.
"""
from __future__ import annotations

import numpy as np
from scipy.io import netcdf_file

LATTICE_NAMES = {1: "fcc", 2: "bcc"}


def _fixed(x: float, w: int, d: int) -> str:
    """Fortran Fw.d of a real: right-justified, asterisks when too wide."""
    s = f"{x:{w}.{d}f}"
    return "*" * w if len(s) > w else s


def write_energy_trajectory(path, steps, energies) -> None:
    """energy_trajectory_writer (metropolis_output.f90:25-47) for a whole
    trajectory."""
    with open(path, "w") as f:
        f.write(" # step_number E\n")
        for st, e in zip(steps, energies):
            f.write(f"{int(st):13d} {_fixed(float(e), 20, 10)}\n")


def write_asro_trajectory(path, steps, asro) -> None:
    """asro_trajectory_writer (metropolis_output.f90:49-87): asro has shape
    (S, S, wc, n_steps) in Fortran order; values are written i fastest."""
    asro = np.asarray(asro)
    with open(path, "w") as f:
        f.write(" # step_number ASRO\n")
        for k, st in enumerate(steps):
            vals = np.ravel(asro[..., k], order="F")
            f.write(f"{int(st):13d} " + "".join(_fixed(float(v), 8, 5) + " " for v in vals) + "\n")


def write_diagnostics(path, temps, energies, C, acceptance) -> None:
    """diagnostics_writer (metropolis_output.f90:99-133)."""
    with open(path, "w") as f:
        f.write(" # T E C acceptance_rate\n")
        for t, e, c, a in zip(temps, energies, C, acceptance):
            f.write(f"{_fixed(float(t), 8, 1)}  {_fixed(float(e), 24, 15)}  "
                    f"{_fixed(float(c), 24, 15)}  {_fixed(float(a), 6, 4)}\n")


def _lattice_attrs(nc, lattice: int, n, concentrations) -> None:
    nc.N_1 = np.int32(n[0])
    nc.N_2 = np.int32(n[1])
    nc.N_3 = np.int32(n[2])
    nc.__setattr__("Number of Species", np.int32(len(concentrations)))
    nc.__setattr__("Lattice Type", LATTICE_NAMES[int(lattice)].encode())


def write_config_nc(path, lattice: int, n, concentrations, config) -> None:
    """ncdf_config_writer (netcdf_io.f90:150-258): the configuration
    config(b, x, y, z) as int16 'configuration' over dimensions (z, y, x, b);
    global attributes N_basis, N_1..N_3, Number of Species, Lattice Type and
    Concentrations (species 0, the empty site, first)."""
    cfg = np.asarray(config)
    if cfg.ndim == 3:
        cfg = cfg[None, ...]                      # (b, x, y, z)
    with netcdf_file(path, "w", version=1) as nc:
        nc.N_basis = np.int32(cfg.shape[0])
        _lattice_attrs(nc, lattice, n, concentrations)
        nc.Concentrations = np.concatenate([[0.0], np.asarray(concentrations, dtype=np.float64)])
        nc.createDimension("b", cfg.shape[0])
        nc.createDimension("x", cfg.shape[1])
        nc.createDimension("y", cfg.shape[2])
        nc.createDimension("z", cfg.shape[3])
        v = nc.createVariable("configuration", "h", ("z", "y", "x", "b"))
        v[:] = np.transpose(cfg, (3, 2, 1, 0)).astype(np.int16)


def write_rho_of_T_nc(path, lattice: int, n, concentrations, interaction_file: str, rho_of_T,
                      shells, temperatures, energies) -> None:
    """ncdf_radial_density_writer (netcdf_io.f90:495-598): rho_of_T(S, S,
    wc, T_steps) (Fortran order) as 'rho data' over (T, r, j, i), with 'r
    data' (shell distances), 'T data' and 'U data' (<E>/N per T)."""
    rho = np.asarray(rho_of_T)
    S, _, wc, nT = rho.shape
    with netcdf_file(path, "w", version=1) as nc:
        _lattice_attrs(nc, lattice, n, concentrations)
        nc.__setattr__("Interaction file", interaction_file.encode())
        nc.Concentrations = np.concatenate([[0.0], np.asarray(concentrations, dtype=np.float64)])
        nc.__setattr__("Warren-Cowley Range", np.int32(wc))
        nc.createDimension("i", S)
        nc.createDimension("j", S)
        nc.createDimension("r", wc)
        nc.createDimension("T", nT)
        nc.createDimension("r_i", wc)
        nc.createDimension("T_i", nT)
        nc.createDimension("U_i", nT)
        v = nc.createVariable("rho data", "d", ("T", "r", "j", "i"))
        v[:] = np.transpose(rho, (3, 2, 1, 0))
        nc.createVariable("r data", "d", ("r_i",))[:] = np.asarray(shells, dtype=np.float64)
        nc.createVariable("T data", "d", ("T_i",))[:] = np.asarray(temperatures, dtype=np.float64)
        nc.createVariable("U data", "d", ("U_i",))[:] = np.asarray(energies, dtype=np.float64)


def read_config_nc(path) -> np.ndarray:
    """The configuration of a reference configuration file as a Fortran-
    ordered int8 (2n1, 2n2, 2n3) array (basis index 1)."""
    with netcdf_file(path, "r", mmap=False) as nc:
        a = nc.variables["configuration"][:].copy()
    return np.asfortranarray(a[:, :, :, 0].transpose(2, 1, 0).astype(np.int8))


def _list_real(x: float) -> str:
    """gfortran's list-directed REAL(8) item: 17 significant digits, fixed
    notation for 0.1 <= |x| < 1e16 (width 25 incl. padding), else
    exponent notation."""
    ax = abs(x)
    if ax == 0.0 or 0.1 <= ax < 1e16:
        digits_before = 0 if ax < 1.0 else int(np.floor(np.log10(ax))) + 1
        s = f"{x:.{max(17 - digits_before, 0)}f}"
        return f" {s:<24}"
    m, e = f"{x:.16E}".split("E")
    return f" {m}E{int(e):+04d}"


def write_ns_energies(path, n_walkers: int, n_at: int, e_limits) -> None:
    """The nested-sampling energies file (nested_sampling.f90:84, 121):
    header 'n_walkers 1 0 False n_at', then 'iteration E_limit' per
    iteration, in gfortran's list-directed layout."""
    with open(path, "w") as f:
        f.write(f"{n_walkers:12d}{1:12d}{0:12d} False{n_at:12d}\n")
        for i, e in enumerate(e_limits, start=1):
            f.write(f"{i:12d} {_list_real(float(e))}\n")


def read_ns_energies(path):
    """(header ints, iterations, energy limits) of an energies file."""
    lines = open(path).read().split("\n")
    head = lines[0].split()
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    it = np.array([int(r[0]) for r in rows])
    e = np.array([float(r[1].replace("D", "E")) for r in rows])
    return (int(head[0]), int(head[1]), int(head[2]), head[3], int(head[4])), it, e


__all__ = ["write_energy_trajectory", "write_asro_trajectory", "write_diagnostics",
           "write_config_nc", "write_rho_of_T_nc", "read_config_nc", "write_ns_energies",
           "read_ns_energies"]
