"""Wang-Landau against the reference's own result: the reference's test 04
(bcc 4^3 AlTiCrMo, 6 shells, 512 bins over [-96, 0] meV/atom, 4 windows with
25 % overlap, one walker per window as in its CI run, ln f from 0.05,
flatness 0.9, 100 sweeps per call) run through the device walkers and the
window machinery (enter_energy_window, pre-sampling, sweeps with window
averages and a replica exchange after each call, dos_average, dos_combine;
the reference's default path), then checked with the reference's own
criterion (tests/04_parallel_wang-landau/wl_test.py): RMSE of ln g against the
golden wl_dos.nc divided by its mean < 1 %.  The run stops once ln f has
fallen below 4e-4 (seven halvings; the reference goes on to 5e-5).  The
walkers' random streams are the reference's static ones, so the run is
deterministic."""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def test_wang_landau_reproduces_reference_dos():
    from brawl_amd import _lib, wang_landau
    _lib.load()
    z = load_golden("ref04_wang_landau.npz")
    wl = wang_landau.WangLandau(int(z["lattice"]), tuple(int(x) for x in z["n"]), z["v_ex"],
                                bins=int(z["bins"]), num_windows=int(z["num_windows"]),
                                walkers_per_window=1, bin_overlap=float(z["bin_overlap"]),
                                energy_min=float(z["energy_min"]), energy_max=float(z["energy_max"]),
                                mc_sweeps=int(z["mc_sweeps"]))
    wl.initial_setup(concentrations=z["concentrations"])
    res = wl.run(wl_f=float(z["wl_f"]), tolerance=4e-4, flatness=float(z["flatness"]),
                 max_sweeps_calls=5000, max_pre_sampling_calls=5000)
    ref = z["lng"]
    nrmse = np.sqrt(np.mean((ref - res.lng) ** 2)) / np.mean(ref)
    print("ln f halvings", res.iterations, "sweeps calls", res.sweeps_calls, "pre-sampling calls",
          res.pre_sampling_calls, "exchanges", res.exchanges, "of", res.exchange_pairs, "nrmse", nrmse,
          flush=True)
    sys.stdout.flush()
    assert res.iterations == 7
    assert nrmse < 0.01
    wl.close()
