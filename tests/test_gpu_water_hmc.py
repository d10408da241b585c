"""The learned-electrostatics workload on the GPU at test size (512 rigid
SPC/E molecules): a model of the k-space electrostatics (exact reciprocal sum
over |k| <= kcut plus a network trained on plain-engine labels) as the
proposal force of hybrid Monte Carlo, judged by exact layer-1 energies, and
as the fast force of r-RESPA 1:4.

What it pins:
* exact-force HMC satisfies the detailed-balance identity
  <exp(-dH/kT)> = 1 (a symplectic, reversible proposal);
* what the learned model gains over the method without it: proposals driven
  by the real-space layer alone miss the exact Hamiltonian by tens of kT (at
  512 molecules; thousands at 4096) and are never accepted, proposals with
  the model are accepted and miss by a fraction of one kT;
* with the model as the fast force, r-RESPA's slow force is only the model's
  error, and its exact energy fluctuates less than with the real-space layer
  as the fast force.
"""
import numpy as np
import pytest

from emdee_amd import surrogate as sg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload(gpu):
    res = sg.water_hmc_workload(gpu, nmol=512, kcut=1.2, samples=30, stride=4, epochs=30, cycles=40, nsteps=10,
                                dt=1.0, seed=3, log=lambda *a: print(*a, flush=True))
    print(res, flush=True)
    return res


def test_exact_hmc_identity(workload):
    h = workload["hmc"]["exact"]
    assert h["acceptance"] > 0.5
    # <exp(-dH)> = 1: with |dH| ~ 0.3 kT the spread of exp(-dH) is ~0.3
    assert abs(h["mean_exp_minus_dH"] - 1.0) < 0.25


def test_model_proposals_gain_over_the_real_space_layer(workload):
    h = workload["hmc"]
    assert h["real_space_only"]["acceptance"] == 0.0
    assert h["real_space_only"]["median_abs_dH_kT"] > 5.0
    assert h["model"]["acceptance"] > 0.3
    assert h["model"]["median_abs_dH_kT"] < 0.1 * h["real_space_only"]["median_abs_dH_kT"]


def test_model_reproduces_held_out_kspace_forces(workload):
    ho = workload["held_out"]
    assert ho["long_range_only"]["atom_force_r2"] > 0.95
    assert ho["model"]["atom_force_r2"] > 0.9


def test_model_as_respa_fast_force(workload):
    assert workload["respa_model_fast_energy_rms_kT"] < workload["respa_real_space_fast_energy_rms_kT"]
