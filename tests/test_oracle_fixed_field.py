"""The CPU fixed-field oracle: its adjoint gradient against central finite differences of its
own energy (the gradient has no reference counterpart; this is what pins it), and the energy's
invariance under the QR cadence."""
import numpy as np

from afqmclab_amd.model import HubbardModel, spin_block_trial
from oracle import fixed_field
from oracle.models import square_lattice_K


def case():
    L, nup = 9, 3
    model = HubbardModel.square(3, 3, 2 * nup, 3.0, 1.0, -0.2)
    phi = spin_block_trial(model, nup, nup)
    K = square_lattice_K(3, 3, 1.0, -0.2)[:L, :L].real
    return K, phi[:L, :nup], phi[L:, nup:]


def test_adjoint_matches_finite_differences():
    K, pu, pd = case()
    args = (pu, pd, 3, 12, 4, 0.05, 5)
    E, kbar, ubar = fixed_field.energy_and_grad(K, 3.0, *args)
    assert abs(E - fixed_field.energy(K, 3.0, *args)) < 1e-12 * abs(E)
    h = 1e-5
    fdU = (fixed_field.energy(K, 3.0 + h, *args) - fixed_field.energy(K, 3.0 - h, *args)).real / (2 * h)
    assert abs(np.sum(ubar).real - fdU) < 1e-7 * max(1, abs(fdU))
    for (i, j) in [(0, 1), (0, 4), (2, 5), (4, 8)]:
        d = np.zeros_like(K)
        d[i, j] = d[j, i] = 1.0
        fd = (fixed_field.energy(K + h * d, 3.0, *args) - fixed_field.energy(K - h * d, 3.0, *args)).real / (2 * h)
        assert abs((kbar[i, j] + kbar[j, i]).real - fd) < 1e-7 * max(1, abs(fd)), (i, j)


def test_energy_does_not_depend_on_the_qr_cadence():
    K, pu, pd = case()
    e1 = fixed_field.energy(K, 3.0, pu, pd, 3, 12, 4, 0.05, 5)
    e2 = fixed_field.energy(K, 3.0, pu, pd, 3, 12, 12, 0.05, 5)
    assert abs(e1 - e2) < 1e-10 * abs(e1)
