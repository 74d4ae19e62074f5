"""End-to-end energy of the constraint-path run on the 4x4 periodic Hubbard lattice (t = 1,
U = 4, 5 up + 5 down, free-electron trial, densitySpin, dt = 0.01, 128 walkers, 40 thermal
+ 200 measurement steps) against exact diagonalisation of the same sector
(tests/golden/hubbard4x4_u4_n5u5d_ed.json, written by make_hubbard4x4_ed.py): the mean of the
block energies agrees with E0 within three standard errors (plus 1e-3 relative for the
constraint's systematic bias, which is below that at this U)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.cp import AfqmcConstraintPath, CPMethod  # noqa: E402
from afqmclab_amd.model import HubbardModel, free_electron_trial  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hubbard4x4_u4_n5u5d_ed.json")


def block_stats(energies):
    e = np.array([(h / d).real for h, d, *_ in energies])
    return e.mean(), e.std(ddof=1) / np.sqrt(len(e)), e


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_cp_energy_4x4_u4_matches_exact_diagonalisation(precision):
    ed = json.load(open(GOLDEN))
    model = HubbardModel.square(4, 4, 10, 4.0)
    phi = free_electron_trial(model)
    m = CPMethod(dt=0.01, thermalSize=40, writeNumber=20, measureNumberPerWrite=2, measureSkipStep=5,
                 walkerSizePerThread=128, decompType="densitySpin", forceType="dynamicForce", forceCap=1.5,
                 mgsStep=10, popControlStep=40, ETAdjustStep=10, ETAdjustMaxSize=40, seed=8501)
    drv = AfqmcConstraintPath(ctx(), model, m, phi, spin_blocks=False, precision=precision)
    drv.run()
    mean, err, e = block_stats(drv.energies)
    print(f"[ed] precision={precision} E = {mean:.5f} +- {err:.5f}, exact {ed['E0']:.6f}")
    assert abs(mean - ed["E0"]) < 3 * err + 1e-3 * abs(ed["E0"])
