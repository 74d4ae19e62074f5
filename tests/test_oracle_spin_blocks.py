"""The spin-block constraint-path oracle against the dense oracle on the same uniforms: fields,
log-weights, walkers (through invariants after QR), comb tables and block energies.  CPU only."""
import cmath

import numpy as np
import pytest

from afqmclab_amd.model import HubbardModel, spin_block_trial
from oracle.blocks import SD, SDSDOperation
from oracle.drivers import AfqmcConstraintPath, CPMethod
from oracle.models import HubbardSOC
from oracle.rng import UniformStream
from oracle.spin_blocks import AfqmcConstraintPathSpinBlocks, BlockSD, BlockSDSDOperation


def make(decomp="densitySpin", U=4.0, t2=0.0, nup=5, ndn=4, seed=7, nw=6, **kw):
    model = HubbardModel.square(4, 4, nup + ndn, U, 1.0, t2)
    om = HubbardSOC(model.L, model.N, model.K, U=model.U)
    phi = spin_block_trial(model, nup, ndn)
    m = dict(dt=0.02, thermalSize=8, writeNumber=3, measureNumberPerWrite=2, measureSkipStep=3,
             walkerSizePerThread=nw, decompType=decomp, mgsStep=4, popControlStep=5, ETAdjustStep=4,
             ETAdjustMaxSize=8, seed=seed)
    m.update(kw)
    dense = AfqmcConstraintPath(CPMethod(**m), om, UniformStream(seed), phiT=SD(wf=phi))
    block = AfqmcConstraintPathSpinBlocks(CPMethod(**m), om, UniformStream(seed), SD(wf=phi), nup)
    return model, phi, dense, block


def test_block_overlap_and_green_equal_dense():
    rng = np.random.default_rng(2)
    model, phi, *_ = make()
    L = model.L
    w = np.zeros_like(phi)
    w[:L, :5] = phi[:L, :5] + 0.4 * (rng.normal(size=(L, 5)) + 1j * rng.normal(size=(L, 5)))
    w[L:, 5:] = phi[L:, 5:] + 0.4 * (rng.normal(size=(L, 4)) + 1j * rng.normal(size=(L, 4)))
    sd = SD(wf=w, logw=0.3 - 2.9j)
    dense = SDSDOperation(SD(wf=phi), sd)
    blk = BlockSDSDOperation(BlockSD.from_dense(SD(wf=phi), 5), BlockSD.from_dense(sd, 5))
    a, b = dense.returnLogOverlap(), blk.returnLogOverlap()
    assert abs(a.real - b.real) < 1e-12 and abs(cmath.exp(1j * (a.imag - b.imag)) - 1) < 1e-12
    assert np.max(np.abs(dense.returnGreenDiagonal() - blk.returnGreenDiagonal())) < 1e-12


@pytest.mark.parametrize("decomp", ["densitySpin", "densityCharge"])
def test_block_run_equals_dense_run(decomp):
    """A whole run (initial QR + comb, ET adjust, QR, combs, measurements): identical uniform
    consumption, identical fields and comb tables, block energies to 1e-10."""
    model, phi, dense, block = make(decomp, t2=-0.2)
    dense.run()
    block.run()
    assert dense.rng.counter == block.rng.counter
    assert len(dense.records) == len(block.records)
    for a, b in zip(dense.records, block.records):
        assert a[0] == b[0]
        assert np.max(np.abs(a[1] - b[1])) < 1e-10
        assert np.array_equal(a[2], b[2])
    assert np.array_equal(dense.last_table, block.last_table)
    for x, y in zip(dense.energies, block.energies):
        for p, q in zip(x, y):
            assert abs(p - q) < 1e-10 * max(1.0, abs(p)), (p, q)
    for wd, wb, al in zip(dense.walker, block.walker, dense.walkerIsAlive):
        if not al:
            continue
        d = wb.to_dense()
        assert abs(wd.logw.real - d.logw.real) < 1e-10 * max(1.0, abs(wd.logw.real))
        # the same columns up to the sign QR may give each of them
        for k in range(d.wf.shape[1]):
            e = min(np.max(np.abs(d.wf[:, k] - wd.wf[:, k])), np.max(np.abs(d.wf[:, k] + wd.wf[:, k])))
            assert e < 1e-9 * max(1.0, np.max(np.abs(wd.wf[:, k])))


def test_block_stabilize_sign_rule():
    """The reference's rule: det R < 0 negates column 0 of Q; logw gains log|det R|."""
    rng = np.random.default_rng(4)
    up = rng.normal(size=(8, 3)) + 1j * rng.normal(size=(8, 3))
    dn = rng.normal(size=(8, 2)) + 1j * rng.normal(size=(8, 2))
    b = BlockSD(up, dn, 0.1j)
    d = b.to_dense()
    b.stabilize()
    d.stabilize()
    assert abs(b.logw - d.logw) < 1e-12
    assert np.max(np.abs(b.to_dense().wf - d.wf)) < 1e-12
