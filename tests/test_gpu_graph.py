"""Batching on the GPU: GEMM tile shapes and the autotuner, HIP-graph capture of the CP step,
and one graph for several independent lattices.  Every replay of a graph is checked against
the same calls run eagerly on an identical propagator (same seed), with eager work and
synchronisation between the replays; results must be identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.cp import Context, CPMethod, CPPropagator  # noqa: E402
from afqmclab_amd.model import HubbardModel, free_electron_trial, spin_block_trial  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402


# the Green GEMM's epilogue needs the 64-column tiles
GREEN_TILES = tuple(t for t, name in enumerate(CPPropagator.TILES)
                    if name.split("x")[1].startswith("64") or name == "128x128")


def make(c, nx, ny, nup, U, precision=0, nw=16, seed=5, t2=0.0, blocks=True):
    model = HubbardModel.square(nx, ny, 2 * nup, U, 1.0, t2)
    phi = spin_block_trial(model, nup, nup) if blocks else free_electron_trial(model)
    m = CPMethod(dt=0.01, walkerSizePerThread=nw, ET=-1.0)
    kw = dict(nup=nup, spin_blocks=True) if blocks else dict(spin_blocks=False)
    prop = CPPropagator(c, model, m, phi, precision=precision, max_walkers=nw, **kw)
    prop.set_seed(seed, 0)
    prop.set_ET(-1.0)
    ws = prop.walkers(nw)
    ws.set_from_trial(phi)
    prop.stabilize(ws)
    return prop, ws


def state(prop, ws):
    wf, lw, al = ws.download()
    e = prop.energy()
    return wf, lw, al, prop.counter(), e["HNum"], e["den"]


def same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
        else:
            assert x == y, (x, y)


def block_of_work(prop, ws):
    """The enqueue-only sequence a driver repeats: steps, QR, comb, measurement."""
    for _ in range(3):
        prop.step(ws)
    prop.stabilize(ws)
    prop.pop_control(ws)
    prop.measure(ws)


@pytest.mark.parametrize("precision", [0, 1])
def test_every_tile_shape_gives_the_same_step(precision):
    ref = None
    for t in range(len(CPPropagator.TILES)):
        prop, ws = make(ctx(), 4, 4, 5, 4.0, precision=precision)
        prop.set_tiles(propagate=t, overlap=t, green=t if t in GREEN_TILES else 0)
        for _ in range(2):
            prop.step(ws)
        prop.measure(ws)
        st = state(prop, ws)
        if ref is None:
            ref = st
        else:
            same(st, ref)
        ws.close()
        prop.close()


def test_autotune_keeps_state_and_results():
    prop_a, ws_a = make(ctx(), 6, 6, 13, 4.0, precision=1, nw=64)
    prop_b, ws_b = make(ctx(), 6, 6, 13, 4.0, precision=1, nw=64)
    before = ws_a.download()
    tiles, ms = prop_a.autotune(ws_a, reps=3)
    after = ws_a.download()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    nt = len(CPPropagator.TILES)
    assert tiles.shape == (3,) and all(0 <= t < nt for t in tiles) and tiles[2] in GREEN_TILES
    assert ms.shape == (3, nt) and np.all(ms[:, 0] > 0)
    assert all(ms[2, t] == -1.0 for t in range(nt) if t not in GREEN_TILES)
    for _ in range(3):
        prop_a.step(ws_a)
        prop_b.step(ws_b)
    same(state(prop_a, ws_a), state(prop_b, ws_b))


@pytest.mark.parametrize("blocks", [True, False])
def test_graph_replays_match_eager_calls(blocks):
    c_graph, c_eager = Context(0), Context(0)
    pg, wg = make(c_graph, 4, 4, 5, 4.0, blocks=blocks)
    pe, we = make(c_eager, 4, 4, 5, 4.0, blocks=blocks)
    c_graph.capture_begin()
    block_of_work(pg, wg)
    g = c_graph.capture_end()
    assert g.nodes() > 10
    # nothing ran during the capture
    same(state(pg, wg)[:4], state(pe, we)[:4])
    for rep in range(3):
        g.launch()
        block_of_work(pe, we)
        c_graph.synchronize()
        same(state(pg, wg), state(pe, we))
        # eager work on the graph's own walkers between replays
        pg.step(wg)
        pe.step(we)
    g.close()


def test_group_graph_of_independent_lattices_matches_per_instance_calls():
    shapes = [(4, 4, 5, 4.0), (6, 6, 13, 3.0), (6, 4, 9, 6.0)]
    origin = Context(0)
    gctx = [Context(0) for _ in shapes]
    ectx = Context(0)
    graph_sets = [make(c, *s, precision=1, seed=11 + i) for i, (c, s) in enumerate(zip(gctx, shapes))]
    eager_sets = [make(ectx, *s, precision=1, seed=11 + i) for i, s in enumerate(shapes)]
    for p, w in graph_sets:
        p.autotune(w, reps=2)
    origin.capture_begin(members=gctx)
    for p, w in graph_sets:
        block_of_work(p, w)
    g = origin.capture_end()
    for rep in range(2):
        g.launch()
        for p, w in eager_sets:
            block_of_work(p, w)
        origin.synchronize()
        for (p1, w1), (p2, w2) in zip(graph_sets, eager_sets):
            same(state(p1, w1), state(p2, w2))
    g.close()


def test_calls_that_return_data_fail_inside_a_capture():
    c = Context(0)
    p, w = make(c, 4, 4, 5, 4.0)
    c.capture_begin()
    with pytest.raises(Exception, match="cannot be recorded"):
        p.log_overlap(w)
    g = c.capture_end()
    g.close()
