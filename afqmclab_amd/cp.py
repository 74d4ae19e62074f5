"""Constraint-path AFQMC on the GPU: device handles and the driver loop.

``AfqmcConstraintPath`` mirrors tutorials/afqmcConstraintPath
(afqmcConstraintPath.cpp:14-158 and afqmcConstraintPathProjection.cpp:11-179):
the same step order, cadences and random-number consumption order, with the
per-walker loop replaced by batched device calls that keep walkers, weights,
ET and the random counter in HBM.
"""
import ctypes as C
import sys
from dataclasses import dataclass

import numpy as np

from ._lib import CPDesc, check, dptr, iptr, lib
from .model import HubbardModel

DECOMP = {"densityCharge": 0, "densitySpin": 1, "hopCharge": 2, "hopSpin": 3}
FORCE = {"dynamicForce": 0, "constForce": 1}


class Context:
    """One device + one HIP stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().afqmc_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().afqmc_ctx_synchronize(self.h))

    REGIONS = ["propagate", "overlap", "gauss_jordan", "green", "hs_sample", "qr", "comb", "measure",
               "ph_force", "ph_sample", "ph_taylor", "hs_apply"]

    def set_timing(self, enable=True):
        """Bracket tagged kernel launches with HIP events on this context's stream (resets totals)."""
        check(lib().afqmc_ctx_set_timing(self.h, int(bool(enable))))

    def timing(self):
        """{region: (total_ms, launches)} since set_timing."""
        n = len(self.REGIONS)
        ms = np.zeros(n)
        cnt = (C.c_longlong * n)()
        check(lib().afqmc_ctx_get_timing(self.h, dptr(ms), cnt, n))
        return {r: (float(ms[i]), int(cnt[i])) for i, r in enumerate(self.REGIONS)}

    # ---- HIP-graph capture of enqueue-only calls --------------------------
    def capture_begin(self, members=()):
        """Record (instead of run) the enqueue-only calls made until capture_end(); calls on the
        member contexts become parallel branches of the same graph."""
        members = list(members)
        if members:
            arr = (C.c_void_p * len(members))(*[m.h for m in members])
            check(lib().afqmc_capture_begin_group(self.h, len(members), arr))
        else:
            check(lib().afqmc_capture_begin(self.h))

    def capture_end(self):
        g = C.c_void_p()
        check(lib().afqmc_capture_end(self.h, C.byref(g)))
        return Graph(g)

    def close(self):
        if self.h:
            lib().afqmc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph of captured calls; launch() enqueues it (asynchronous)."""

    def __init__(self, h):
        self.h = h

    def launch(self):
        check(lib().afqmc_graph_launch(self.h))

    def nodes(self):
        return lib().afqmc_graph_node_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class CPGroup:
    """Several constraint-path runs (a propagator and one of its walker sets each, all on one
    context) advanced together: step / stabilize / pop_control act on every member as the
    propagator's own calls do, with each kernel of a phase launched once over all members."""

    def __init__(self, members, tiles=None):
        self.members = list(members)
        n = len(self.members)
        cps = (C.c_void_p * n)(*[p.h for p, _ in self.members])
        ws = (C.c_void_p * n)(*[w.h for _, w in self.members])
        h = C.c_void_p()
        check(lib().afqmc_cp_group_create(n, cps, ws, C.byref(h)))
        self.h = h
        if tiles is not None:
            self.set_tiles(*tiles)

    def set_tiles(self, propagate=0, overlap=0, green=0):
        t = np.array([propagate, overlap, green], dtype=np.int32)
        check(lib().afqmc_cp_group_set_tiles(self.h, iptr(t)))

    def step(self):
        check(lib().afqmc_cp_group_step(self.h))

    def stabilize(self):
        check(lib().afqmc_cp_group_stabilize(self.h))

    def pop_control(self):
        check(lib().afqmc_cp_group_pop_control(self.h))

    def table_count(self):
        return lib().afqmc_cp_group_table_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_cp_group_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def walkers_to_buffer(wfs):
    """List/array of 2L x N walkers -> (nw, N, 2L) C-contiguous buffer (each walker column-major)."""
    a = np.asarray(wfs, dtype=np.complex128)
    return np.ascontiguousarray(np.transpose(a, (0, 2, 1)))


def buffer_to_walkers(buf):
    return np.transpose(buf, (0, 2, 1))


@dataclass
class CPMethod:
    """The afqmc_param keys of the CP driver (afqmcConstraintPathMethod.cpp:158-180)."""
    dt: float = 0.01
    thermalSize: int = 200
    writeNumber: int = 60
    measureNumberPerWrite: int = 2
    measureSkipStep: int = 5
    walkerSizePerThread: int = 300
    decompType: str = "densitySpin"
    forceType: str = "dynamicForce"
    forceCap: float = 1.5
    mgsStep: int = 10
    popControlStep: int = 10
    ET: float = -10.0
    ETAdjustStep: int = 10
    ETAdjustMaxSize: int = 100
    seed: int = 985456376

    def analysis(self):
        """afqmcConstraintPathMethod.cpp:182-229."""
        if self.dt < 0.0:
            raise ValueError("dt must be postive!")
        if self.measureNumberPerWrite == 0:
            raise ValueError("measureNumberPerWrite = 0, code will not measure anything!")
        if self.measureSkipStep == 0:
            raise ValueError("measureSkipStep ==0, code will not measure anything!")
        if self.walkerSizePerThread <= 0:
            raise ValueError("walkerSizePerThread is not a positive number!")
        if self.forceCap < 0.0:
            raise ValueError("forceCap must be positive or zero!")
        if self.ETAdjustMaxSize > self.thermalSize:
            raise ValueError("We should not adjust ET after thermalizing!")


class CPPropagator:
    """Device-side model, propagators, HS constants, trial and workspace (afqmc_cp)."""

    def __init__(self, ctx: Context, model: HubbardModel, method: CPMethod, phiT, nup=None, spin_blocks=None,
                 precision=0, max_walkers=None, const_force=None, phiT_logw=0j):
        self.ctx = ctx
        self.model = model
        self.method = method
        L, N = model.L, model.N
        phiT = np.asfortranarray(np.asarray(phiT, dtype=np.complex128))
        if phiT.shape != (2 * L, N):
            raise ValueError("phiT must be 2L x N")
        if spin_blocks is None:
            spin_blocks = nup is not None and model.is_spin_block() and method.decompType in ("densityCharge", "densitySpin")
        self.spin_blocks = bool(spin_blocks)
        self.nup = int(nup) if nup is not None else 0
        self.precision = int(precision)
        self._keep = []

        def keep(a):
            a = np.asfortranarray(np.asarray(a))
            self._keep.append(a)
            return dptr(a)

        d = CPDesc()
        d.L, d.N, d.Nup = L, N, self.nup
        d.spin_blocks = int(self.spin_blocks)
        d.precision = self.precision
        d.decomp_type = DECOMP[method.decompType]
        d.force_type = FORCE[method.forceType]
        d.dt = method.dt
        d.force_cap = method.forceCap
        d.K = keep(model.K)
        d.exp_minus_dt_k = keep(model.exp_minus_alpha_k(method.dt))
        d.exp_minus_half_dt_k = keep(model.exp_minus_alpha_k(method.dt * 0.5))
        d.exp_half_dt_k = keep(model.exp_minus_alpha_k(-method.dt * 0.5))
        for k in range(6):
            d.hop_logw[k] = 0.0
        d.U = keep(model.U)
        d.mu = keep(model.mu)
        d.hx = keep(model.hx)
        d.hy = keep(model.hy)
        d.hz = keep(model.hz)
        d.phiT = keep(phiT)
        d.phiT_logw[0] = complex(phiT_logw).real
        d.phiT_logw[1] = complex(phiT_logw).imag
        cf = np.zeros(L) if const_force is None else np.asarray(const_force, dtype=np.float64)
        d.const_force = keep(cf)
        d.max_walkers = int(max_walkers or method.walkerSizePerThread)
        h = C.c_void_p()
        check(lib().afqmc_cp_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.phiT = phiT
        self.max_walkers = d.max_walkers

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_cp_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def hs_constants(self):
        L = self.model.L
        out = np.zeros(5 * L, dtype=np.complex128)
        check(lib().afqmc_cp_get_hs_constants(self.h, dptr(out)))
        return out[:L], out[L:].reshape(4, L)

    def set_seed(self, seed, counter=0):
        check(lib().afqmc_cp_set_seed(self.h, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), C.c_uint64(counter)))

    def counter(self):
        c = C.c_uint64()
        check(lib().afqmc_cp_get_counter(self.h, C.byref(c)))
        return c.value

    def set_uniforms(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        check(lib().afqmc_cp_set_uniforms(self.h, dptr(u) if u.size else None, u.size))

    def walker_steps(self):
        c = C.c_uint64()
        check(lib().afqmc_cp_get_walker_steps(self.h, C.byref(c)))
        return c.value

    def skip_uniforms(self, n):
        check(lib().afqmc_cp_skip_uniforms(self.h, C.c_uint64(int(n))))

    def set_ET(self, et):
        check(lib().afqmc_cp_set_ET(self.h, float(et)))

    def get_ET(self):
        v = np.zeros(1)
        check(lib().afqmc_cp_get_ET(self.h, dptr(v)))
        return float(v[0])

    def walkers(self, nw):
        return WalkerSet(self, nw)

    # ---- GEMM tile shapes -------------------------------------------------
    TILES = ["64x64", "32x64", "64x32", "32x32", "128x64", "64x64k32", "64x32k32", "32x64k32", "128x128", "256x64",
             "64x128"]

    def set_tiles(self, propagate=0, overlap=0, green=0):
        t = np.array([propagate, overlap, green], dtype=np.int32)
        check(lib().afqmc_cp_set_tiles(self.h, iptr(t)))

    def autotune(self, w, reps=5):
        """Time every tile shape on this walker set's propagate / overlap / Green GEMMs, keep the
        fastest; returns (tiles, ms[3][n_tiles])."""
        n = len(self.TILES)
        t = np.zeros(3, dtype=np.int32)
        ms = np.zeros(3 * n)
        check(lib().afqmc_cp_autotune(self.h, w.h, int(reps), iptr(t), dptr(ms)))
        return t, ms.reshape(3, n)

    # ---- batched drop-ins -------------------------------------------------
    def log_overlap(self, w):
        out = np.zeros(w.nw, dtype=np.complex128)
        check(lib().afqmc_cp_log_overlap(self.h, w.h, dptr(out)))
        return out

    def green_diagonal(self, w):
        L2 = 2 * self.model.L
        gd = np.zeros((w.nw, L2), dtype=np.complex128)
        go = np.zeros((w.nw, L2), dtype=np.complex128)
        check(lib().afqmc_cp_green_diagonal(self.h, w.h, dptr(gd), dptr(go)))
        return gd, go

    def force(self, w):
        f = np.zeros((w.nw, self.model.L))
        check(lib().afqmc_cp_force(self.h, w.h, dptr(f)))
        return f

    def onebody(self, w, which=0):
        check(lib().afqmc_cp_onebody(self.h, w.h, int(which)))

    def hs_apply(self, w, diag4, sample_logw):
        d = np.ascontiguousarray(diag4, dtype=np.complex128)
        lw = np.ascontiguousarray(sample_logw, dtype=np.complex128)
        check(lib().afqmc_cp_hs_apply(self.h, w.h, dptr(d), dptr(lw)))

    def qr_stats(self):
        """(walkers redone by shifted CholeskyQR, walkers marked dead as rank deficient)."""
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().afqmc_cp_get_qr_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stabilize(self, w):
        check(lib().afqmc_cp_stabilize(self.h, w.h))

    def apply_table(self, w, table):
        t = np.ascontiguousarray(table, dtype=np.int32)
        check(lib().afqmc_cp_apply_table(self.h, w.h, iptr(t)))

    def step(self, w, record=False):
        if record:
            f = np.zeros((w.nw, self.model.L))
            a = np.zeros((w.nw, self.model.L), dtype=np.int32)
            check(lib().afqmc_cp_step(self.h, w.h, dptr(f), iptr(a)))
            return f, a
        check(lib().afqmc_cp_step(self.h, w.h, None, None))
        return None

    def pop_control(self, w, record=False):
        if record:
            wt = np.zeros(w.nw)
            tb = np.zeros(w.nw, dtype=np.int32)
            st = np.zeros(3)
            check(lib().afqmc_cp_pop_control(self.h, w.h, dptr(wt), iptr(tb), dptr(st)))
            return wt, tb, st
        check(lib().afqmc_cp_pop_control(self.h, w.h, None, None, None))
        return None

    def step_fields(self, w, aux):
        """One step with the fields given (nw x L, +-1) instead of sampled: no uniforms drawn."""
        a = np.ascontiguousarray(aux, dtype=np.int32)
        if a.shape != (w.nw, self.model.L):
            raise ValueError("aux must be nw x L")
        check(lib().afqmc_cp_step_fields(self.h, w.h, iptr(a)))

    def pop_control_table(self, w, table):
        """popControl with the comb's table given; returns the weights."""
        t = np.ascontiguousarray(table, dtype=np.int32)
        wt = np.zeros(w.nw)
        check(lib().afqmc_cp_pop_control_table(self.h, w.h, iptr(t), dptr(wt)))
        return wt

    def measure(self, w):
        check(lib().afqmc_cp_measure(self.h, w.h))

    def energy(self):
        v = np.zeros(7, dtype=np.complex128)
        check(lib().afqmc_cp_energy(self.h, dptr(v)))
        return dict(zip(["den", "HNum", "KNum", "VNum", "RNum", "NupTotNum", "NdnTotNum"], v))

    def site_measure(self):
        v = np.zeros(4 * self.model.L, dtype=np.complex128)
        check(lib().afqmc_cp_site_measure(self.h, dptr(v)))
        return v.reshape(4, self.model.L)

    def reset_measure(self):
        check(lib().afqmc_cp_reset_measure(self.h))

    def adjust_ET(self):
        check(lib().afqmc_cp_adjust_ET(self.h))


def comb(ctx: Context, weight, uniforms):
    """popConfiguration(1, weight) on the device; returns (table, (min, max, mean))."""
    w = np.ascontiguousarray(weight, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    t = np.zeros(w.size, dtype=np.int32)
    st = np.zeros(3)
    check(lib().afqmc_comb(ctx.h, w.size, dptr(w), dptr(u), iptr(t), dptr(st)))
    return t, tuple(st)


class WalkerSet:
    """Device-resident std::vector<SD> of the driver."""

    def __init__(self, prop: CPPropagator, nw):
        h = C.c_void_p()
        check(lib().afqmc_cp_walkers_create(prop.h, int(nw), C.byref(h)))
        self.h = h
        self.nw = int(nw)
        self.prop = prop

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_walkers_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def upload(self, wfs, logw=None, alive=None):
        buf = walkers_to_buffer(wfs)
        lw = np.zeros(self.nw, dtype=np.complex128) if logw is None else np.ascontiguousarray(logw, dtype=np.complex128)
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.int32)
        check(lib().afqmc_walkers_upload(self.h, dptr(buf), dptr(lw), iptr(al)))

    def download(self):
        L2, N = 2 * self.prop.model.L, self.prop.model.N
        buf = np.zeros((self.nw, N, L2), dtype=np.complex128)
        lw = np.zeros(self.nw, dtype=np.complex128)
        al = np.zeros(self.nw, dtype=np.int32)
        check(lib().afqmc_walkers_download(self.h, dptr(buf), dptr(lw), iptr(al)))
        return buffer_to_walkers(buf), lw, al

    def set_from_trial(self, phiT=None, logw=0j):
        """setWalkerFromPhiT (SDSDOperation.cpp:176-186): every walker = phiT, alive.  Without phiT,
        the propagator's trial is copied on the device (no host walker buffers)."""
        if phiT is None:
            lw = np.array([complex(logw)], dtype=np.complex128)
            check(lib().afqmc_cp_walkers_from_trial(self.prop.h, self.h, dptr(lw)))
            return
        self.upload([phiT] * self.nw, np.full(self.nw, complex(logw)))

    def download_some(self, index):
        """The walkers index[...] as (wf list, logw, alive)."""
        idx = np.ascontiguousarray(index, dtype=np.int32)
        L2, N = 2 * self.prop.model.L, self.prop.model.N
        buf = np.zeros((len(idx), N, L2), dtype=np.complex128)
        lw = np.zeros(len(idx), dtype=np.complex128)
        al = np.zeros(len(idx), dtype=np.int32)
        check(lib().afqmc_walkers_download_some(self.h, len(idx), iptr(idx), dptr(buf), dptr(lw), iptr(al)))
        return buffer_to_walkers(buf), lw, al

    def alive(self):
        """The alive mask of every walker (no walker data is copied)."""
        al = np.zeros(self.nw, dtype=np.int32)
        check(lib().afqmc_walkers_download(self.h, None, None, iptr(al)))
        return al

    def set_alive(self, alive):
        check(lib().afqmc_walkers_set_alive(self.h, iptr(np.ascontiguousarray(alive, dtype=np.int32))))


class AfqmcConstraintPath:
    """Device-resident CP driver with the reference's step loop.

    run() follows afqmcConstraintPath.cpp:14-30: initial walkers
    (sampleFromPhiT) -> QR + comb -> one aux sample skipped
    (estimateMemory) -> measureWithProjection -> writes of (HNum, den) per
    block.  ``records`` keeps the per-block energies."""

    def __init__(self, ctx, model, method: CPMethod, phiT, nup=None, spin_blocks=None, precision=0,
                 uniforms=None, seed=None):
        method.analysis()
        self.ctx = ctx
        self.model = model
        self.method = method
        self.prop = CPPropagator(ctx, model, method, phiT, nup=nup, spin_blocks=spin_blocks, precision=precision,
                                 max_walkers=method.walkerSizePerThread)
        if uniforms is not None:
            self.prop.set_uniforms(uniforms)
        else:
            self.prop.set_seed(method.seed if seed is None else seed, 0)
        self.prop.set_ET(method.ET)
        self.walker = self.prop.walkers(method.walkerSizePerThread)
        self.energies = []
        self.steps_done = 0
        self.walker_steps = 0

    def initialWalker(self):
        self.walker.set_from_trial(self.prop.phiT)
        self.prop.stabilize(self.walker)
        self.prop.pop_control(self.walker)

    def estimateMemory(self):
        """The reference draws one aux sample here (afqmcConstraintPath.cpp:66): Ls uniforms."""
        self.prop.skip_uniforms(self.model.L)

    def projectOneStep(self, idx):
        self.prop.step(self.walker)
        self.steps_done += 1
        self.walker_steps += self.walker.nw
        idx[0] += 1
        if idx[0] == self.method.mgsStep:
            self.prop.stabilize(self.walker)
            idx[0] = 0
        idx[1] += 1
        if idx[1] == self.method.popControlStep:
            self.prop.pop_control(self.walker)
            idx[1] = 0

    def measureWithProjection(self, on_block=None):
        m = self.method
        self.prop.onebody(self.walker, 1)
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAdjustMaxSize and i % m.ETAdjustStep == 0:
                self.prop.measure(self.walker)
                self.prop.adjust_ET()
            self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                self.prop.measure(self.walker)
                for k in range(m.measureSkipStep):
                    self.projectOneStep(idx)
            e = self.prop.energy()
            self.energies.append((e["HNum"], e["den"], e["KNum"], e["VNum"], e["RNum"]))
            self.prop.reset_measure()
            if on_block:
                on_block(i, e)

    def run(self):
        self.initialWalker()
        self.estimateMemory()
        self.measureWithProjection()
        return self.energies

    def energy(self):
        H = sum(e[0] for e in self.energies)
        D = sum(e[1] for e in self.energies)
        return H / D


# ------------------------------------------------------------------ back-propagation
@dataclass
class BPMethod(CPMethod):
    """The afqmc_param keys of the back-propagation driver: the CP keys and backPropagationStep
    (tutorials/afqmcConstraintPathBP/source/afqmcConstraintPathMethod.cpp:19-46, default 100)."""
    backPropagationStep: int = 100

    def analysis(self):
        super().analysis()
        if self.backPropagationStep < 0:
            raise ValueError("backPropagationStep must be non-negative")


BP_MATRICES = ("greenMatrixNum", "densityDensityNum", "splusSminusNum", "sminusSplusNum", "spairSpairNum")


class BackPropagation:
    """Device records and accumulators of the back-propagated estimator (afqmc_bp)."""

    def __init__(self, prop: CPPropagator, nw, bp_steps, measure_skip, mgs):
        h = C.c_void_p()
        check(lib().afqmc_bp_create(prop.h, nw, bp_steps, measure_skip, mgs, C.byref(h)))
        self.h = h
        self.L = prop.model.L

    def begin(self):
        check(lib().afqmc_bp_begin(self.h))

    def step(self, w):
        check(lib().afqmc_bp_step(self.h, w.h))

    def pop_control(self, w):
        check(lib().afqmc_bp_pop_control(self.h, w.h, None))

    def measure(self, w):
        check(lib().afqmc_bp_measure(self.h, w.h))

    def get(self):
        """{den, HNum, KNum, VNum, RNum, greenMatrixNum, densityDensityNum, splusSminusNum,
        sminusSplusNum, spairSpairNum} as HubbardSOCMeasureObserveSDSD holds them."""
        L = self.L
        sc = np.zeros(5, dtype=np.complex128)
        g = np.zeros((2 * L, 2 * L), dtype=np.complex128, order="F")
        dd = np.zeros((2 * L, 2 * L), dtype=np.complex128, order="F")
        pm, mp, sp = (np.zeros((L, L), dtype=np.complex128, order="F") for _ in range(3))
        check(lib().afqmc_bp_get(self.h, dptr(sc), dptr(g), dptr(dd), dptr(pm), dptr(mp), dptr(sp)))
        out = dict(zip(["den", "HNum", "KNum", "VNum", "RNum"], sc))
        out.update(zip(BP_MATRICES, (g, dd, pm, mp, sp)))
        return out

    def reset(self):
        check(lib().afqmc_bp_reset(self.h))

    def steps(self):
        t, m = C.c_int64(), C.c_int64()
        check(lib().afqmc_bp_get_step(self.h, C.byref(t), C.byref(m)))
        return t.value, m.value

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_bp_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class AfqmcConstraintPathBP(AfqmcConstraintPath):
    """The back-propagation driver (tutorials/afqmcConstraintPathBP/source/afqmcConstraintPath.cpp:115-180):
    the constraint-path step loop with every step after thermalisation recorded on the device and a
    back-propagated measurement where the reference's addMeasurement fires.  ``writes`` holds one
    dict of accumulated numerators per write, as HubbardSOCMeasureObserveSDSD::write puts them on
    disk; ``betas`` the matching beta.dat values."""

    def __init__(self, ctx, model, method: BPMethod, phiT, nup=None, spin_blocks=None, precision=0, uniforms=None,
                 seed=None):
        super().__init__(ctx, model, method, phiT, nup=nup, spin_blocks=spin_blocks, precision=precision,
                         uniforms=uniforms, seed=seed)
        self.bp = BackPropagation(self.prop, method.walkerSizePerThread, method.backPropagationStep,
                                  method.measureSkipStep, method.mgsStep)
        self.recording = False
        self.isBP = False
        self.BPIndex = 0
        self.writes = []
        self.betas = []

    def projectOneStep(self, idx):
        if self.recording:
            self.bp.step(self.walker)
        else:
            self.prop.step(self.walker)
        self.steps_done += 1
        self.walker_steps += self.walker.nw
        idx[0] += 1
        if idx[0] == self.method.mgsStep:
            self.prop.stabilize(self.walker)
            idx[0] = 0
        idx[1] += 1
        if idx[1] == self.method.popControlStep:
            if self.recording:
                self.bp.pop_control(self.walker)
            else:
                self.prop.pop_control(self.walker)
            idx[1] = 0
        if self.isBP:
            self.BPIndex += 1

    def prepareBackPropagation(self):
        """afqmcConstraintPathBackPropagation.cpp:11-20; the device keeps recording from the first call."""
        if not self.recording:
            self.bp.begin()
            self.recording = True
        self.isBP = True
        self.BPIndex = 0

    def addMeasurement(self):
        """afqmcConstraintPathMeasure.cpp:41-72 with the window bookkeeping of
        prepareBackPropagationMeasAndNextBackup (afqmcConstraintPathBackPropagation.cpp:59-76)."""
        if not self.isBP:
            return
        m = self.method
        self.bp.measure(self.walker)
        if m.backPropagationStep >= m.measureSkipStep:
            self.BPIndex = m.backPropagationStep - m.measureSkipStep
        else:
            self.isBP = False

    def measureWithProjection(self, on_block=None):
        m = self.method
        nbp, skip = m.backPropagationStep, m.measureSkipStep
        self.prop.onebody(self.walker, 1)
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAdjustMaxSize and i % m.ETAdjustStep == 0:
                self.prop.measure(self.walker)
                self.prop.adjust_ET()
            self.projectOneStep(idx)
        self.prepareBackPropagation()
        if nbp >= skip:
            for _ in range(nbp):
                self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                for k in range(skip):
                    if self.BPIndex == nbp:
                        self.addMeasurement()
                    self.projectOneStep(idx)
                if nbp < skip:
                    self.prepareBackPropagation()
            self.betas.append((m.thermalSize + (i + 0.5) * m.measureNumberPerWrite * skip - 0.5) * m.dt)
            w = self.bp.get()
            self.writes.append(w)
            self.bp.reset()
            if on_block:
                on_block(i, w)

    def energy(self):
        H = sum(w["HNum"] for w in self.writes)
        D = sum(w["den"] for w in self.writes)
        return H / D
