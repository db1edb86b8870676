"""Python handle on one sample in device memory (the C-ABI's rf_ctx).

Mirrors the reference's test drivers: create the context from a sample,
build rows and bonds (the t = 0 part of h_CalculationUniaxialCompression,
md_calculationuniaxialcompression.cu:155-212), select a platen test and
step or run it.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib

STATE_FIELDS = [("R", 3, np.float64, False), ("V", 3, np.float64, False),
                ("W", 3, np.float64, False), ("F", 3, np.float64, False),
                ("M", 3, np.float64, False), ("nbr", 1, np.int32, True),
                ("type", 1, np.uint8, True), ("inv_L0", 1, np.float64, True),
                ("prev_n", 3, np.float64, True), ("o_t", 3, np.float64, True),
                ("m_n", 3, np.float64, True), ("m_t", 3, np.float64, True),
                ("sigma_c", 1, np.float64, True), ("tau_c", 1, np.float64, True),
                ("bidx", 1, np.int8, True)]

_CT = {np.float64: C.c_double, np.int32: C.c_int32, np.uint8: C.c_uint8, np.int8: C.c_int8}


def alloc_state(n, slots):
    st = {}
    for name, comps, dt, per_slot in STATE_FIELDS:
        st[name] = np.zeros(n * slots * comps if per_slot else 3 * n, dtype=dt)
    return st


def state_view(st):
    v = abi.StateView()
    for name, _, dt, _ in STATE_FIELDS:
        a = st.get(name)
        if a is not None:
            setattr(v, name, a.ctypes.data_as(C.POINTER(_CT[dt])))
    return v


def derive(mat, r_min, r_max):
    d = abi.Derived()
    check(lib().rf_derive(C.byref(mat), float(r_min), float(r_max), C.byref(d)), "rf_derive")
    return d


def box_reference(space, cell):
    b = abi.Box()
    sp = (C.c_double * 3)(*[float(v) for v in space])
    check(lib().rf_box_reference(sp, float(cell), C.byref(b)), "rf_box_reference")
    return b


def _curve_ptr(curve):
    c = np.ascontiguousarray(np.asarray(curve, dtype=np.float64).reshape(-1, 3))
    return c, c.ctypes.data_as(C.POINTER(C.c_double))


def linearity_deviation(curve, peak_strain, peak_stress):
    """CalculateLinearityDeviation of a curve (n, 3) in the rf_get_curve layout."""
    c, ptr = _curve_ptr(curve)
    out = C.c_double()
    check(lib().rf_linearity_deviation(ptr, len(c), float(peak_strain), float(peak_stress),
                                       C.byref(out)), "rf_linearity_deviation")
    return out.value


def curve_moduli(curve, peak_strain, peak_stress, lo=0.1, hi=0.5):
    """Secant modulus and least-squares modulus over [lo, hi] x peak stress of
    the ascending branch (rf_curve_moduli)."""
    c, ptr = _curve_ptr(curve)
    out = abi.Moduli()
    check(lib().rf_curve_moduli(ptr, len(c), float(peak_strain), float(peak_stress), float(lo),
                                float(hi), C.byref(out)), "rf_curve_moduli")
    return out


class Context:
    """One sample on the GPU."""

    def __init__(self, xyz, r, mat, box, slots=32, cand_cap=32, skin=-1.0, device=0,
                 stream=None, order=abi.RF_ORDER_SPATIAL):
        """stream: a hipStream_t (as an int handle, e.g. torch.cuda.Stream's
        cuda_stream) the context runs on instead of its own.  order: how the
        context stores the particles (RF_ORDER_SPATIAL or RF_ORDER_GIVEN);
        every call speaks in the caller's ids either way."""
        self.n = int(len(r))
        self.slots = int(slots)
        self._x = np.ascontiguousarray(xyz[:, 0], dtype=np.float64)
        self._y = np.ascontiguousarray(xyz[:, 1], dtype=np.float64)
        self._z = np.ascontiguousarray(xyz[:, 2], dtype=np.float64)
        self._r = np.ascontiguousarray(r, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        s = abi.Sample(self.n, self._x.ctypes.data_as(dp), self._y.ctypes.data_as(dp),
                       self._z.ctypes.data_as(dp), self._r.ctypes.data_as(dp))
        opt = abi.Options(slots, cand_cap, skin, device, 256, stream, order, 0)
        h = C.c_void_p()
        self.h = None
        check(lib().rf_create(C.byref(s), C.byref(mat), C.byref(box), C.byref(opt), C.byref(h)),
              "rf_create")
        self.h = h
        self.mat = mat
        self.box = box

    def close(self):
        if self.h:
            lib().rf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_bonds(self):
        check(lib().rf_build_bonds(self.h), "rf_build_bonds")

    def set_test(self, cfg):
        check(lib().rf_set_test(self.h, C.byref(cfg)), "rf_set_test")

    def step(self, nsteps):
        done = C.c_int64()
        check(lib().rf_step(self.h, int(nsteps), C.byref(done)), "rf_step")
        return done.value

    def run(self):
        t = abi.TestState()
        check(lib().rf_run(self.h, C.byref(t)), "rf_run")
        return t

    def run_uniaxial(self, cfg):
        """Whole uniaxial test (rf_run_uniaxial): an abi.Result."""
        out = abi.Result()
        check(lib().rf_run_uniaxial(self.h, C.byref(cfg), C.byref(out)), "rf_run_uniaxial")
        return out

    def run_brazilian(self, cfg):
        """Whole Brazilian test (rf_run_brazilian): an abi.Result."""
        out = abi.Result()
        check(lib().rf_run_brazilian(self.h, C.byref(cfg), C.byref(out)), "rf_run_brazilian")
        return out

    def adjoint_run(self, params, steps):
        """Loading window of `steps` steps with tangents along the rf_param
        directions (rf_adjoint_run): returns (values (2,), grad (2, nparam))
        for the mean axial stress and the secant modulus, reduced units."""
        cfg = abi.AdjointCfg()
        cfg.nparam = len(params)
        for k, q in enumerate(params):
            cfg.params[k] = int(q)
        cfg.steps = int(steps)
        vals = np.zeros(2)
        grad = np.zeros(2 * len(params))
        check(lib().rf_adjoint_run(self.h, C.byref(cfg), vals.ctypes.data_as(C.POINTER(C.c_double)),
                                   grad.ctypes.data_as(C.POINTER(C.c_double))), "rf_adjoint_run")
        self._sens_np = len(params)
        return vals, grad.reshape(2, len(params))

    def test_state(self):
        t = abi.TestState()
        check(lib().rf_get_test_state(self.h, C.byref(t)), "rf_get_test_state")
        return t

    def curve(self):
        n = C.c_int64()
        check(lib().rf_get_curve(self.h, None, 0, C.byref(n)), "rf_get_curve")
        cap = 1 << 24
        out = np.zeros((min(cap, max(1, self.test_state().n_curve)), 3))
        check(lib().rf_get_curve(self.h, out.ctypes.data_as(C.c_void_p), len(out), C.byref(n)),
              "rf_get_curve")
        return out[:n.value].copy()

    def bond_counts(self):
        m = max(1, self.test_state().n_curve)
        out = np.zeros(m, dtype=np.int64)
        n = C.c_int64()
        check(lib().rf_get_bond_counts(self.h, out.ctypes.data_as(C.c_void_p), m, C.byref(n)),
              "rf_get_bond_counts")
        return out[:n.value].copy()

    def counters(self):
        c = abi.Counters()
        check(lib().rf_get_counters(self.h, C.byref(c)), "rf_get_counters")
        return c

    def get_state(self):
        st = alloc_state(self.n, self.slots)
        v = state_view(st)
        check(lib().rf_get_state(self.h, C.byref(v)), "rf_get_state")
        return st

    def set_state(self, st):
        v = state_view(st)
        check(lib().rf_set_state(self.h, C.byref(v)), "rf_set_state")

    def set_timing(self, every=1):
        """Bracket every `every`-th step-kernel launch with HIP events (0: off)."""
        check(lib().rf_set_timing(self.h, int(every)), "rf_set_timing")

    def sens_begin(self, params):
        """Start a sensitivity window along the given rf_param directions."""
        arr = (C.c_int32 * len(params))(*[int(q) for q in params])
        check(lib().rf_sens_begin(self.h, len(params), arr), "rf_sens_begin")
        self._sens_np = len(params)

    def sens_branches(self, mode, steps=0, recorded=None):
        """Frozen-branch windows (rf_sens_branches): abi.RF_BRANCH_RECORD the
        next windows' decisions for up to `steps` steps, RF_BRANCH_REPLAY
        those another context recorded, RF_BRANCH_DECIDE."""
        check(lib().rf_sens_branches(self.h, int(mode), int(steps),
                                     recorded.h if recorded is not None else None),
              "rf_sens_branches")
        self._branch_src = recorded  # keep the recording alive

    def sens_step(self, nsteps):
        done = C.c_int64()
        check(lib().rf_sens_step(self.h, int(nsteps), C.byref(done)), "rf_sens_step")
        return done.value

    def sens_curve(self):
        """(points, nparam, 2): d(bottom, top platen stress)/d(param) at the
        curve points recorded since sens_begin (reduced stress units)."""
        cap = 1 << 20
        buf = np.zeros(cap * self._sens_np * 2)
        n = C.c_int64()
        check(lib().rf_sens_get_curve(self.h, buf.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                      C.byref(n)), "rf_sens_get_curve")
        return buf[: n.value * self._sens_np * 2].reshape(n.value, self._sens_np, 2)

    def set_graph(self, on=True):
        """Replay full rf_step chunks as captured HIP graphs (default on)."""
        check(lib().rf_set_graph(self.h, 1 if on else 0), "rf_set_graph")

    def set_chunk(self, steps):
        check(lib().rf_set_chunk(self.h, int(steps)), "rf_set_chunk")

    def capture_chunk(self, steps):
        """Set the chunk length and capture its graph now (rf_capture_chunk);
        no step runs."""
        check(lib().rf_capture_chunk(self.h, int(steps)), "rf_capture_chunk")

    def last_timing(self):
        """(summed ms of the bracketed step kernels, bracketed launches,
        wall ms of the last step() call, steps it executed)"""
        k, n, t, s = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        check(lib().rf_last_timing(self.h, C.byref(k), C.byref(n), C.byref(t), C.byref(s)),
              "rf_last_timing")
        return k.value, n.value, t.value, s.value


def material_tangent(mat, param, wall_modulus_factor=0.0):
    """Derivatives of (E, G, hm_E, hm_G, mu, mu_r, p_A, E_wall, G_wall,
    ratio, bond radius) with respect to one rf_param."""
    out = (C.c_double * 11)()
    check(lib().rf_sens_material_tangent(C.byref(mat), float(wall_modulus_factor), int(param),
                                          out), "rf_sens_material_tangent")
    return np.array(out[:])


class Batch:
    """Several contexts advanced together (rf_batch_*): one step launch and
    one controller launch per step for all samples; each sample's results
    are the same as when it runs alone."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        check(lib().rf_batch_create(arr, len(self.contexts), C.byref(h)), "rf_batch_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().rf_batch_destroy(self.h)
            self.h = None

    def step(self, nsteps):
        done = (C.c_int64 * len(self.contexts))()
        check(lib().rf_batch_step(self.h, int(nsteps), done), "rf_batch_step")
        return list(done)

    def run(self):
        out = (abi.TestState * len(self.contexts))()
        check(lib().rf_batch_run(self.h, out), "rf_batch_run")
        return list(out)

    def last_ms(self):
        ms = C.c_double()
        check(lib().rf_batch_last_ms(self.h, C.byref(ms)), "rf_batch_last_ms")
        return ms.value

    def set_form(self, form):
        """Fix the launch form (0-4, as in tuning()) or return to the
        autotuner (-1)."""
        check(lib().rf_batch_set_form(self.h, int(form)), "rf_batch_set_form")

    def tuning(self):
        """(choice or -1, [ms per step of each way]): 0-2 one shared launch
        with the blocks by sample, interleaved, largest sample first; 3 each
        sample alone, one after another; 4 each sample on its own stream,
        all launched before any is waited for."""
        order = C.c_int32()
        ms = (C.c_double * 5)()
        check(lib().rf_batch_tuning(self.h, C.byref(order), ms), "rf_batch_tuning")
        return order.value, list(ms)
