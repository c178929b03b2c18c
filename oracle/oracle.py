"""TEST INFRASTRUCTURE ONLY: ctypes wrappers of the CPU checker (oracle/liboracle.so) and of
the reference's own CPU simulator (oracle/_ref/libspice_ref.so, built from the reference's
sources by `make -C oracle ref`). The product never imports this module."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
# SPICE_ORACLE_LIB: another build of the checker (the mutation tests load deliberately broken
# builds made from edited copies of the model headers)
ORACLE_LIB = os.environ.get("SPICE_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libspice_ref.so")
REF_IEEE_LIB = os.path.join(_HERE, "_ref", "libspice_ref_ieee.so")

_c = ctypes
_P = _c.c_void_p


class Connection(_c.Structure):
    _fields_ = [("src_pop", _c.c_uint64), ("dst_pop", _c.c_uint64), ("p", _c.c_float)]


class Edge(_c.Structure):
    _fields_ = [("src_first", _c.c_int32), ("src_last", _c.c_int32),
                ("dst_first", _c.c_int32), ("dst_last", _c.c_int32), ("p", _c.c_float)]


MODEL_IDS = {"vogels_abbott": 0, "brunel": 1, "brunel_with_plasticity": 2, "synth": 3,
             "brunel_3pop": 4, "brunel_3pop_stdp": 5, "brunel_3pop_stdp_deferred": 6}
NEURON_ATTRS = {
    "vogels_abbott": [("V", np.float32), ("Gex", np.float32), ("Gin", np.float32), ("Twait", np.int32)],
    "brunel": [("V", np.float32), ("Twait", np.int32)],
    "brunel_with_plasticity": [("V", np.float32), ("Twait", np.int32)],
    "synth": [("N", np.int32)],
    "brunel_3pop": [("V", np.float32), ("Twait", np.int32)],
    "brunel_3pop_stdp": [("V", np.float32), ("Twait", np.int32)],
    "brunel_3pop_stdp_deferred": [("V", np.float32), ("Twait", np.int32)],
}
SYNAPSE_ATTRS = {
    "brunel_with_plasticity": [("W", np.float32), ("Zpre", np.float32), ("Zpost", np.float32)],
    "brunel_3pop_stdp": [("W", np.float32)],
    "brunel_3pop_stdp_deferred": [("W", np.float32)],
}

_lib = None
_refs = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        _lib = _c.CDLL(ORACLE_LIB)
        _lib.oracle_last_error.restype = _c.c_char_p
    return _lib


def ref_available(ieee=False):
    return os.path.exists(REF_IEEE_LIB if ieee else REF_LIB)


def ref(ieee=False):
    """The reference's CPU build: its own flags (-ffast-math) or the IEEE variant."""
    path = REF_IEEE_LIB if ieee else REF_LIB
    if path not in _refs:
        L = _c.CDLL(path)
        L.ref_last_error.restype = _c.c_char_p
        L.ref_num_neurons.restype = _c.c_uint64
        L.ref_width.restype = _c.c_uint64
        _refs[path] = L
    return _refs[path]


def _check(st, L=None):
    if st != 0:
        L = L or lib()
        msg = (L.oracle_last_error() if hasattr(L, "oracle_last_error") else L.ref_last_error()).decode()
        if st == 1:
            raise ValueError(msg)
        if st == 2:
            raise MemoryError(msg)
        raise RuntimeError(msg)


def _layout_args(pops, conns):
    pa = np.asarray(pops, dtype=np.uint64)
    ca = (Connection * max(1, len(conns)))(*[Connection(s, d, p) for s, d, p in conns])
    return pa, ca


def _normalize_layout(layout):
    """Accepts (n, p) or (pops, conns) or an object with .pops/.conns."""
    if hasattr(layout, "pops"):
        return layout.pops, layout.conns
    a, b = layout
    if np.isscalar(a):
        return [int(a)], [(0, 0, float(b))]
    return list(a), list(b)


def _ptr(a, t):
    return a.ctypes.data_as(_c.POINTER(t))


# ---- primitives ----
def philox(ctr, key):
    c = (_c.c_uint32 * 4)(*ctr)
    k = (_c.c_uint32 * 2)(*key)
    o = (_c.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def _vec(fn, x, out_dtype=np.float64, in_dtype=np.float64):
    x = np.ascontiguousarray(x, dtype=in_dtype)
    y = np.empty(x.shape, dtype=out_dtype)
    getattr(lib(), fn)(_c.c_void_p(x.ctypes.data), _c.c_void_p(y.ctypes.data), _c.c_int64(x.size))
    return y


def det_log(x):
    return _vec("oracle_det_log", x)


def det_sqrt(x):
    return _vec("oracle_det_sqrt", x)


def det_cos2pi(x):
    return _vec("oracle_det_cos2pi", x)


def exp_fix(x):
    return _vec("oracle_exp_fix", x, np.uint32, np.uint32)


def kahan(dt, nsteps):
    d = np.empty(nsteps, np.float32)
    s = np.empty(nsteps, np.float32)
    lib().oracle_kahan(_c.c_float(dt), _c.c_int64(nsteps), _c.c_void_p(d.ctypes.data), _c.c_void_p(s.ctypes.data))
    return d, s


def layout_flatten(pops, conns):
    pa, ca = _layout_args(pops, conns)
    edges = (Edge * max(1, len(conns)))()
    ne, n, md = _c.c_int32(), _c.c_uint64(), _c.c_uint64()
    _check(lib().oracle_layout_flatten(_ptr(pa, _c.c_uint64), len(pops), ca, len(conns), edges,
                                       _c.byref(ne), _c.byref(n), _c.byref(md)))
    return [(e.src_first, e.src_last, e.dst_first, e.dst_last, e.p) for e in edges[: ne.value]], n.value, md.value


def generate_csr(layout, seed):
    """Host build of the generator: (row_ptr int64[N+1], col int32[E], width)."""
    pops, conns = _normalize_layout(layout)
    pa, ca = _layout_args(pops, conns)
    n, w, nnz = _c.c_int64(), _c.c_int64(), _c.c_int64()
    args = (_ptr(pa, _c.c_uint64), len(pops), ca, len(conns), _c.c_uint64(seed), _c.byref(n), _c.byref(w), _c.byref(nnz))
    _check(lib().oracle_generate_csr(*args, None, None))
    rp = np.empty(n.value + 1, np.int64)
    col = np.empty(max(nnz.value, 1), np.int32)
    _check(lib().oracle_generate_csr(*args, _c.c_void_p(rp.ctypes.data), _c.c_void_p(col.ctypes.data)))
    return rp, col[: nnz.value], w.value


def generate_rows(layout, seed, row_first, row_last, cols=True):
    """Host build of rows [row_first, row_last) only: (absolute offsets int64[k + 1], targets
    int32 of those rows, or None with cols=False)."""
    pops, conns = _normalize_layout(layout)
    pa, ca = _layout_args(pops, conns)
    k = int(row_last) - int(row_first)
    rp = np.empty(max(k, 0) + 1, np.int64)
    args = (_ptr(pa, _c.c_uint64), len(pops), ca, len(conns), _c.c_uint64(seed), _c.c_int64(row_first),
            _c.c_int64(row_last), _c.c_void_p(rp.ctypes.data))
    _check(lib().oracle_generate_rows(*args, None, _c.c_int64(0)))
    if not cols:
        return rp, None
    col = np.empty(max(int(rp[-1] - rp[0]), 1), np.int32)
    _check(lib().oracle_generate_rows(*args, _c.c_void_p(col.ctypes.data), _c.c_int64(col.size)))
    return rp, col[: int(rp[-1] - rp[0])]


def degrees(layout, seed):
    """Per (row, flattened connection) clipped degrees; -1 where the connection does not cover the row."""
    pops, conns = _normalize_layout(layout)
    pa, ca = _layout_args(pops, conns)
    n = int(sum(pops))
    out = np.empty((n, len(conns)), np.int32)
    _check(lib().oracle_degrees(_ptr(pa, _c.c_uint64), len(pops), ca, len(conns), _c.c_uint64(seed),
                                _c.c_void_p(out.ctypes.data)))
    return out


FLAG_FIXED_POINT = 1
FLAG_REF_RNG = 1 << 16


def _flags(fixed_point, ref_rng):
    return (FLAG_FIXED_POINT if fixed_point else 0) | (FLAG_REF_RNG if ref_rng else 0)


class CpuSnn:
    """The CPU checker: cpu::snn semantics (dense plasticity, sequential fp32 receive).

    Random numbers come from Philox keyed on (seed, neuron, step) like the GPU path, or, with
    ref_rng=True, from the reference CPU backend's sequential xoshiro128+ stream seeded with
    `seed` (to reproduce the reference build's trajectories)."""

    def __init__(self, model, layout=None, dt=1e-4, delay=1, seed=1337, fixed_point=False, shift=40, _handle=None,
                 ref_rng=False):
        self.model = model
        self._mid = MODEL_IDS[model]
        self._dt, self._delay = float(dt), int(delay)
        self._h = _c.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            pops, conns = _normalize_layout(layout)
            pa, ca = _layout_args(pops, conns)
            _check(lib().oracle_create(self._mid, _ptr(pa, _c.c_uint64), len(pops), ca, len(conns), _c.c_float(dt),
                                       int(delay), _c.c_uint64(seed), _c.c_uint32(_flags(fixed_point, ref_rng)),
                                       int(shift), _c.byref(self._h)))

    @classmethod
    def from_csr(cls, model, row_ptr, col, width, dt, delay=1, seed=1337, fixed_point=False, shift=40, ref_rng=False):
        rp = np.ascontiguousarray(row_ptr, np.int64)
        cl = np.ascontiguousarray(col, np.int32)
        if cl.size == 0:
            cl = np.zeros(1, np.int32)
        h = _c.c_void_p()
        _check(lib().oracle_create_csr(MODEL_IDS[model], _c.c_void_p(rp.ctypes.data), _c.c_void_p(cl.ctypes.data),
                                       _c.c_int64(rp.size - 1), _c.c_int64(width), _c.c_float(dt), int(delay),
                                       _c.c_uint64(seed), _c.c_uint32(_flags(fixed_point, ref_rng)), int(shift),
                                       _c.byref(h)))
        return cls(model, None, dt, delay, seed, fixed_point, shift, _handle=h)

    @classmethod
    def from_ell(cls, model, ell, width, dt, delay=1, seed=1337, fixed_point=False, shift=40, ref_rng=False):
        e = np.ascontiguousarray(ell, np.int32)
        h = _c.c_void_p()
        _check(lib().oracle_create_ell(MODEL_IDS[model], _c.c_void_p(e.ctypes.data), _c.c_int64(e.size // width),
                                       _c.c_int64(width), _c.c_float(dt), int(delay), _c.c_uint64(seed),
                                       _c.c_uint32(_flags(fixed_point, ref_rng)), int(shift), _c.byref(h)))
        return cls(model, None, dt, delay, seed, fixed_point, shift, _handle=h)

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = _c.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _info(self):
        n, w, nnz, na, ns = _c.c_int64(), _c.c_int64(), _c.c_int64(), _c.c_int32(), _c.c_int32()
        lib().oracle_info(self._h, _c.byref(n), _c.byref(w), _c.byref(nnz), _c.byref(na), _c.byref(ns))
        return n.value, w.value, nnz.value

    def num_neurons(self):
        return self._info()[0]

    def width(self):
        return self._info()[1]

    def num_edges(self):
        return self._info()[2]

    def num_synapses(self):
        n, w, _ = self._info()
        return n * w

    def dt(self):
        return self._dt

    def delay(self):
        return self._delay

    def step(self, spikes=False):
        if not spikes:
            _check(lib().oracle_step(self._h, None, _c.c_uint64(0), None))
            return None
        n = self.num_neurons()
        buf = np.empty(max(n, 1), np.int32)
        cnt = _c.c_uint64()
        _check(lib().oracle_step(self._h, _c.c_void_p(buf.ctypes.data), _c.c_uint64(n), _c.byref(cnt)))
        return buf[: cnt.value].copy()

    def run(self, nsteps):
        _check(lib().oracle_run(self._h, _c.c_int64(nsteps)))

    def run_record(self, nsteps, cap=None):
        """Runs nsteps steps in one call (the GIL is released meanwhile); returns (offsets,
        ids): the spikes of step k (counted from the current step) are ids[offsets[k]:offsets[k+1]]."""
        if cap is None:
            cap = self.num_neurons() * nsteps // 16 + 4096
        ids = np.empty(max(cap, 1), np.int32)
        off = np.empty(nsteps + 1, np.uint64)
        _check(lib().oracle_run_record(self._h, _c.c_int64(nsteps), _c.c_void_p(ids.ctypes.data), _c.c_uint64(cap),
                                       _c.c_void_p(off.ctypes.data)))
        return off, ids[: int(off[-1])].copy()

    def csr(self):
        n, _, nnz = self._info()
        rp = np.empty(n + 1, np.int64)
        col = np.empty(max(nnz, 1), np.int32)
        lib().oracle_export_csr(self._h, _c.c_void_p(rp.ctypes.data), _c.c_void_p(col.ctypes.data))
        return rp, col[:nnz]

    def adj(self):
        rp, col = self.csr()
        n, w, _ = self._info()
        ell = np.full(n * w, np.iinfo(np.int32).max, np.int32)
        for r in range(n):
            ell[r * w: r * w + rp[r + 1] - rp[r]] = col[rp[r]: rp[r + 1]]
        return ell, w

    def neurons(self):
        n = self.num_neurons()
        out = {}
        for a, (name, dt) in enumerate(NEURON_ATTRS[self.model]):
            arr = np.empty(n, dt)
            _check(lib().oracle_export_neuron_attr(self._h, a, _c.c_void_p(arr.ctypes.data)))
            out[name] = arr
        return out

    def set_neurons(self, state):
        for a, (name, dt) in enumerate(NEURON_ATTRS[self.model]):
            arr = np.ascontiguousarray(state[name], dt)
            _check(lib().oracle_import_neuron_attr(self._h, a, _c.c_void_p(arr.ctypes.data)))

    def synapses(self, ell=True):
        _, w, nnz = self._info()
        out = {}
        for a, (name, dt) in enumerate(SYNAPSE_ATTRS.get(self.model, [])):
            arr = np.empty(max(nnz, 1), dt)
            _check(lib().oracle_export_synapse_attr(self._h, a, _c.c_void_p(arr.ctypes.data)))
            arr = arr[:nnz]
            if ell:
                rp, _ = self.csr()
                n = rp.size - 1
                e = np.zeros(n * w, dt)
                for r in range(n):
                    e[r * w: r * w + rp[r + 1] - rp[r]] = arr[rp[r]: rp[r + 1]]
                arr = e
            out[name] = arr
        return out

    def set_synapses(self, state, ell=False):
        if ell:
            raise NotImplementedError("CSR order only")
        for a, (name, dt) in enumerate(SYNAPSE_ATTRS.get(self.model, [])):
            arr = np.ascontiguousarray(state[name], dt)
            _check(lib().oracle_import_synapse_attr(self._h, a, _c.c_void_p(arr.ctypes.data)))

    def set_params(self, params):
        """Per-network model parameters (snn_info::param), set before the first step."""
        p = list(params or [])[:4]
        p += [0.0] * (4 - len(p))
        lib().oracle_set_params(self._h, (_c.c_float * 4)(*p))

    def traces(self):
        """Neuron-local (pre, post) traces of trace-synapse models."""
        n = self.num_neurons()
        x, y = np.empty(n, np.float32), np.empty(n, np.float32)
        if lib().oracle_export_traces(self._h, _c.c_void_p(x.ctypes.data), _c.c_void_p(y.ctypes.data)):
            raise ValueError("model has no neuron-local traces")
        return x, y

    def clock(self):
        i, s, c = _c.c_int64(), _c.c_float(), _c.c_float()
        lib().oracle_get_clock(self._h, _c.byref(i), _c.byref(s), _c.byref(c))
        return i.value, s.value, c.value

    def set_clock(self, step, sum_, comp):
        lib().oracle_set_clock(self._h, _c.c_int64(step), _c.c_float(sum_), _c.c_float(comp))

    def set_in_flight(self, steps):
        """steps: spike-id lists of the last len(steps) (<= delay) steps, oldest first: the
        spikes the next steps deliver (replaces the net's own record of them)."""
        counts = np.array([len(x) for x in steps], np.int64)
        ids = np.concatenate([np.asarray(x, np.int32) for x in steps] + [np.zeros(1, np.int32)])
        _check(lib().oracle_set_in_flight(self._h, _c.c_int64(len(steps)), _c.c_void_p(counts.ctypes.data),
                                          _c.c_void_p(ids.ctypes.data)))


class RefSnn:
    """The reference's own cpu::snn (oracle/_ref), for the four reference models.

    The reference seeds the random stream of each constructed net from a process-wide counter
    starting at 1337 (spice/cpu/snn.h.cpp:16, 76); `self.seed` is the value this net got."""

    _created = {}

    def __init__(self, model, layout, dt, delay=1, ieee=False):
        assert model in ("vogels_abbott", "brunel", "brunel_with_plasticity", "synth")
        self.model = model
        pops, conns = _normalize_layout(layout)
        pa = np.asarray(pops, np.uint64)
        cs = np.asarray([c[0] for c in conns] or [0], np.uint64)
        cd = np.asarray([c[1] for c in conns] or [0], np.uint64)
        cp = np.asarray([c[2] for c in conns] or [0], np.float32)
        self._h = _c.c_void_p()
        self._L = L = ref(ieee)
        st = L.ref_create(MODEL_IDS[model], _c.c_void_p(pa.ctypes.data), len(pops), _c.c_void_p(cs.ctypes.data),
                              _c.c_void_p(cd.ctypes.data), _c.c_void_p(cp.ctypes.data), len(conns), _c.c_float(dt),
                              int(delay), _c.byref(self._h))
        if st:
            raise ValueError(L.ref_last_error().decode())
        k = RefSnn._created.get(ieee, 0)
        self.seed = 1337 + k
        RefSnn._created[ieee] = k + 1
        self._dt, self._delay = float(dt), int(delay)

    def __del__(self):
        try:
            if self._h:
                self._L.ref_destroy(self._h)
        except Exception:
            pass

    def num_neurons(self):
        return self._L.ref_num_neurons(self._h)

    def width(self):
        return self._L.ref_width(self._h)

    def step(self, spikes=True):
        n = self.num_neurons()
        buf = np.empty(max(n, 1), np.int32)
        cnt = _c.c_uint64()
        self._L.ref_step(self._h, _c.c_void_p(buf.ctypes.data), _c.c_uint64(n), _c.byref(cnt))
        return buf[: cnt.value].copy()

    def adj(self):
        n, w = self.num_neurons(), self.width()
        e = np.empty(n * w, np.int32)
        self._L.ref_export_adj(self._h, _c.c_void_p(e.ctypes.data))
        return e, w

    def neurons(self):
        n = self.num_neurons()
        out = {}
        for a, (name, dt) in enumerate(NEURON_ATTRS[self.model]):
            arr = np.empty(n, dt)
            self._L.ref_export_neurons(self._h, a, _c.c_void_p(arr.ctypes.data))
            out[name] = arr
        return out

    def synapses(self):
        n, w = self.num_neurons(), self.width()
        out = {}
        for a, (name, dt) in enumerate(SYNAPSE_ATTRS.get(self.model, [])):
            arr = np.empty(n * w, dt)
            self._L.ref_export_synapses(self._h, a, _c.c_void_p(arr.ctypes.data))
            out[name] = arr
        return out


def ref_layout(pops, conns):
    pa = np.asarray(pops, np.uint64)
    cs = np.asarray([c[0] for c in conns] or [0], np.uint64)
    cd = np.asarray([c[1] for c in conns] or [0], np.uint64)
    cp = np.asarray([c[2] for c in conns] or [0], np.float32)
    out = np.empty((max(1, len(conns)), 4), np.int32)
    op = np.empty(max(1, len(conns)), np.float32)
    n, md = _c.c_uint64(), _c.c_uint64()
    st = ref().ref_layout(_c.c_void_p(pa.ctypes.data), len(pops), _c.c_void_p(cs.ctypes.data),
                          _c.c_void_p(cd.ctypes.data), _c.c_void_p(cp.ctypes.data), len(conns),
                          _c.c_void_p(out.ctypes.data), _c.c_void_p(op.ctypes.data), _c.byref(n), _c.byref(md))
    if st:
        raise ValueError(ref().ref_last_error().decode())
    return [tuple(int(x) for x in out[i]) + (float(op[i]),) for i in range(len(conns))], n.value, md.value


def ref_generate(pops, conns):
    _, n, w = ref_layout(pops, conns)
    pa = np.asarray(pops, np.uint64)
    cs = np.asarray([c[0] for c in conns] or [0], np.uint64)
    cd = np.asarray([c[1] for c in conns] or [0], np.uint64)
    cp = np.asarray([c[2] for c in conns] or [0], np.float32)
    out = np.empty(max(n * w, 1), np.int32)
    st = ref().ref_generate(_c.c_void_p(pa.ctypes.data), len(pops), _c.c_void_p(cs.ctypes.data),
                            _c.c_void_p(cd.ctypes.data), _c.c_void_p(cp.ctypes.data), len(conns),
                            _c.c_void_p(out.ctypes.data))
    if st:
        raise ValueError(ref().ref_last_error().decode())
    return out[: n * w], w
