"""Python mirror of the reference's network API on top of the gfx950 C ABI.

`Layout` mirrors util::layout (spice/util/layout.h:24-57) and `Snn` mirrors
spice::snn<Model> / cuda::snn<Model> (spice/snn.h:16-45, spice/cuda/snn.h:16-77): same
constructors (layout, ELL adjacency, conversion from another network), step() with optional
spike readback, num_neurons(), num_synapses() (= N x ELL width), dt(), delay(), adj(),
neurons(), synapses(). Error behaviour follows the reference's exception types
(ValueError <- std::invalid_argument, MemoryError <- std::bad_alloc, ...).
"""
import ctypes

import numpy as np

from . import _lib
from .models import MODELS, model_id, neuron_attrs, synapse_attrs

_c = ctypes


def _u64p(a):
    return a.ctypes.data_as(_c.POINTER(_c.c_uint64))


class Layout:
    """Population sizes plus (src pop, dst pop, p) connections.

    Layout(n, p) is one population of n neurons connected to itself with probability p;
    Layout([n0, n1, ...], [(src, dst, p), ...]) is the general form.
    """

    def __init__(self, pops, connections=None):
        if connections is None and np.isscalar(pops):
            raise ValueError("Layout(n, p) needs a connection probability")
        if np.isscalar(pops):
            n = int(pops)
            if n <= 0:
                raise ValueError("layout must contain at least 1 neuron")
            pops, connections = [n], [(0, 0, float(connections))]
        self.pops = [int(x) for x in pops]
        self.conns = [(int(s), int(d), float(p)) for (s, d, p) in connections]
        if any(x < 0 for x in self.pops):
            raise ValueError("invalid population size")
        if any(s < 0 or d < 0 for s, d, _ in self.conns):
            raise ValueError("invalid index in connections matrix")
        L = _lib.lib()
        pops_a = np.asarray(self.pops, dtype=np.uint64)
        conns_a = (_lib.Connection * max(1, len(self.conns)))(*[_lib.Connection(s, d, p) for s, d, p in self.conns])
        edges = (_lib.Edge * max(1, len(self.conns)))()
        ne, n, md = _c.c_int32(), _c.c_uint64(), _c.c_uint64()
        _lib.check(L.spice_layout_flatten(_u64p(pops_a), len(self.pops), conns_a, len(self.conns), edges,
                                          _c.byref(ne), _c.byref(n), _c.byref(md)))
        self._edges = [(e.src_first, e.src_last, e.dst_first, e.dst_last, e.p) for e in edges[: ne.value]]
        self._n = n.value
        self._max_degree = md.value

    def size(self):
        return self._n

    def connections(self):
        return list(self._edges)

    def max_degree(self):
        return self._max_degree

    def _c_args(self):
        pops_a = np.asarray(self.pops, dtype=np.uint64)
        conns_a = (_lib.Connection * max(1, len(self.conns)))(*[_lib.Connection(s, d, p) for s, d, p in self.conns])
        return pops_a, conns_a


def generate_graph(layout, seed, d_row_ptr, d_col=None, col_cap=0, noalloc=False):
    """Random graph of `layout` (the generator of Snn, same seed -> same graph) written to
    caller-owned device memory: d_row_ptr / d_col are device addresses (ints) of int64[N + 1]
    and int32[col_cap] buffers (d_col None: offsets only). Returns (nnz, ELL width); raises
    ValueError if d_col is given and holds fewer than nnz entries. noalloc: the entry point
    that allocates no device memory (its scratch is the library's fixed 1 MiB region per GPU:
    at most 1024 flattened connections and 130,023,424 neurons; overlapping calls on one GPU
    race)."""
    if not isinstance(layout, Layout):
        layout = Layout(*layout) if isinstance(layout, tuple) else Layout(layout)
    pops_a, conns_a = layout._c_args()
    nnz, width = _c.c_uint64(), _c.c_uint64()
    fn = _lib.lib().spice_hip_generate_graph_noalloc if noalloc else _lib.lib().spice_hip_generate_graph
    _lib.check(fn(_u64p(pops_a), len(layout.pops), conns_a, len(layout.conns), int(seed), _c.c_void_p(int(d_row_ptr)),
                  _c.c_void_p(int(d_col)) if d_col else None, int(col_cap), _c.byref(nnz), _c.byref(width)))
    return nnz.value, width.value


def probe_random_lines(nbytes, accesses, rmw, reps=20):
    """Microseconds per launch of `accesses` 4-byte accesses at pseudo-random places of a fresh
    device array of `nbytes` bytes (a load each; with rmw, the store of the changed value after
    it), averaged over `reps` launches: the device's rate of random memory lines
    (spice_hip_probe_random_lines)."""
    us = _c.c_double()
    _lib.check(_lib.lib().spice_hip_probe_random_lines(int(nbytes), int(accesses), 1 if rmw else 0, int(reps), _c.byref(us)))
    return us.value


def _desc(layout, delay, seed, params, keep):
    """NetDesc for a layout; `keep` collects the ctypes buffers the descriptor points to."""
    if not isinstance(layout, Layout):
        layout = Layout(*layout) if isinstance(layout, tuple) else Layout(layout)
    pops_a, conns_a = layout._c_args()
    keep.extend([pops_a, conns_a])
    p = list(params or [])[:4]
    p += [0.0] * (4 - len(p))
    return _lib.NetDesc(_u64p(pops_a), len(layout.pops), conns_a, len(layout.conns), int(delay), int(seed),
                        (_c.c_float * 4)(*p))


class Snn:
    """One network simulated on the GPU (cuda::snn<Model> counterpart).

    params: up to 4 per-network model parameters (snn_info::param; e.g. brunel_3pop's Poisson
    rate in Hz), 0 = the model's default."""

    def __init__(self, model, layout, dt, delay=1, seed=1337, fixed_point=False, _handle=None, flags=0, params=None,
                 _borrowed=False):
        self.model = model
        self._mid = model_id(model)
        self._h = _c.c_void_p()
        self._borrowed = _borrowed
        L = _lib.lib()
        if _handle is not None:
            self._h = _handle
        else:
            keep = []
            d = _desc(layout, delay, seed, params, keep)
            _lib.check(L.spice_hip_create_desc(self._mid, _c.byref(d), float(dt), self._flags(fixed_point) | int(flags),
                                               _c.byref(self._h)))
        self._nattrs = neuron_attrs(model)
        self._sattrs = synapse_attrs(model)

    @staticmethod
    def _flags(fixed_point):
        return _lib.SPICE_FLAG_FIXED_POINT if fixed_point else 0

    @classmethod
    def from_ell(cls, model, edges, width, dt, delay=1, seed=1337, fixed_point=False):
        """cuda::snn(adj, width, dt, delay): ELL rows padded with INT_MAX."""
        e = np.ascontiguousarray(edges, dtype=np.int32)
        if width <= 0 or e.size % width:
            raise ValueError("adjacency size must be a multiple of width")
        h = _c.c_void_p()
        _lib.check(_lib.lib().spice_hip_create_from_ell(model_id(model), e.ctypes.data_as(_c.POINTER(_c.c_int32)),
                                                        e.size // width, width, float(dt), int(delay), int(seed),
                                                        cls._flags(fixed_point), _c.byref(h)))
        return cls(model, None, dt, delay, seed, fixed_point, _handle=h)

    @classmethod
    def from_csr(cls, model, row_ptr, col, width, dt, delay=1, seed=1337, fixed_point=False):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        h = _c.c_void_p()
        _lib.check(_lib.lib().spice_hip_create_from_csr(model_id(model), rp.ctypes.data_as(_c.POINTER(_c.c_int64)),
                                                        cl.ctypes.data_as(_c.POINTER(_c.c_int32)), rp.size - 1,
                                                        int(width), float(dt), int(delay), int(seed),
                                                        cls._flags(fixed_point), _c.byref(h)))
        return cls(model, None, dt, delay, seed, fixed_point, _handle=h)

    @classmethod
    def from_net(cls, net, seed=1337, fixed_point=False):
        """Conversion constructor (cuda::snn(spice::snn const&), spice/cuda/snn.d.cpp:123-143):
        copies graph, neuron and synapse state and the clock of any network exposing the same
        accessors (another Snn or the CPU checker)."""
        rp, col = net.csr()
        out = cls.from_csr(net.model, rp, col, net.width(), net.dt(), net.delay(), seed, fixed_point)
        out.set_neurons(net.neurons())
        if out._sattrs:
            out.set_synapses(net.synapses(ell=False), ell=False)
        out.set_clock(*net.clock())
        return out

    # ---- lifetime ----
    def close(self):
        if self._h and not self._borrowed:
            _lib.check(_lib.lib().spice_hip_destroy(self._h))
        self._h = _c.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- simulation ----
    def step(self, spikes=False):
        """Advances one step; returns this step's spike ids (unspecified order) if asked."""
        L = _lib.lib()
        _lib.check(L.spice_hip_step(self._h, None, None))
        if not spikes:
            return None
        n = self.num_neurons()
        buf = np.empty(n, dtype=np.int32)
        cnt = _c.c_uint64()
        _lib.check(L.spice_hip_copy_spikes(self._h, buf.ctypes.data_as(_c.POINTER(_c.c_int32)), n, _c.byref(cnt)))
        return buf[: cnt.value].copy()

    def step_spikes_of_last(self):
        """Spike ids of the most recent step (without stepping)."""
        n = self.num_neurons()
        buf = np.empty(n, dtype=np.int32)
        cnt = _c.c_uint64()
        _lib.check(_lib.lib().spice_hip_copy_spikes(self._h, buf.ctypes.data_as(_c.POINTER(_c.c_int32)), n,
                                                    _c.byref(cnt)))
        return buf[: cnt.value].copy()

    def step_device(self):
        """step(int** d_spikes, uint** d_count): returns device pointers (valid until the next step)."""
        ds, dc = _c.c_void_p(), _c.c_void_p()
        _lib.check(_lib.lib().spice_hip_step(self._h, _c.byref(ds), _c.byref(dc)))
        return ds.value, dc.value

    def run(self, nsteps):
        _lib.check(_lib.lib().spice_hip_run(self._h, int(nsteps)))

    def run_profiled(self, nsteps):
        """Runs nsteps eagerly, each kernel launched with HIP events attached to its dispatch
        (its own duration, without the launch gap); returns summed device ms per kernel:
        {"receive": ..., "update": ..., "step": ...}. Two-kernel steps: "receive" is the
        delivery (with the previous step's plasticity work of plastic models), "update" the
        neuron update, "step" 0. Overlapped steps: "receive" is the first delivery, "update"
        the last update (with the passes of the step before it), "step" the nsteps - 1
        overlapped kernels (update of step T, delivery of T + 1, plasticity work of T - 1).
        Fused steps (static, delay-1 nets at shapes with one delivery workgroup per tile):
        "receive" and "update" 0, "step" the nsteps kernels that each deliver and update."""
        ms = (_c.c_double * 3)(0.0, 0.0, 0.0)
        _lib.check(_lib.lib().spice_hip_run_profiled(self._h, int(nsteps), ms))
        return {"receive": ms[0], "update": ms[1], "step": ms[2]}

    def synchronize(self):
        _lib.check(_lib.lib().spice_hip_synchronize(self._h))

    # ---- queries ----
    def _q(self, fn, ctype):
        v = ctype()
        _lib.check(getattr(_lib.lib(), fn)(self._h, _c.byref(v)))
        return v.value

    def num_neurons(self):
        return self._q("spice_hip_num_neurons", _c.c_uint64)

    def num_synapses(self):
        return self._q("spice_hip_num_synapses", _c.c_uint64)

    def width(self):
        return self._q("spice_hip_width", _c.c_uint64)

    def num_edges(self):
        return self._q("spice_hip_num_edges", _c.c_uint64)

    def dt(self):
        return self._q("spice_hip_dt", _c.c_float)

    def delay(self):
        return self._q("spice_hip_delay", _c.c_int32)

    def info(self):
        return {"num_neurons": self.num_neurons()}

    def clock(self):
        i, s, c = _c.c_int64(), _c.c_float(), _c.c_float()
        _lib.check(_lib.lib().spice_hip_get_clock(self._h, _c.byref(i), _c.byref(s), _c.byref(c)))
        return i.value, s.value, c.value

    def set_clock(self, step, sum_, comp):
        _lib.check(_lib.lib().spice_hip_set_clock(self._h, int(step), float(sum_), float(comp)))

    # ---- launch shapes ----
    @staticmethod
    def _launch_dict(c):
        return {"tile_w": c.tile_w, "flush_slots": c.flush_slots, "flush_grid": c.flush_grid, "recv_block": c.recv_block,
                "recv_parts": c.recv_parts}

    def launch(self):
        """Current launch shapes: tile_w (targets per delivery tile), flush_slots (target slots
        per plasticity-flush work item), flush_grid (flush workgroups), recv_block (threads per
        delivery workgroup), recv_parts (delivery workgroups per tile)."""
        c = _lib.Launch()
        _lib.check(_lib.lib().spice_hip_get_launch(self._h, _c.byref(c)))
        return self._launch_dict(c)

    def set_launch(self, tile_w=0, flush_slots=0, flush_grid=0, recv_block=0, recv_parts=0):
        """Sets launch shapes (0 = default); results do not depend on them (in fp32 mode the
        summation order of delivered inputs may change, as it does between runs)."""
        c = _lib.Launch(int(tile_w), int(flush_slots), int(flush_grid), int(recv_block), int(recv_parts))
        _lib.check(_lib.lib().spice_hip_set_launch(self._h, _c.byref(c)))

    def autotune(self, warmup=500, steps=100):
        """Times candidate launch shapes on a copy of this network and applies the fastest;
        returns (launch dict, step time in us). The network itself does not advance."""
        c = _lib.Launch()
        us = _c.c_double()
        _lib.check(_lib.lib().spice_hip_autotune(self._h, int(warmup), int(steps), _c.byref(c), _c.byref(us)))
        return self._launch_dict(c), us.value

    def stats(self):
        s = _lib.Stats()
        _lib.check(_lib.lib().spice_hip_get_stats(self._h, _c.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.Stats._fields_}

    def reset_stats(self):
        _lib.check(_lib.lib().spice_hip_reset_stats(self._h))

    # ---- export / import ----
    def adj(self):
        """(ELL edges int32[N * width], width), INT_MAX-padded."""
        n, w = self.num_neurons(), self.width()
        out = np.empty(n * w, dtype=np.int32)
        _lib.check(_lib.lib().spice_hip_export_adj(self._h, out.ctypes.data_as(_c.POINTER(_c.c_int32)), out.size))
        return out, w

    def csr(self):
        n, e = self.num_neurons(), self.num_edges()
        rp = np.empty(n + 1, dtype=np.int64)
        col = np.empty(max(e, 1), dtype=np.int32)
        _lib.check(_lib.lib().spice_hip_export_csr(self._h, rp.ctypes.data_as(_c.POINTER(_c.c_int64)),
                                                   col.ctypes.data_as(_c.POINTER(_c.c_int32))))
        return rp, col[:e]

    def csr_rows(self, row_first, row_last, cols=True):
        """Rows [row_first, row_last) of the CSR adjacency: (absolute offsets int64[k + 1],
        target ids int32 of those rows, or None with cols=False)."""
        L = _lib.lib()
        k = int(row_last) - int(row_first)
        rp = np.empty(max(k, 0) + 1, dtype=np.int64)
        pi64 = rp.ctypes.data_as(_c.POINTER(_c.c_int64))
        _lib.check(L.spice_hip_export_csr_rows(self._h, int(row_first), int(row_last), pi64, None, 0))
        if not cols:
            return rp, None
        col = np.empty(max(int(rp[-1] - rp[0]), 1), dtype=np.int32)
        _lib.check(L.spice_hip_export_csr_rows(self._h, int(row_first), int(row_last), pi64,
                                               col.ctypes.data_as(_c.POINTER(_c.c_int32)), col.size))
        return rp, col[: int(rp[-1] - rp[0])]

    def neurons(self):
        """Neuron state as {attribute name: array[N]}."""
        n = self.num_neurons()
        out = {}
        for a, (name, dt) in enumerate(self._nattrs):
            arr = np.empty(n, dtype=dt)
            _lib.check(_lib.lib().spice_hip_export_neuron_attr(self._h, a, arr.ctypes.data, arr.nbytes))
            out[name] = arr
        return out

    def set_neurons(self, state):
        for a, (name, dt) in enumerate(self._nattrs):
            arr = np.ascontiguousarray(state[name], dtype=dt)
            _lib.check(_lib.lib().spice_hip_import_neuron_attr(self._h, a, arr.ctypes.data, arr.nbytes))

    def synapses(self, ell=True):
        """Synapse state as {attribute name: array}, ELL slot order (N x width) or CSR order."""
        size = self.num_synapses() if ell else self.num_edges()
        out = {}
        for a, (name, dt) in enumerate(self._sattrs):
            arr = np.empty(size, dtype=dt)
            _lib.check(_lib.lib().spice_hip_export_synapse_attr(self._h, a, arr.ctypes.data, arr.nbytes, int(ell)))
            out[name] = arr
        return out

    def set_synapses(self, state, ell=True):
        for a, (name, dt) in enumerate(self._sattrs):
            arr = np.ascontiguousarray(state[name], dtype=dt)
            _lib.check(_lib.lib().spice_hip_import_synapse_attr(self._h, a, arr.ctypes.data, arr.nbytes, int(ell)))


class Batch:
    """Many independent networks of one static model stepped together (one launch per step
    phase for all of them). nets: list of (layout, delay, seed[, params]) tuples. net(i) is a
    borrowed Snn view of network i (queries, exports and spike readback; no stepping)."""

    def __init__(self, model, nets, dt, fixed_point=False, flags=0):
        self.model = model
        keep = []
        descs = (_lib.NetDesc * len(nets))(*[_desc(n[0], n[1], n[2], n[3] if len(n) > 3 else None, keep) for n in nets])
        self._h = _c.c_void_p()
        _lib.check(_lib.lib().spice_batch_create(model_id(model), len(nets), descs, float(dt),
                                                 Snn._flags(fixed_point) | int(flags), _c.byref(self._h)))
        self._dt = dt
        self._nets = []
        for i in range(len(nets)):
            h = _c.c_void_p()
            _lib.check(_lib.lib().spice_batch_net(self._h, i, _c.byref(h)))
            self._nets.append(Snn(model, None, dt, _handle=h, _borrowed=True))

    def __len__(self):
        return len(self._nets)

    def net(self, i):
        return self._nets[i]

    def run(self, nsteps):
        _lib.check(_lib.lib().spice_batch_run(self._h, int(nsteps)))

    def synchronize(self):
        _lib.check(_lib.lib().spice_batch_synchronize(self._h))

    def close(self):
        if self._h:
            for n in self._nets:
                n._h = _c.c_void_p()
            _lib.check(_lib.lib().spice_batch_destroy(self._h))
            self._h = _c.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
