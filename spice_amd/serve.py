"""Serving loop: simulation jobs admitted against a device-memory budget as they arrive,
kept resident on the GPU, stepped together in batched chunks, spikes streamed back.

Python mirror of the spice_server_* entry points of include/spice_hip.h (no counterpart in
the reference, whose unit of work is one cuda::snn stepped by its owner). The loop itself
(admission, chunking, batching, spike streaming) runs natively on a worker thread of the
library; this module only submits, waits and receives spikes.
"""
import ctypes as _c
import threading

import numpy as np

from . import _lib
from .models import model_id
from .snn import Layout, Snn, _desc

STATUS = {_lib.JOB_QUEUED: "queued", _lib.JOB_RUNNING: "running", _lib.JOB_DONE: "done",
          _lib.JOB_FAILED: "failed", _lib.JOB_CANCELLED: "cancelled"}


def mix64(x):
    """splitmix64 finalizer on a uint64 array (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def spike_digest(steps):
    """Digest of a spike stream given as an iterable of (step, ids): the wrapping sum over
    its spikes of mix64(step << 32 | id) -- the server's spice_job_result.digest."""
    d = np.uint64(0)
    with np.errstate(over="ignore"):
        for t, ids in steps:
            ids = np.asarray(ids, dtype=np.int64)
            if ids.size:
                keys = (np.uint64(t) << np.uint64(32)) | (ids.astype(np.uint64) & np.uint64(0xFFFFFFFF))
                d = np.uint64(d + np.sum(mix64(keys), dtype=np.uint64))
    return int(d)


class Server:
    """A spice_server. budget_bytes: admission budget of device memory; chunk_steps: the
    longest chunk of steps between spike deliveries. on_spikes(job, first_step, offsets, ids)
    is called on the library's worker thread after every chunk with numpy copies (offsets has
    nsteps + 1 entries); keep_spikes=True also keeps every job's stream for spikes(job).
    spike_jobs: None (every job), or a set of job ids (the caller may add to it, e.g. the id
    the next submit will return, as ids count up from 1) whose spikes reach on_spikes and
    keep_spikes; the other jobs' spikes are still streamed to host memory and digested by the
    library (spice_job_result.spikes / .digest) but not copied into Python. The callback runs
    on the server's worker thread, between chunks: work done in it delays the next chunk."""

    def __init__(self, budget_bytes=250e9, chunk_steps=1000, on_spikes=None, keep_spikes=False, spike_jobs=None):
        self._on_spikes = on_spikes
        self._keep = keep_spikes
        self._spike_jobs = spike_jobs
        self._streams = {}
        self._lock = threading.Lock()
        self._chunk = int(chunk_steps)
        self._sink = _lib.SPIKE_SINK(self._deliver)  # referenced for the server's lifetime
        self._h = _c.c_void_p()
        _lib.check(_lib.lib().spice_server_create(int(budget_bytes), self._chunk, self._sink, None, _c.byref(self._h)))

    def _deliver(self, user, job, first_step, nsteps, offsets, ids):
        if self._on_spikes is None and not self._keep:
            return
        if self._spike_jobs is not None and int(job) not in self._spike_jobs:
            return
        off = np.ctypeslib.as_array(offsets, shape=(nsteps + 1,)).copy()
        n = int(off[-1])
        spk = np.ctypeslib.as_array(ids, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        if self._keep:
            with self._lock:
                self._streams.setdefault(int(job), []).append((int(first_step), off, spk))
        if self._on_spikes is not None:
            self._on_spikes(int(job), int(first_step), off, spk)

    @staticmethod
    def _job_desc(model, layout, dt, steps, delay, seed, params, fixed_point, flags, keep):
        if not isinstance(layout, Layout):
            layout = Layout(*layout) if isinstance(layout, tuple) else Layout(layout)
        nd = _desc(layout, delay, seed, params, keep)
        return _lib.JobDesc(model_id(model), nd, float(dt), Snn._flags(fixed_point) | int(flags), int(steps))

    def submit(self, model, layout, dt, steps, delay=1, seed=1337, params=None, fixed_point=False, flags=0):
        """Queues a job of `steps` steps; returns its id. ValueError if it can never fit."""
        keep = []
        d = self._job_desc(model, layout, dt, steps, delay, seed, params, fixed_point, flags, keep)
        out = _c.c_uint64()
        _lib.check(_lib.lib().spice_server_submit(self._h, _c.byref(d), _c.byref(out)))
        return int(out.value)

    def estimate(self, model, layout, dt, steps, delay=1, seed=1337, params=None, fixed_point=False, flags=0):
        """Device-memory reservation (bytes) the server would make for the job."""
        keep = []
        d = self._job_desc(model, layout, dt, steps, delay, seed, params, fixed_point, flags, keep)
        out = _c.c_uint64()
        _lib.check(_lib.lib().spice_server_estimate(_c.byref(d), self._chunk, _c.byref(out)))
        return int(out.value)

    def wait(self, job, timeout=None):
        """Waits for the job (timeout in seconds, None = no limit); returns its record as a
        dict (status 'done' / 'failed' / 'cancelled', or 'queued' / 'running' on timeout)."""
        r = _lib.JobResult()
        fin = _c.c_int32()
        _lib.check(_lib.lib().spice_server_wait(self._h, int(job), -1.0 if timeout is None else float(timeout),
                                                _c.byref(r), _c.byref(fin)))
        rec = {k: getattr(r, k) for k, _ in _lib.JobResult._fields_ if k != "reserved0"}
        rec["status"] = STATUS[r.status]
        rec["finished"] = bool(fin.value)
        rec["latency_s"] = r.t_done - r.t_submit if fin.value else None
        if r.status == _lib.JOB_FAILED:
            buf = _c.create_string_buffer(1024)
            _lib.check(_lib.lib().spice_server_job_error(self._h, int(job), buf, 1024))
            rec["error"] = buf.value.decode(errors="replace")
        return rec

    def cancel(self, job):
        _lib.check(_lib.lib().spice_server_cancel(self._h, int(job)))

    def drain(self):
        """Waits until no job is queued or resident."""
        _lib.check(_lib.lib().spice_server_drain(self._h))

    def info(self):
        i = _lib.ServerInfo()
        _lib.check(_lib.lib().spice_server_get_info(self._h, _c.byref(i)))
        return {k: getattr(i, k) for k, _ in _lib.ServerInfo._fields_}

    def spikes(self, job):
        """The kept spike stream of a job: list over its steps of int32 id arrays (the order
        within a step is the device's compaction order)."""
        with self._lock:
            chunks = list(self._streams.get(int(job), []))
        out = []
        for _, off, ids in sorted(chunks, key=lambda c: c[0]):
            for k in range(len(off) - 1):
                out.append(ids[off[k]:off[k + 1]])
        return out

    def close(self):
        if self._h:
            _lib.check(_lib.lib().spice_server_destroy(self._h))
            self._h = _c.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
