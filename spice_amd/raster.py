"""Spike rasters: per-step spike-id streams turned into the reference's text raster and into
per-population statistics (firing rate, ISI coefficient of variation).

Text format (samples/_main.cpp:31-54 of the reference): the first line holds the neuron
count, then one line per step with that step's spike ids sorted ascending, comma-separated
(an empty line for a silent step).

Statistics are accumulated incrementally (a chunk of steps at a time, e.g. from a server's
spike sink), so a raster of a long run never has to be held in memory.
"""
import io

import numpy as np


class Raster:
    """Accumulates the spikes of one network. pops: population sizes (default: one
    population of n neurons); keep_lines: also keep the text raster."""

    def __init__(self, n, pops=None, keep_lines=False):
        self.n = int(n)
        self.pops = list(pops) if pops is not None else [self.n]
        if sum(self.pops) != self.n:
            raise ValueError("population sizes must add up to the neuron count")
        self.steps = 0
        self.count = np.zeros(self.n, np.int64)
        self._last = np.full(self.n, -1, np.int64)
        self._isi_n = np.zeros(self.n, np.int64)
        self._isi_s = np.zeros(self.n, np.float64)
        self._isi_q = np.zeros(self.n, np.float64)
        self._lines = [] if keep_lines else None

    def add_step(self, step, ids):
        """Spikes of one step (any order)."""
        ids = np.asarray(ids, dtype=np.int64)
        self.add(step, np.array([0, ids.size], np.uint64), ids)

    def add(self, first_step, offsets, ids):
        """Spikes of steps first_step, first_step + 1, ...: ids[offsets[k]:offsets[k+1]] of step
        first_step + k. Steps must arrive in order."""
        offsets = np.asarray(offsets, dtype=np.int64)
        ids = np.asarray(ids, dtype=np.int64)
        nsteps = offsets.size - 1
        if first_step != self.steps:
            raise ValueError(f"steps must be added in order (expected step {self.steps}, got {first_step})")
        if ids.size and (ids.min() < 0 or ids.max() >= self.n):
            raise ValueError("spike id out of range")
        if self._lines is not None:
            for k in range(nsteps):
                self._lines.append(",".join(map(str, np.sort(ids[offsets[k]:offsets[k + 1]]))))
        self.steps += nsteps
        if ids.size == 0:
            return
        t = first_step + np.repeat(np.arange(nsteps, dtype=np.int64), np.diff(offsets))
        order = np.lexsort((t, ids))
        i, t = ids[order], t[order]
        np.add.at(self.count, i, 1)
        # intervals inside the chunk, plus each neuron's first spike here vs its last before
        same = i[1:] == i[:-1]
        prev_t = np.concatenate(([-1], t[:-1]))
        first = np.concatenate(([True], ~same))
        prev_t[first] = self._last[i[first]]
        ok = prev_t >= 0
        isi = (t - prev_t)[ok].astype(np.float64)
        np.add.at(self._isi_n, i[ok], 1)
        np.add.at(self._isi_s, i[ok], isi)
        np.add.at(self._isi_q, i[ok], isi * isi)
        last = np.concatenate((~same, [True]))
        self._last[i[last]] = t[last]

    def _pop_slices(self):
        b = np.cumsum([0] + self.pops)
        return [slice(b[k], b[k + 1]) for k in range(len(self.pops))]

    def rates(self, dt):
        """Mean firing rate (Hz) of every population over the recorded steps."""
        T = self.steps * dt
        return [float(self.count[s].sum() / (max(1, s.stop - s.start) * T)) if T > 0 else 0.0
                for s in self._pop_slices()]

    def isi_cv(self, min_intervals=2):
        """Mean over the population's neurons (with >= min_intervals intervals) of
        std(ISI) / mean(ISI); NaN for a population without such neurons."""
        out = []
        for s in self._pop_slices():
            n = self._isi_n[s]
            m = n >= min_intervals
            if not m.any():
                out.append(float("nan"))
                continue
            mean = self._isi_s[s][m] / n[m]
            var = np.maximum(self._isi_q[s][m] / n[m] - mean * mean, 0.0)
            out.append(float(np.mean(np.sqrt(var) / mean)))
        return out

    def text(self):
        """The kept raster in the reference's text format."""
        if self._lines is None:
            raise ValueError("raster lines were not kept (keep_lines=False)")
        buf = io.StringIO()
        buf.write(f"{self.n}\n")
        for line in self._lines:
            buf.write(line + "\n")
        return buf.getvalue()

    def write(self, path):
        with open(path, "w") as f:
            f.write(self.text())


def read_raster(text):
    """Parses the text format: returns (n, list of per-step int32 id arrays)."""
    lines = text.split("\n")
    n = int(lines[0])
    body = lines[1:]
    if body and body[-1] == "":
        body = body[:-1]
    return n, [np.array([int(x) for x in ln.split(",")], np.int32) if ln else np.zeros(0, np.int32) for ln in body]
