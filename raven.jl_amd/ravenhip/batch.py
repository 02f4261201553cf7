"""Batches of independent solves (rv_batch_* in include/ravenhip.h).

``Batch(dgs)`` groups DGSEM operators (same dim; affine bricks, Rusanov flux);
``stage`` runs one fused LSRK stage of every member (kinds large enough to fill
the device launch on their own, the small 2-D kinds share one persistent
launch), ``step`` one whole LSRK step, and ``lsrk_run`` replays whole LSRK steps
as captured hipGraphs.  ``autotune`` picks the launch layout by timing it.
Raven advances one solver per ``solve!`` call (odesolvers.jl:9-30); every
member's result here is the one its own operator gives.
"""
import ctypes

import numpy as np

from . import _lib as L


def _ptrs(tensors):
    arr = (ctypes.c_void_p * len(tensors))()
    for k, t in enumerate(tensors):
        arr[k] = t.data_ptr()
    return arr


class Batch:
    def __init__(self, dgs):
        self.dgs = list(dgs)
        hs = (L.H * len(self.dgs))(*[dg.handle for dg in self.dgs])
        b = L.H()
        L.check_batch(L.lib().rv_batch_create(hs, len(self.dgs), ctypes.byref(b)), None)
        self._b = b

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            L.lib().rv_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.dgs)

    @property
    def info(self):
        ni, nw, sb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        L.check_batch(L.lib().rv_batch_info(self._b, ctypes.byref(ni), ctypes.byref(nw),
                                            ctypes.byref(sb)), self._b)
        return {"items": ni.value, "waves": nw.value, "lds_bytes_per_wave": sb.value}

    def _check(self, states):
        if len(states) != len(self.dgs):
            raise ValueError("one state per member")
        return [dg._check_state(t) for dg, t in zip(self.dgs, states)]

    def stage(self, q_out, q_in, dq, a, b_dt):
        """One fused LSRK stage of every member (lists of per-member states and
        coefficients)."""
        self._check(q_out), self._check(q_in), self._check(dq)
        av = np.ascontiguousarray(a, dtype=np.float64)
        bv = np.ascontiguousarray(b_dt, dtype=np.float64)
        if av.size != len(self.dgs) or bv.size != len(self.dgs):
            raise ValueError("one coefficient per member")
        L.check_batch(L.lib().rv_batch_stage(self._b, _ptrs(q_out), _ptrs(q_in), _ptrs(dq),
                                             L.ptr(av, ctypes.c_double),
                                             L.ptr(bv, ctypes.c_double)), self._b)

    def lsrk_run(self, q, dq, dt, nsteps, scheme=L.RV_LSRK54):
        """nsteps whole LSRK steps of every member with its own dt (dq starts
        at zero as in rv_lsrk_run); q is updated in place."""
        self._check(q), self._check(dq)
        dv = np.ascontiguousarray(dt, dtype=np.float64)
        if dv.size != len(self.dgs):
            raise ValueError("one dt per member")
        L.check_batch(L.lib().rv_batch_lsrk_run(self._b, _ptrs(q), _ptrs(dq), int(scheme),
                                                L.ptr(dv, ctypes.c_double), int(nsteps)),
                      self._b)

    def step(self, q0, q1, dq, cur, dt, scheme=L.RV_LSRK54):
        """One whole LSRK step of every member with its own dt, ping-ponging
        between q0[m] and q1[m] (read from q1[m] when cur[m] is 1); members with
        dt[m] == 0 are left untouched."""
        n = len(self.dgs)
        dv = np.ascontiguousarray(dt, dtype=np.float64)
        cv = np.ascontiguousarray(cur, dtype=np.int32)
        if dv.size != n or cv.size != n:
            raise ValueError("one dt and one parity per member")
        on = [k for k in range(n) if dv[k] != 0.0]
        for k in on:
            for t in (q0[k], q1[k], dq[k]):
                self.dgs[k]._check_state(t)
        ptrs = [(ctypes.c_void_p * n)(*[t.data_ptr() if t is not None else None for t in ts])
                for ts in (q0, q1, dq)]
        L.check_batch(L.lib().rv_batch_step(self._b, *ptrs, L.ptr(cv, ctypes.c_int32),
                                            L.ptr(dv, ctypes.c_double), int(scheme)), self._b)

    def configure(self, mixed_waves_per_block=2, own_launch_items=1024, streams=None):
        """Launch layout: waves per block of the mixed launch, the item count
        from which a 2-D kind launches on its own (rv_batch_configure) and,
        when given, the number of concurrent streams the launch chains are
        spread over (rv_batch_set_streams, 1 to 4; the library starts at 2)."""
        L.check_batch(L.lib().rv_batch_configure(self._b, int(mixed_waves_per_block),
                                                 int(own_launch_items)), self._b)
        if streams is not None:
            L.check_batch(L.lib().rv_batch_set_streams(self._b, int(streams)), self._b)

    def autotune(self, q, dt, mixed_waves=(1, 2, 4), own_items=(512, 1024, 4096, 1 << 40),
                 streams=(1, 2, 4), member_waves=(1, 2), steps=4, scheme=L.RV_LSRK54,
                 repeats=3):
        """Pick the launch layout by timing what a run does: whole LSRK steps
        replayed as hipGraphs (lsrk_run), on scratch copies of the states q
        with step sizes dt (q is not modified).

        First each kind's waves per block (the single operator's launch shape,
        timed on the kind's largest member with DGSEM.autotune and applied to
        every member of the kind), then the mixed launch's waves per block and
        the own-launch threshold at streams[1] concurrent streams (streams[0]
        if there is one choice), then the stream count at the best of those:
        `repeats` rounds over the layouts of a pass in turn (so a drift of the
        clock touches them alike), each setting the layout, building its
        graphs with one run of `steps` steps and timing a second; the layout
        with the lowest median is kept.  Returns {(mixed_waves, own_items,
        streams): (median, min, max) ms per step}."""
        import torch
        by_kind = {}
        for k, dg in enumerate(self.dgs):
            key = (dg.info.dim, dg.info.nq, dg.law.law_id, dg.dtype)
            if key not in by_kind or dg.K > self.dgs[by_kind[key]].K:
                by_kind[key] = k
        for key, k in by_kind.items():
            self.dgs[k].autotune(q[k], choices=member_waves)
            w = self.dgs[k].waves_per_block
            for dg in self.dgs:
                if (dg.info.dim, dg.info.nq, dg.law.law_id, dg.dtype) == key:
                    dg.waves_per_block = w
        qs = [t.clone() for t in q]
        dqs = [torch.empty_like(t) for t in q]

        def timed(layouts):
            runs = {k: [] for k in layouts}
            for _ in range(repeats):
                for k in layouts:
                    # a new layout drops the captured graphs: one run captures
                    # them, the next is timed
                    self.configure(*k)
                    self.lsrk_run(qs, dqs, dt, steps, scheme)
                    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                    ev[0].record()
                    self.lsrk_run(qs, dqs, dt, steps, scheme)
                    ev[1].record()
                    ev[1].synchronize()
                    runs[k].append(ev[0].elapsed_time(ev[1]) / steps)
            return {k: (float(np.median(v)), min(v), max(v)) for k, v in runs.items()}

        s0 = streams[1] if len(streams) > 1 else streams[0]
        times = timed([(w, t, s0) for w in mixed_waves for t in own_items])
        w, t, _ = min(times, key=lambda k: times[k][0])
        times.update(timed([(w, t, s) for s in streams if s != s0]))
        best = min(times, key=lambda k: times[k][0])
        self.configure(*best)
        return times
