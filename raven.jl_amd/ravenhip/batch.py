"""Batches of independent solves over one persistent launch per stage
(rv_batch_* in include/ravenhip.h).

``Batch(dgs)`` groups DGSEM operators (same dim; affine bricks, Rusanov flux);
``stage`` runs one fused LSRK stage of every member in one launch and
``lsrk_run`` replays whole LSRK steps as captured hipGraphs.  Raven advances
one solver per ``solve!`` call (odesolvers.jl:9-30); every member's result here
is the one its own operator gives.
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
        """nsteps whole LSRK steps of every member with its own dt (dq zeroed
        first); q is updated in place."""
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
