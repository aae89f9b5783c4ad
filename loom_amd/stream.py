"""The device-resident stream (loom_stream_* in include/loom_amd.h): thread
slots whose registers stay on the GPU between steps, finished threads'
outputs exchanged for new inputs on the device.

Python mirror of the reference's stream protocol (src/vm/simd.hpp:2797-2903;
driven by src/tests/stream_wrapper.hpp:33-80): `set_inputs`, `step` (the
reference's `s()`), `sort`, `exchange`, `done`, `drain` (its
`get_output_zipped`).  Outputs come with the sequence number of their
thread (inputs are numbered in the order given), so `run_all` can return
them in input order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import RunResult, abi, init, kernel, make_params


class Stream:
    def __init__(self, k, slots: int, **params):
        init()
        self.kernel = kernel(k)
        params.setdefault("live_threshold", 32)
        self.params = make_params(mode=abi.MODE_HYBRID, **params)
        self.slots = (max(int(slots), 1) + 63) // 64 * 64
        h = C.c_void_p()
        abi.check(abi.load().loom_stream_create(self.kernel.index, self.slots,
                                                C.byref(self.params), C.byref(h)),
                  "loom_stream_create")
        self.handle = h

    def _in(self, inputs, m):
        cols = [np.ascontiguousarray(a[:m], dtype=dt) for a, dt in zip(inputs, self.kernel.inputs)]
        return cols, (C.c_void_p * abi.MAX_IO)(*[c.ctypes.data for c in cols])

    def _out(self, m):
        cols = [np.zeros(m, dt) for dt in self.kernel.outputs]
        ids = np.zeros(m, np.uint64)
        return cols, ids, (C.c_void_p * abi.MAX_IO)(*[c.ctypes.data for c in cols])

    def set_inputs(self, inputs) -> int:
        """Start the first rows of `inputs` (columns); returns how many."""
        n = len(inputs[0]) if inputs else 0
        cols, ptrs = self._in(inputs, min(n, self.slots))
        taken = C.c_uint64()
        abi.check(abi.load().loom_stream_set_inputs(self.handle, ptrs, len(cols[0]) if cols else 0,
                                                    C.byref(taken)), "loom_stream_set_inputs")
        return taken.value

    def step(self) -> RunResult:
        st = abi.RunStats()
        abi.check(abi.load().loom_stream_step(self.handle, C.byref(st)), "loom_stream_step")
        return RunResult(outputs=None, ms_total=st.ms_total, ms_main=st.ms_main,
                         main_launches=st.main_launches, passes=st.passes, errors=st.errors,
                         grid_waves=st.grid_waves, counters=list(st.counters))

    def sort(self) -> bool:
        avail = C.c_int()
        abi.check(abi.load().loom_stream_sort(self.handle, C.byref(avail)), "loom_stream_sort")
        return bool(avail.value)

    def state(self):
        live, halted = C.c_uint64(), C.c_uint64()
        abi.check(abi.load().loom_stream_state(self.handle, C.byref(live), C.byref(halted)),
                  "loom_stream_state")
        return live.value, halted.value

    @property
    def done(self) -> bool:
        return self.state()[0] == 0

    def exchange(self, inputs):
        """Outputs (columns) and sequence numbers of finished threads, their
        slots given the next rows of `inputs`; as many as both allow."""
        _, halted = self.state()
        m = min(halted, len(inputs[0]) if inputs else 0)
        cols, ptrs = self._in(inputs, m)
        outs, ids, optrs = self._out(m)
        got = C.c_uint64()
        abi.check(abi.load().loom_stream_exchange(self.handle, ptrs, m, optrs, ids.ctypes.data,
                                                  C.byref(got)), "loom_stream_exchange")
        return [o[:got.value] for o in outs], ids[:got.value]

    def drain(self):
        """Outputs and sequence numbers of every finished thread still held."""
        _, halted = self.state()
        outs, ids, optrs = self._out(halted)
        got = C.c_uint64()
        abi.check(abi.load().loom_stream_drain(self.handle, optrs, ids.ctypes.data, halted,
                                               C.byref(got)), "loom_stream_drain")
        return [o[:got.value] for o in outs], ids[:got.value]

    def counters(self) -> dict:
        c = abi.StreamCounters()
        abi.check(abi.load().loom_stream_counters_get(self.handle, C.byref(c)),
                  "loom_stream_counters_get")
        return {f: getattr(c, f) for f, _ in abi.StreamCounters._fields_}

    def run_all(self, inputs):
        """The reference's stream driver (StreamWrapper): start, then
        exchange after every step until the input is consumed, step until
        done, drain; returns the outputs in input order."""
        n = len(inputs[0])
        result = [np.zeros(n, dt) for dt in self.kernel.outputs]
        seen = np.zeros(n, bool)

        def put(outs, ids):
            idx = ids.astype(np.int64)
            assert not seen[idx].any(), "a thread returned twice"
            seen[idx] = True
            for r, o in zip(result, outs):
                r[idx] = o

        i = self.set_inputs(inputs)
        self.step()
        self.sort()
        while i < n:
            outs, ids = self.exchange([a[i:] for a in inputs])
            i += len(ids)
            put(outs, ids)
            self.step()
            self.sort()
        while not self.done:
            self.step()
            self.sort()
        put(*self.drain())
        assert seen.all(), "a thread never returned"
        return result

    def close(self):
        if self.handle:
            abi.load().loom_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
