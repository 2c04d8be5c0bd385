"""ctypes wrapper of the serving loop (orqg_server_* in include/orqg.h):
secret-shared sort / GROUP BY SUM requests with continuous batching."""
import ctypes

import numpy as np

from . import OrqgError, lib

REQ_SORT, REQ_GROUPBY_SUM = 0, 1


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("op", ctypes.c_int), ("width", ctypes.c_int),
                ("n", ctypes.c_size_t), ("key_shares", ctypes.c_void_p),
                ("val_shares", ctypes.c_void_p), ("keys", ctypes.c_uint8 * 320),
                ("out_key", ctypes.c_void_p), ("out_val", ctypes.c_void_p),
                ("out_valid", ctypes.c_void_p)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("slot", ctypes.c_int),
                ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double), ("msg", ctypes.c_char * 192)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("batch_size", ctypes.c_long), ("batch_rows", ctypes.c_size_t),
                ("batch_window_us", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("max_running", ctypes.c_uint64),
                ("peak_reserved", ctypes.c_uint64), ("peak_live_bytes", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("mallocs", ctypes.c_uint64),
                ("pool_trims", ctypes.c_uint64), ("sort_batches", ctypes.c_uint64),
                ("sorts_batched", ctypes.c_uint64)]


class Job:
    """One request with its host buffers ([3][n] share arrays in and out)."""

    def __init__(self, rid, op, key_shares, val_shares, keys):
        key_shares = np.ascontiguousarray(key_shares)
        val_shares = np.ascontiguousarray(val_shares)
        n = key_shares.shape[1]
        self.id, self.op, self.n = rid, op, n
        self.key_in, self.val_in = key_shares, val_shares
        self.out_key = np.zeros_like(key_shares)
        self.out_val = np.zeros_like(val_shares)
        self.out_valid = np.zeros_like(key_shares)
        r = Request()
        r.id, r.op, r.width, r.n = rid, op, key_shares.dtype.itemsize, n
        r.key_shares = key_shares.ctypes.data
        r.val_shares = val_shares.ctypes.data
        ctypes.memmove(r.keys, bytes(keys), 320)
        r.out_key = self.out_key.ctypes.data
        r.out_val = self.out_val.ctypes.data
        r.out_valid = self.out_valid.ctypes.data
        self.req = r
        self.completion = None


class Server:
    """batch_rows > 0: radix sorts of at most that many rows from concurrent
    requests run as one batched sort (gathered for up to batch_window_us)."""

    def __init__(self, slots=4, hbm_budget=0, device=0, batch_size=-12, batch_rows=0,
                 batch_window_us=2000):
        cfg = ServerConfig(device, slots, hbm_budget, batch_size, batch_rows, batch_window_us)
        self.h = ctypes.c_void_p()
        rc = lib().orqg_server_create(ctypes.byref(cfg), ctypes.byref(self.h))
        if rc:
            raise OrqgError(f"orqg_server_create failed ({rc})")
        self.jobs = {}

    def now_ms(self):
        return lib().orqg_server_now_ms(self.h)

    def submit(self, job):
        self.jobs[job.id] = job
        rc = lib().orqg_server_submit(self.h, ctypes.byref(job.req))
        if rc:
            raise OrqgError(f"orqg_server_submit failed ({rc})")

    def poll(self, wait_ms=0, max_n=64):
        out = (Completion * max_n)()
        cnt = ctypes.c_int()
        lib().orqg_server_poll(self.h, out, max_n, ctypes.byref(cnt), wait_ms)
        res = []
        for i in range(cnt.value):
            c = out[i]
            job = self.jobs.pop(c.id)
            job.completion = (c.status, c.slot, c.t_submit_ms, c.t_start_ms, c.t_done_ms)
            job.error = c.msg.decode(errors="replace") if c.status else ""
            res.append(job)
        return res

    def pending(self):
        q, r = ctypes.c_int(), ctypes.c_int()
        lib().orqg_server_pending(self.h, ctypes.byref(q), ctypes.byref(r))
        return q.value, r.value

    def stats(self):
        st = ServerStats()
        lib().orqg_server_get_stats(self.h, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def drain(self):
        """Wait until every submitted request has completed; returns them."""
        done = []
        while self.jobs:
            done += self.poll(wait_ms=100)
        return done

    def close(self):
        if self.h:
            lib().orqg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
