"""Serving: a device-resident reference panel answering query batches, with continuous batching.

A query batch (1-512 sequences) is appended to the panel set and scored against every earlier sequence -- the panel,
all earlier queries, and the earlier queries of its own batch -- by one TN93G_APPEND run over its index range.
Those are exactly the pairs a tn93 all-pairs run over everything received so far adds for the batch, each computed
with the earlier sequence first as tn93 does (rows before columns), so a batch's result does not depend on what
else is running.

Admission: a batch waits until a worker is free and the device memory of the runs in flight plus its own
(tn93g_run_bytes) fits in the HBM budget.  Each worker thread owns a context (its own stream and run buffers), so
admitted runs overlap on the device while later batches are appended and admitted.  ctypes releases the GIL during
the library calls.

Merging: batches that are waiting together when a worker frees up are admitted as one group (up to `merge_rows`
queries) and computed by ONE APPEND run over the group's whole range [n1 of the first, n2 of the last).  That run's
pairs are exactly the union of the batches' own pairs (pair (i, j) with i < j belongs to the batch holding j), so a
batch's links are the run's links whose column lies in its range, bit for bit the same as alone.

Results are certified: each run's link distances must print with %g as tn93 prints them at both ends of their
enclosures (tn93g_result_certify), or the run is repeated with FLAG_EXACT.
"""
import collections
import threading
import time

import numpy as np

from . import APPEND, FLAG_EXACT, RESOLVE, Engine
from . import gen


class Request:
    def __init__(self, rid, codes, arrival):
        self.rid, self.codes, self.arrival = rid, codes, arrival
        self.n1 = self.n2 = None  # index range in the panel set after the append
        self.bytes = 0
        self.started = self.done = None
        self.result = None  # {"link_row", "link_col", "link_dist"}: the batch's links in tn93's order
        self.group = 1      # batches computed by the same APPEND run (this one included)
        self.error = None
        self.event = threading.Event()

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class Server:
    """Panel set on device 0; `workers` contexts run the admitted batches concurrently."""

    def __init__(self, panel, capacity, workers=4, hbm_budget=32 << 30, run_kw=None, merge_rows=0):
        panel = np.ascontiguousarray(panel, dtype=np.uint8)
        self.merge_rows = merge_rows  # 0: every batch is its own run
        self.L = panel.shape[1]
        self.owner = Engine(0)
        self.set = self.owner.create(self.L, capacity)
        self.set.append(panel)
        self.run_kw = dict(mode=RESOLVE, threshold=0.015, min_overlap=500, resolve_fraction=0.05)
        if run_kw:
            self.run_kw.update(run_kw)
        self.budget = hbm_budget
        self.in_flight = 0
        self.lock = threading.Condition()
        self.pending = collections.deque()
        self.free = [Engine(0) for _ in range(workers)]
        self.stopping = False
        self.threads = []
        self.admitter = threading.Thread(target=self._admit, daemon=True)
        self.admitter.start()

    # -- arrival: append in arrival order (this defines "earlier"), then queue for admission
    def submit(self, rid, codes, arrival=None):
        req = Request(rid, np.ascontiguousarray(codes, dtype=np.uint8), time.perf_counter() if arrival is None
                      else arrival)
        with self.lock:
            req.n1 = self.set.append(req.codes)
            req.n2 = req.n1 + req.codes.shape[0]
            req.bytes = self.run_bytes(req.n1, req.n2)
            self.pending.append(req)
            self.lock.notify_all()
        return req

    def run_bytes(self, n1, n2):
        """Device memory an APPEND run over [n1, n2) may work in: the fast path's, or the exact path's when
        certification sends it there (the larger of the two, tn93g_run_bytes)."""
        kw = dict(self.run_kw, kind=APPEND, n1=n1, n2=n2)
        fast = self.owner.run_bytes(self.set, **kw)
        exact = self.owner.run_bytes(self.set, **dict(kw, flags=kw.get("flags", 0) | FLAG_EXACT))
        return max(fast, exact)

    def _admissible(self):
        # a run larger than the whole budget is admitted alone
        return bool(self.pending and self.free and
                    (self.in_flight == 0 or self.in_flight + self.pending[0].bytes <= self.budget))

    def _admit(self):
        while True:
            with self.lock:
                while not self._admissible() and not (self.stopping and not self.pending):
                    self.lock.wait()
                if not self._admissible():  # stopping, nothing left
                    return
                group = [self.pending.popleft()]
                rows = group[0].n2 - group[0].n1
                nbytes = group[0].bytes
                # later batches waiting behind it join the group while the merged run stays within the budget
                while self.merge_rows and self.pending and rows + self.pending[0].n2 - self.pending[0].n1 <= \
                        self.merge_rows:
                    nxt = self.pending[0]
                    b = self.run_bytes(group[0].n1, nxt.n2)
                    if self.in_flight and self.in_flight + b > self.budget:
                        break
                    group.append(self.pending.popleft())
                    rows += nxt.n2 - nxt.n1
                    nbytes = b
                eng = self.free.pop()
                self.in_flight += nbytes
            t = threading.Thread(target=self._work, args=(eng, group, nbytes), daemon=True)
            t.start()
            self.threads.append(t)

    def _work(self, eng, group, nbytes):
        t0 = time.perf_counter()
        for req in group:
            req.started = t0
            req.group = len(group)
        try:
            r = eng.run(self.set, kind=APPEND, n1=group[0].n1, n2=group[-1].n2, certify=True, **self.run_kw)
            for req, part in zip(group, split_links(r, [q.n1 for q in group] + [group[-1].n2])):
                req.result = part
        except Exception as e:  # reported with the requests
            for req in group:
                req.error = e
        t1 = time.perf_counter()
        for req in group:
            req.done = t1
        with self.lock:
            self.in_flight -= nbytes
            self.free.append(eng)
            self.lock.notify_all()
        for req in group:
            req.event.set()

    def run_alone(self, req):
        """The same batch computed again, alone, on the owner context (for the alone-vs-mixed equality check)."""
        r = self.owner.run(self.set, kind=APPEND, n1=req.n1, n2=req.n2, certify=True, **self.run_kw)
        return split_links(r, [req.n1, req.n2])[0]

    def close(self):
        with self.lock:
            self.stopping = True
            self.lock.notify_all()
        self.admitter.join()
        for t in self.threads:
            t.join()
        for e in self.free:
            e.close()
        self.set.free()
        self.owner.close()


def split_links(r, bounds):
    """The links of an APPEND run over [bounds[0], bounds[-1]) split by column into the ranges between consecutive
    bounds (the links stay in tn93's order within each part)."""
    col = r["link_col"]
    parts = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        m = (col >= a) & (col < b)
        parts.append({"link_row": np.asarray(r["link_row"])[m].copy(), "link_col": np.asarray(col)[m].copy(),
                      "link_dist": np.asarray(r["link_dist"])[m].copy(), "certified": r["certified"],
                      "exact": r["exact"]})
    return parts


def query_stream(panel, seed=8, rate=20.0, seconds=60.0, max_batch=512):
    """Poisson arrivals at `rate` batches/s for `seconds`; batch sizes uniform in 1..max_batch; each query is a
    panel sequence with each clean site substituted with probability 0.5-2 % (drawn per query), and 1 % ambiguity
    codes (the c1 generator's proportions)."""
    rs = gen.Stream(seed)
    t, out = 0.0, []
    n, L = panel.shape
    while True:
        t += -np.log(1.0 - rs.uniform(1)[0]) / rate
        if t >= seconds:
            return out
        q = 1 + int(rs.integers(max_batch, 1)[0])
        src = rs.integers(n, q)
        codes = panel[src].copy()
        # substitutions at clean sites only, each site of query k with probability rate_k
        sub = 0.005 + 0.015 * rs.uniform(q)
        hit = (rs.uniform(q * L).reshape(q, L) < sub[:, None]) & (codes < 4)
        shift = (rs.integers(3, int(hit.sum())) + 1).astype(np.uint8)
        codes[hit] = (codes[hit] + shift) % 4
        w = list(gen._C1_WEIGHTS)
        w += [sum(w) * 0.005 / (1.0 - 0.005) / 4.0] * 4
        gen._ambiguities(rs, codes, 0.01, gen._C1_POOL, w)
        out.append((t, codes))


def serve(server, stream, clock_start=None):
    """Plays an arrival stream against the server in real time; returns the requests when all are answered."""
    t0 = time.perf_counter() if clock_start is None else clock_start
    reqs = []
    for k, (t, codes) in enumerate(stream):
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        reqs.append(server.submit(k, codes, arrival=t0 + t))
    for r in reqs:
        r.event.wait()
    return reqs
