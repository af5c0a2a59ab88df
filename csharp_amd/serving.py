"""Open-loop serving driver: Poisson query arrivals against a Server.

Latency of a query = its submit time -> the time its last distance is in
host memory (the server's own steady-clock stamps); GCUPS = DP cells of the
completed queries / (last completion - first submit); admitted concurrency =
queries admitted to a batch and not yet done, averaged over that span and at
its peak; queue wait = submit -> admission.
"""
import threading
import time

import numpy as np


def run_poisson(server, queries, rate, seed=0, timeout_s=600.0):
    """Submit `queries` (list of bytes) with exponential inter-arrival times
    of mean 1/rate seconds; returns a dict of latency percentiles and GCUPS."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(queries))
    arrivals = np.cumsum(gaps) - gaps[0]
    tickets = [None] * len(queries)
    err = []

    def client():
        t0 = time.perf_counter()
        try:
            for k, q in enumerate(queries):
                delay = arrivals[k] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets[k] = server.submit(q)
        except Exception as e:  # surfaced to the caller below
            err.append(e)

    n0, cells0, _ = server.stats()
    th = threading.Thread(target=client)
    th.start()
    th.join()
    if err:
        raise err[0]
    deadline = time.time() + timeout_s
    for t in tickets:
        while not server.wait(t, 1000):
            if time.time() > deadline:
                raise TimeoutError("serving run did not drain")
    n1, cells1, _ = server.stats()
    sub, adm, done = [], [], []
    for t in tickets:
        s, a, d = server.times(t)
        sub.append(s)
        adm.append(a)
        done.append(d)
        server.release(t)
    sub = np.array(sub)
    adm = np.array(adm)
    done = np.array(done)
    lat = done - sub
    span_s = (done.max() - sub.min()) / 1e3
    # in flight between admission and completion: +1 at admit, -1 at done
    ev = np.concatenate([np.stack([adm, np.ones_like(adm)], 1), np.stack([done, -np.ones_like(done)], 1)])
    ev = ev[np.lexsort((ev[:, 1], ev[:, 0]))]
    peak = int(np.cumsum(ev[:, 1]).max())
    return {
        "rate_qps": rate,
        "queries": len(queries),
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "mean_ms": round(float(lat.mean()), 2),
        "gcups": round((cells1 - cells0) / span_s / 1e9, 1),
        "achieved_qps": round(len(queries) / span_s, 1),
        "queue_wait_p50_ms": round(float(np.percentile(adm - sub, 50)), 2),
        "admitted_concurrency_mean": round(float((done - adm).sum() / 1e3 / span_s), 2),
        "admitted_concurrency_max": peak,
    }
