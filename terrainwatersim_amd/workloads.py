"""Benchmark workloads beside the headline plain-grid step (used by bench.py).

small    256^2 fBm terrain (seed 1), dry start, rain disc r = 16 (R^2 = 256)
         adding 1e-3 m/step at the centre before every step, closed walls,
         1000 steps: the latency/launch-bound case (state fits in L2/MALL).
batched  256 independent terrains, sides log-uniform 128..2048 (seed 3),
         per-grid random rain discs (r in [8, 32], 1e-3 m/step), 100 steps per
         captured launch sequence, in three layouts: the batch's default
         launch (TWS_KERNEL_AUTO); the one launch the autotuner picks for the
         whole mix; three size classes (sides <= 512, <= 1024, larger), each
         on the (kernel, k) the autotuner picks for its grids, as concurrent
         launch chains.  The fastest is the batched rate.  Compared with the
         same grids stepped one at a time.
served   Poisson stream of solves, sides log-uniform 256..4096, steps uniform
         100..5000, dam-break (20 m column over a centred square of side/8)
         or rain (r = 16 at the centre, 1e-3 m/step) 50/50, arrival rate 70%
         of the measured capacity; the continuous-batching server against the
         plain path taking one request at a time in arrival order.  Latency is
         arrival -> result on the host.  A second server pass submits all
         requests at once: its rate is the server's sustained capacity.
large    16384^2 dam-break (20 m over the centred 2048^2 square) on fBm
         terrain, closed walls, 10,000 steps: throughput and volume drift
         (tests/test_large_gpu.py checks a 1024^2 window against the
         restatements).
"""
from __future__ import annotations

import hashlib
import math
import time

import numpy as np

import terrainwatersim_amd as tws

F32 = np.float32
PARAMS = dict(dt=0.02, gravity=9.81, cell_size=1.0, pipe_area=1.0, damping=0.995)


def _splitmix(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


# ---------------------------------------------------------------------------
def small_state(steps=1000):
    """The small workload's initial state: (b, d, rain, steps)."""
    N = 256
    b = tws.generate_fbm_terrain(N, N, seed=1)
    return b, np.zeros((N, N), F32), (128.0, 128.0, 1e-3, 256.0), steps


def small(steps=1000, repeats=5, log=print, return_depth=False):
    p = tws.params_physical(**PARAMS)
    b, d0, rain, steps = small_state(steps)
    N = b.shape[0]
    log("small: 256^2 rain, 1000 steps")
    times = []
    with tws.Simulation(N, N, p) as sim:
        kernel, k = sim.get_launch()
        for _ in range(repeats + 1):
            sim.upload(b, d0)
            sim.set_rain(*rain)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(steps)
            sim.sync()
            times.append(time.perf_counter() - t0)
        vol = sim.volume()
        depth = sim.download(flux=False)["depth"]
    t = float(np.median(times[1:]))
    rec = {"grid": N, "steps": steps, "kernel": kernel, "steps_per_launch": k,
           "cell_updates_per_s": N * N * steps / t, "ms_per_1000_steps": t * 1e3 * 1000 / steps,
           "volume_m3": vol}
    return (rec, depth) if return_depth else rec


def _draw_sides(rng, n, lo, hi):
    return np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), n))).astype(int)


def batch_sides(n_grids=256, seed=3):
    """Sides of the batched workload's grids (log-uniform 128..2048)."""
    return _draw_sides(np.random.default_rng(seed), n_grids, 128, 2048)


def batch_grids(n_grids=256, seed=3, only=None):
    """The batched workload's grids: (side, terrain, rain disc) per grid; sides
    as batch_sides, terrain seed splitmix(seed * 1000003 + i), rain disc of
    radius U[8, 32] at a uniform centre, 1e-3 m/step.  `only` picks grid
    indices (the terrain of the others is not generated)."""
    rng = np.random.default_rng(seed)
    sides = _draw_sides(rng, n_grids, 128, 2048)
    grids = []
    for i, s in enumerate(sides):
        r = float(rng.uniform(8, 32))
        rain = (float(rng.uniform(0, s)), float(rng.uniform(0, s)), 1e-3, r * r)
        if only is not None and i not in only:
            continue
        b = tws.generate_fbm_terrain(int(s), int(s), seed=_splitmix(seed * 1000003 + i))
        grids.append((int(s), b, rain))
    return grids


# Size classes of the batched workload, by grid side: <= 512, <= 1024, larger.
BATCH_CLASS_SIDES = (512, 1024)


def _time_runs(batch, steps, repeats):
    """Median seconds of `repeats` runs of `steps` steps after one warm run
    (which captures the launch graph)."""
    batch.step(steps)
    batch.sync()
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        batch.step(steps)
        batch.sync()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def batched(n_grids=256, steps=100, seed=3, repeats=3, log=print):
    p = tws.params_physical(**PARAMS)
    grids = batch_grids(n_grids, seed)
    cells = int(sum(s * s for s, _, _ in grids))
    log(f"batched: {n_grids} grids, {cells / 1e6:.1f} M cells; autotuning")
    # tuned on the step count of the runs it serves: a trial whose steps leave
    # a partial last launch would misjudge k
    best, trials = tws.autotune([(s, s) for s, _, _ in grids], steps=steps)
    classes, class_recs, lo = [], [], 0
    for hi in [c * c for c in BATCH_CLASS_SIDES] + [None]:
        shapes = [(s, s) for s, _, _ in grids if s * s > lo and (hi is None or s * s <= hi)]
        if shapes:
            cb, _ = tws.autotune(shapes, steps=steps)
            classes.append((hi or 0, cb["kernel"], cb["k"]))
            class_recs.append({"max_cells": hi, "grids": len(shapes),
                               "cells": int(sum(w * h for w, h in shapes)), "autotune_best": cb})
        lo = hi or lo
    if classes:
        classes[-1] = (0,) + classes[-1][1:]
    pool = sum(tws.grid_footprint(s, s) for s, _, _ in grids)
    out = {"grids": n_grids, "cells": cells, "steps_per_run": steps,
           "autotune_best": best, "autotune_trials": trials, "classes": class_recs}
    with tws.Batch(n_grids, pool) as batch:
        ids = [batch.add(s, s, p) for s, _, _ in grids]

        def load():
            for gid, (s, b, rain) in zip(ids, grids):
                batch.upload(gid, b, np.zeros((s, s), F32))
                batch.set_rain(gid, *rain)

        # every layout runs the same steps from the same state, so their
        # results must agree bit for bit
        times, layout_checks = {}, {}
        for layout in ("default", "one_launch", "classes"):
            load()
            if layout == "default":
                batch.set_launch(tws.KERNEL_AUTO, 0)
            elif layout == "one_launch":
                batch.set_launch(best["kernel"], best["k"])
            else:
                batch.set_launch_classes(classes)
            times[layout] = _time_runs(batch, steps, repeats)
            out[f"batched_{layout}_cell_updates_per_s"] = cells * steps / times[layout]
            layout_checks[layout] = [batch.download(g, flux=False)["depth"] for g in ids[:4]]
        out["graph"] = batch.stats()
        fastest = min(times, key=times.get)
        out["batched_layout"] = fastest
        out["batched_cell_updates_per_s"] = cells * steps / times[fastest]
        out["batched_ms_per_run"] = times[fastest] * 1e3
        checks = layout_checks[fastest]
        out["layouts_equal"] = all(
            np.array_equal(a, b) for lc in layout_checks.values() for a, b in zip(lc, checks))
    # The same grids one at a time, each stepped and synchronised alone: first
    # on its own default launch (what a caller of the plain path gets; the
    # tuner's table is cleared so the batch's result does not leak into it),
    # then forced onto the batch's tuned launch.
    for leg, forced in (("plain", False), ("plain_forced", True)):
        log(f"batched: one grid at a time ({'batch launch' if forced else 'default launches'})")
        tws.tune_forget()
        t, same, launches = _one_at_a_time(grids, p, steps, repeats, checks,
                                           (best["kernel"], best["k"]) if forced else None)
        out[f"{leg}_cell_updates_per_s"] = cells * steps / t
        out[f"{leg}_ms_per_run"] = t * 1e3
        out[f"{leg}_equals_batched"] = bool(same)
        if not forced:
            out["plain_launches"] = launches
    out["batched_over_plain"] = out["batched_cell_updates_per_s"] / out["plain_cell_updates_per_s"]
    out["classes_over_one_launch"] = (out["batched_classes_cell_updates_per_s"]
                                      / out["batched_one_launch_cell_updates_per_s"])
    out["batched_over_plain_forced"] = (out["batched_cell_updates_per_s"]
                                        / out["plain_forced_cell_updates_per_s"])
    out["plain_forced_note"] = ("each grid alone on the tuner's one launch for the whole mix "
                                "(%d/%d); on the streaming kernel a small grid alone is one "
                                "wavefront's serial sweep per segment, so this ratio is not "
                                "the fair one (batched_over_plain is)" % (best["kernel"], best["k"]))
    out["batched_equals_plain"] = bool(out["plain_equals_batched"]
                                       and out["plain_forced_equals_batched"]
                                       and out["layouts_equal"])
    return out


def _one_at_a_time(grids, p, steps, repeats, checks, launch=None):
    """Median seconds to step every grid alone (one Simulation each, stepped
    and synchronised in turn), whether the first grids' depths equal
    `checks`, and the launches used ({"kernel/k": grids})."""
    sims, launches = [], {}
    try:
        for s, b, rain in grids:
            sim = tws.Simulation(s, s, p)
            if launch is not None:
                sim.set_launch(*launch)
            kl = "%d/%d" % sim.get_launch()
            launches[kl] = launches.get(kl, 0) + 1
            sim.upload(b, np.zeros((s, s), F32))
            sim.set_rain(*rain)
            sims.append(sim)
        for sim in sims:
            sim.step(steps)
            sim.sync()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            for sim in sims:
                sim.step(steps)
                sim.sync()
            times.append(time.perf_counter() - t0)
        same = all(np.array_equal(c, sim.download(flux=False)["depth"])
                   for c, sim in zip(checks, sims[:len(checks)]))
    finally:
        for sim in sims:
            sim.close()
    return float(np.median(times)), same, launches


# ---------------------------------------------------------------------------
class Requests:
    """The served request stream: terrain windows cut from one fBm field."""

    def __init__(self, n, seed, rate):
        rng = np.random.default_rng(seed)
        self.sides = _draw_sides(rng, n, 256, 4096)
        self.steps = rng.integers(100, 5001, n)
        self.kind = rng.integers(0, 2, n)  # 0 dam-break, 1 rain
        self.arrival = np.cumsum(rng.exponential(1.0 / rate, n))
        self.arrival -= self.arrival[0]
        self.offsets = rng.integers(0, 4096, (n, 2))
        self.field = tws.generate_fbm_terrain(4096, 4096, seed=seed)

    def state(self, i):
        s = int(self.sides[i])
        oy, ox = self.offsets[i]
        ys = (np.arange(s) + oy) % 4096
        xs = (np.arange(s) + ox) % 4096
        b = np.ascontiguousarray(self.field[np.ix_(ys, xs)])
        d = np.zeros((s, s), F32)
        rain = None
        if self.kind[i] == 0:
            a = s // 8
            c0 = (s - a) // 2
            d[c0:c0 + a, c0:c0 + a] = 20.0
        else:
            rain = (s / 2.0, s / 2.0, 1e-3, 256.0)
        return b, d, rain

    def work(self):
        return float(np.sum(self.sides.astype(np.float64) ** 2 * self.steps))


def _pcts(lat):
    lat = np.asarray(lat)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


# Requests resident on the device at once in the served workload.  MI355X,
# saturated capacity 4.36-4.38e11 cell-updates/s at 128 against 4.30e11 at 64,
# with the same latency at 70% load (p50 0.143 s, p99 0.76 s); the pool's
# footprint test still bounds the resident set.
MAX_RESIDENT = 128


def served(capacity_cups, n_requests=200, seed=4, load=0.7, chunk_steps=48, log=print,
           device_cap=None):
    p = tws.params_physical(**PARAMS)
    probe = Requests(n_requests, seed, 1.0)
    mean_work = probe.work() / n_requests
    rate = load * capacity_cups / mean_work
    req = Requests(n_requests, seed, rate)
    log(f"served: {n_requests} requests, mean work {mean_work:.3g} cell-updates, "
        f"rate {rate:.2f}/s ({load:.0%} of {capacity_cups:.3g} cell-updates/s)")
    states = [req.state(i) for i in range(n_requests)]
    pool = int(min(96 << 30, 40 * tws.grid_footprint(4096, 4096)))
    digests_served = {}
    lat_served = []
    with tws.Server(pool, max_resident=MAX_RESIDENT, chunk_steps=chunk_steps) as srv:
        tickets = {}
        t0 = time.perf_counter()
        base = srv.now()
        for i in range(n_requests):
            wait = req.arrival[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            b, d, rain = states[i]
            t, out = srv.submit(b, int(req.steps[i]), p, depth=d, rain=rain, want_flux=False)
            tickets[t] = (i, out)
        srv.drain()
        t_end = 0.0
        for t, (i, out) in tickets.items():
            st = srv.poll(t)
            assert st.state == tws.REQ_DONE, (i, st.state, st.error)
            lat_served.append(st.t_done - st.t_submit)
            t_end = max(t_end, st.t_done - base)
            digests_served[i] = hashlib.sha1(out["depth"].tobytes()).hexdigest()
            srv.release(t)
        stats = srv.stats()
    work = req.work()
    res = {"requests": n_requests, "arrival_rate_per_s": rate, "load": load,
           "chunk_steps": chunk_steps, "max_resident": MAX_RESIDENT, "cell_updates": work,
           "served": {"cell_updates_per_s": work / t_end, "makespan_s": t_end,
                      **_pcts(lat_served), "batched_runs": stats["runs"],
                      "gpu_busy_s": stats["busy_seconds"]}}
    # capacity: the same requests all at once (the server saturated)
    log(f"served: done in {t_end:.2f}s; saturated pass")
    with tws.Server(pool, max_resident=MAX_RESIDENT, chunk_steps=chunk_steps) as srv:
        base = srv.now()
        tickets = []
        for i in range(n_requests):
            b, d, rain = states[i]
            tickets.append(srv.submit(b, int(req.steps[i]), p, depth=d, rain=rain,
                                      want_flux=False)[0])
        srv.drain()
        t_sat = max(srv.poll(t).t_done for t in tickets) - base
        st = srv.stats()
    res["served_saturated"] = {"cell_updates_per_s": work / t_sat, "makespan_s": t_sat,
                               "gpu_busy_s": st["busy_seconds"]}
    log("served: plain one-at-a-time")
    lat_plain, same = [], True
    t0 = time.perf_counter()
    for i in range(n_requests):
        wait = req.arrival[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        b, d, rain = states[i]
        s = b.shape[0]
        with tws.Simulation(s, s, p) as sim:
            sim.upload(b, d)
            if rain:
                sim.set_rain(*rain)
            sim.step(int(req.steps[i]))
            dd = sim.download(flux=False)["depth"]
        lat_plain.append(time.perf_counter() - t0 - req.arrival[i])
        same &= hashlib.sha1(dd.tobytes()).hexdigest() == digests_served[i]
    t_plain = time.perf_counter() - t0
    res["plain"] = {"cell_updates_per_s": work / t_plain, "makespan_s": t_plain,
                    **_pcts(lat_plain)}
    # at 70% load the served rate is the offered load; the plain path is
    # saturated (its makespan far exceeds the arrival span), so the capacity
    # ratio compares the saturated server with it
    res["served_over_plain"] = res["served"]["cell_updates_per_s"] / res["plain"]["cell_updates_per_s"]
    res["served_capacity_over_plain"] = (res["served_saturated"]["cell_updates_per_s"]
                                         / res["plain"]["cell_updates_per_s"])
    res["served_equals_plain"] = bool(same)
    if device_cap:
        # No schedule steps the stream faster than the device's best rate on
        # its sizes, so served/plain can reach at most cap / plain rate.
        res["served_over_plain_bound"] = {
            "device_cap_cell_updates_per_s": device_cap,
            "max_ratio": device_cap / res["plain"]["cell_updates_per_s"],
            "cap_source": "the fastest measured rate on the stream's sizes (the 4096^2 "
                          "single-grid step, or the batched mix if faster)"}
    return res


# ---------------------------------------------------------------------------
def large_state(N=16384, seed=5):
    """fBm terrain and the dam-break depth: 20 m over the centred 2048^2 square."""
    b = tws.generate_fbm_terrain(N, N, seed=seed)
    d = np.zeros((N, N), F32)
    c0 = (N - 2048) // 2
    d[c0:c0 + 2048, c0:c0 + 2048] = 20.0
    return b, d


def large(steps=10000, N=16384, seed=5, log=print):
    p = tws.params_physical(**PARAMS)
    log(f"large: {N}^2 fBm terrain, dam-break")
    b, d = large_state(N, seed)
    out = {"grid": N, "steps": steps, "footprint_bytes": tws.grid_footprint(N, N)}
    with tws.Simulation(N, N, p) as sim:
        sim.upload(b, d)
        del b, d
        v0 = sim.volume()
        sim.step(50)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps - 50)
        sim.sync()
        el = time.perf_counter() - t0
        v1 = sim.volume()
        ok = sim.check_finite()
        kernel, k = sim.get_launch()
    out.update({"cell_updates_per_s": N * N * (steps - 50) / el, "seconds": el,
                "kernel": kernel, "steps_per_launch": k, "finite": ok,
                "volume_initial_m3": v0, "volume_final_m3": v1,
                "volume_rel_drift": abs(v1 - v0) / v0})
    return out


def batched_default(n_grids=256, steps=100, runs=3, seed=3, log=print):
    """The batched mix on the batch's default launch alone: one untimed run
    (the graph capture), then `runs` timed runs of `steps` steps.  The short
    form that a profiler wraps (no autotuning, no one-at-a-time legs)."""
    p = tws.params_physical(**PARAMS)
    grids = batch_grids(n_grids, seed)
    cells = int(sum(s * s for s, _, _ in grids))
    pool = sum(tws.grid_footprint(s, s) for s, _, _ in grids)
    with tws.Batch(n_grids, pool) as batch:
        for s, b, rain in grids:
            gid = batch.add(s, s, p)
            batch.upload(gid, b, np.zeros((s, s), F32))
            batch.set_rain(gid, *rain)
        t = _time_runs(batch, steps, runs)
        log(f"batched_default: {n_grids} grids, {cells / 1e6:.1f} M cells, {steps} steps per "
            f"run: {cells * steps / t:.4g} cell-updates/s ({t * 1e3:.2f} ms per run)")
        return {"grids": n_grids, "cells": cells, "steps_per_run": steps,
                "cell_updates_per_s": cells * steps / t, "ms_per_run": t * 1e3,
                "graph": batch.stats()}


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Run one workload of this module and print its "
                                             "record as JSON.")
    ap.add_argument("workload", choices=["small", "batched", "batched_default", "served", "large"])
    ap.add_argument("--runs", type=int, default=3, help="timed runs (batched_default)")
    ap.add_argument("--capacity", type=float, default=4.3e11,
                    help="served: cell-updates/s the arrival rate is a fraction of")
    ap.add_argument("--requests", type=int, default=200, help="served: requests")
    a = ap.parse_args()
    if a.workload == "batched_default":
        rec = batched_default(runs=a.runs)
    elif a.workload == "served":
        rec = served(a.capacity, n_requests=a.requests)
    else:
        rec = globals()[a.workload]()
    print(json.dumps(rec, default=float))
