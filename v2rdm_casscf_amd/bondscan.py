"""Batched bond scans: many independent Hamiltonians solved together on one GPU.

solve_scan() puts every point into one batch of the engine (batch.py): each
kernel launch covers all points, points leave the batch as they converge.
solve_scan_threads() is the older way: one engine handle and HIP stream per
point, driven from a pool of host threads.

make_hchain_scan(count, seed) draws the workload of the batched benchmark:
chain lengths uniform in {8, 10, 12, 14, 16} and H-H spacings uniform in
[0.7, 3.0] Angstrom, numpy default_rng(seed).
"""
import concurrent.futures as cf
import time

import numpy as np

from .hamiltonians import ANGSTROM_TO_BOHR, hydrogen_chain
from .solver import V2RDMSolver


def make_hchain_scan(count=64, seed=0, sizes=(8, 10, 12, 14, 16), lo=0.7, hi=3.0):
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(sizes), count)
    rs = rng.uniform(lo, hi, count)
    return [(int(n), float(r)) for n, r in zip(ns, rs)]


def _hamiltonians(points):
    hams = {}
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = {ex.submit(hydrogen_chain, n, r * ANGSTROM_TO_BOHR): i
                for i, (n, r) in enumerate(points)}
        for f in cf.as_completed(futs):
            hams[futs[f]] = f.result()
    return [hams[i] for i in range(len(points))]


def solve_scan(points, positivity="DQG", device="cuda:0", options=None, max_seconds=0.0,
               progress=None):
    """Solve every (n, spacing_angstrom) chain as one batch; returns
    (results, wall seconds of the solve)."""
    from .batch import BatchSolver
    hams = _hamiltonians(points)
    bs = BatchSolver([dict(amopi=d["amopi"], nalpha=d["nalpha"], nbeta=d["nbeta"],
                           positivity=positivity) for d in hams], device=device, options=options)
    bs.set_integrals([d["h1"] for d in hams], [d["eri"] for d in hams])
    if progress:
        progress(-1, dict(stage=f"{len(points)} problems in one batch"))
    t0 = time.perf_counter()
    out = bs.compute_energy(enucs=[d["enuc"] for d in hams], max_seconds=max_seconds,
                            progress=progress is not None)
    wall = time.perf_counter() - t0
    for i, (r, d) in enumerate(zip(out, hams)):
        r.update(n=points[i][0], spacing_angstrom=points[i][1], e_rhf=d["e_rhf"])
        if progress:
            progress(i, r)
    return out, wall


def solve_scan_threads(points, positivity="DQG", workers=16, device="cuda:0", options=None,
                       max_seconds=0.0, progress=None):
    """Solve every (n, spacing_angstrom) chain; returns (results, wall seconds).
    progress(i, result) is called as each point finishes (and, with running=True
    in the result, at most every 20 s while points are iterating)."""
    hams = _hamiltonians(points)
    order = sorted(range(len(points)), key=lambda i: -points[i][0])
    solvers = {}
    for i in order:
        d = hams[i]
        solvers[i] = V2RDMSolver(d["amopi"], d["nalpha"], d["nbeta"], 1, positivity,
                                 device=device, options=options)
        solvers[i].GetIntegrals(d["h1"], d["eri"])
        solvers[i].BuildConstraints()
        solvers[i].Guess()

    if progress:
        progress(-1, dict(stage=f"{len(points)} problems set up"))
    last = [time.perf_counter()]

    def run(i):
        d = hams[i]
        if progress:
            def report(oiter, st):
                # at most one line every 20 s across all points
                now = time.perf_counter()
                if now - last[0] > 20.0:
                    last[0] = now
                    progress(i, dict(st, n=points[i][0], spacing_angstrom=points[i][1],
                                     energy=st["energy_primal"] + d["enuc"], converged=0,
                                     running=True))
                return False
            solvers[i].set_iter_callback(report)
        out = solvers[i].compute_energy(enuc=d["enuc"], guess=False, max_seconds=max_seconds)
        out.update(n=points[i][0], spacing_angstrom=points[i][1], e_rhf=d["e_rhf"])
        if progress:
            progress(i, out)
        return i, out

    t0 = time.perf_counter()
    results = [None] * len(points)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for i, out in ex.map(run, order):
            results[i] = out
    wall = time.perf_counter() - t0
    return results, wall
