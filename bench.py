"""Benchmark: BPSDP outer iterations on one MI355X.

Workload (default): linear H50 / STO-3G, spacing 1.8 bohr, D2h, (50e,50o)
full active space, DQG + spin constraints, FP64 -- the configuration the
headline metric names.  A step is one outer BPSDP iteration (rhs build, CG on
A A^T, Update_xz eigen split, primal/dual errors), exactly as
v2RDMSolver::compute_energy does it; `value` is outer iterations per second
with every vector resident in HBM.

The Update_xz eigen split runs warm: the eigenvectors of the previous outer
iteration seed one block-Jacobi sweep (one cyclic-Jacobi sweep per 64 x 64 pair
sub-problem), which tracks the invariant subspaces from iteration to iteration;
on H24 this reaches R/E_CONVERGENCE 1e-6 in the same number of outer
iterations, at the same energy, as fully converged eigendecompositions.

Also reported on the same JSON line:
  legs          the same timing for each CG variant (plain CG as in the
                reference, Jacobi-preconditioned CG, learned preconditioner
                trained during the warmup), each from the same RANDOM guess;
  roofline      the A^T.y + A.x map kernels (the CG operator), timed with HIP
                events on the engine's stream: algorithmic bytes per launch
                8 (dimx + ncon) / mean launch time, against 8 TB/s HBM;
  eig           the Update_xz eigen split as the loop runs it (wall time per
                outer iteration on the engine stream), against the algorithmic FP64
                flops of a full eigendecomposition, sum (16/3) N^3, and 78.6 TFLOP/s;
                one converged warm call, timed with HIP events, is reported beside it;
  cpu_baseline  the CPU oracle (OpenMP maps + MKL dsyev) on the host, on a
                bounded sample (one CG operator application and one Update_xz
                timed; an outer iteration projected with the GPU run's mean CG
                iterations per outer iteration).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-6Eh & outer iters/s, (50e,50o) DQG: learned-PC vs Jacobi-CG vs CPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _workload(name):
    from v2rdm_casscf_amd.hamiltonians import hydrogen_chain, synthetic_cholesky
    if name == "h50":
        d = hydrogen_chain(50, 1.8)
        return dict(name="H50 STO-3G r=1.8 bohr D2h (50e,50o) DQG+spin", amopi=d["amopi"],
                    na=25, nb=25, h1=d["h1"], eri=d["eri"], enuc=d["enuc"])
    if name.startswith("h"):
        n = int(name[1:])
        d = hydrogen_chain(n, 1.8)
        return dict(name=f"H{n} STO-3G r=1.8 bohr D2h ({n}e,{n}o) DQG+spin", amopi=d["amopi"],
                    na=n // 2, nb=n // 2, h1=d["h1"], eri=d["eri"], enuc=d["enuc"])
    if name.startswith("syn"):
        # synN or synNsK: n = N, seed K (default 0), L = 3n Cholesky vectors
        body = name[3:]
        n, seed = (int(body.split("s")[0]), int(body.split("s")[1])) if "s" in body else (int(body), 0)
        h1, eri = synthetic_cholesky(n, seed)
        return dict(name=f"synthetic Cholesky n={n} seed {seed} C1 ({n}e,{n}o)", amopi=[n],
                    na=n // 2, nb=n // 2, h1=h1, eri=eri, enuc=0.0)
    raise ValueError(name)


def _map_traffic(workload):
    """HBM bytes per A^T.y / A.x launch from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by a plain copy kernel)."""
    path = os.path.join(ROOT, "profiles", f"maps_{workload}_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            kv = dict(tok.split("=") for tok in line.split()[1:])
            return {k: int(v) for k, v in kv.items()}
    return None


def _eig_flops(dims):
    return sum((16.0 / 3.0) * float(n) ** 3 for n in dims)


def cpu_baseline(w, gpu_cg_per_outer, positivity="DQG"):
    """Time the oracle's CG operator and Update_xz on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    L = orc.Layout(w["amopi"], w["na"], w["nb"], None, positivity)
    c = L.build_c(w["h1"], w["eri"])
    x, z, y = L.guess_random()
    t0 = time.time()
    L.cg_Ax(y)
    t_op = time.time() - t0
    t0 = time.time()
    L.update_xz(1.0, y, c, x)
    t_xz = time.time() - t0
    per_outer = t_xz + (gpu_cg_per_outer + 4.0) * t_op
    return dict(value=1.0 / per_outer, unit="outer iter/s", cores=orc.num_threads(),
                kind="port", lapack="MKL dsyev" if orc.lapack_backend() else "built-in Jacobi",
                sample=(f"oracle on the same workload: one A(A^T y) = {t_op:.3f} s and one "
                        f"Update_xz = {t_xz:.3f} s timed; outer iteration projected as "
                        f"Update_xz + ({gpu_cg_per_outer:.1f} CG + 4 map) applications"))


def _maps_only(s, stream, reps, w):
    """A^T.y and A.x alone (plus a reference copy) for hardware-counter runs."""
    import ctypes
    import torch
    dev = s.y.device
    ncon, dimx = s.nconstraints, s.dimx
    yv = s.y.clone()
    aty = torch.empty(dimx, dtype=torch.float64, device=dev)
    ax = torch.empty(ncon, dtype=torch.float64, device=dev)
    src = torch.empty(dimx + ncon, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    with torch.cuda.stream(stream):
        ev[0].record(stream)
        for _ in range(reps):
            s.lib.v2g_ATu(s.h, ctypes.c_void_p(yv.data_ptr()), ctypes.c_void_p(aty.data_ptr()))
        ev[1].record(stream)
        for _ in range(reps):
            s.lib.v2g_Au(s.h, ctypes.c_void_p(aty.data_ptr()), ctypes.c_void_p(ax.data_ptr()))
        ev[2].record(stream)
        for _ in range(reps):
            dst.copy_(src)
        ev[3].record(stream)
    ev[3].synchronize()
    out = dict(workload=w["name"], dimx=dimx, ncon=ncon, reps=reps,
               bytes_map=8.0 * (dimx + ncon), bytes_copy=2 * 8.0 * (dimx + ncon),
               t_atu_ms=ev[0].elapsed_time(ev[1]) / reps, t_au_ms=ev[1].elapsed_time(ev[2]) / reps,
               t_copy_ms=ev[2].elapsed_time(ev[3]) / reps)
    print(json.dumps(out))


def _scan(args):
    """Many independent H-chains at once (one handle and stream per chain)."""
    from v2rdm_casscf_amd.bondscan import make_hchain_scan, solve_scan
    pts = make_hchain_scan(args.scan, seed=0)
    opts = dict(PRECONDITIONER=args.precond.upper(), EIG_WARM_SWEEPS=args.eig_sweeps,
                EIG_INNER_SWEEPS=args.inner_sweeps, MAXITER=10 ** 6)

    def progress(i, r):
        if i < 0:
            print(f"# {r['stage']}", file=sys.stderr, flush=True)
            return
        print(f"# point {i}: H{r['n']} r={r['spacing_angstrom']:.3f} E={r['energy']:.8f} "
              f"oiter={r['oiter']} converged={r['converged']}", file=sys.stderr, flush=True)

    res, wall = solve_scan(pts, positivity=args.positivity, workers=len(pts), options=opts,
                           max_seconds=args.converge, progress=progress)
    conv = sum(int(r["converged"]) for r in res)
    tot = sum(r["oiter"] for r in res)
    print(json.dumps(dict(mode="bond-scan", points=len(pts), converged=conv, seconds=round(wall, 2),
                          outer_iterations=tot, outer_iterations_per_s=round(tot / wall, 1),
                          max_outer=max(r["oiter"] for r in res),
                          per_point=[dict(n=r["n"], r=round(r["spacing_angstrom"], 4),
                                          energy=round(r["energy"], 8), oiter=r["oiter"],
                                          converged=r["converged"]) for r in res])))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="h50")
    ap.add_argument("--precond", default="none", choices=["none", "jacobi", "learned"],
                    help="CG of the timed leg that gives `value`: none = the reference's plain CG")
    ap.add_argument("--legs", default="none,jacobi,learned",
                    help="CG variants timed side by side (same guess, warmup and steps each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--map-reps", type=int, default=20)
    ap.add_argument("--eig-sweeps", type=int, default=1,
                    help="block-Jacobi sweeps per warm Update_xz inside the loop (0: converge); "
                         "one sweep tracks the eigenvectors from iteration to iteration with the "
                         "same outer convergence as fully converged splits")
    ap.add_argument("--inner-sweeps", type=int, default=1,
                    help="cyclic-Jacobi sweeps per 64x64 pair sub-problem of a warm call (0: 12)")
    ap.add_argument("--positivity", default="DQG", help="D, DQ, DG, DQG, DQGT1, DQGT2, DQGT")
    ap.add_argument("--converge", type=float, default=0.0,
                    help="instead of timing K steps: run from the RANDOM guess to R/E_CONVERGENCE "
                         "1e-6 (stopping after this many seconds) and report time-to-converge")
    ap.add_argument("--scan", type=int, default=0,
                    help="batched bond scan: this many H-chains (n in 8..16, spacing 0.7-3.0 A, "
                         "seed 0) solved concurrently to R/E_CONVERGENCE 1e-6; --converge caps it")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="with --converge: time this many CPU-oracle outer iterations too")
    ap.add_argument("--maps-only", action="store_true",
                    help="only the A^T.y / A.x kernels (map-reps calls each) and a device copy "
                         "of 8 (dimx + ncon) bytes as the traffic reference, for counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X (--gpus 1); --gpus {args.gpus} is not supported: "
              "the engine runs one process on one GPU and has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from v2rdm_casscf_amd import V2RDMSolver
    if args.scan > 0:
        _scan(args)
        return

    dev = torch.device("cuda:0")
    w = _workload(args.workload)
    s = V2RDMSolver(w["amopi"], w["na"], w["nb"], 1, args.positivity, device=dev,
                    options=dict(PRECONDITIONER=args.precond.upper(),
                                 EIG_WARM_SWEEPS=args.eig_sweeps,
                                 EIG_INNER_SWEEPS=args.inner_sweeps))
    stream = torch.cuda.Stream(device=dev)
    s.lib.v2g_set_stream(s.h, __import__("ctypes").c_void_p(stream.cuda_stream))
    s.GetIntegrals(w["h1"], w["eri"])
    s.BuildConstraints()
    s.Guess()
    torch.cuda.synchronize()
    if args.maps_only:
        _maps_only(s, stream, args.map_reps, w)
        return
    if args.converge > 0:
        s.options["MAXITER"] = 10 ** 7
        t_last = [time.perf_counter()]

        def progress(oiter, st):
            # one line every ~20 s, so a long solve shows it is alive
            if time.perf_counter() - t_last[0] > 20.0:
                t_last[0] = time.perf_counter()
                print(f"# oiter {oiter} cg {st['cg_iters']} E {st['energy_primal']:.8f} "
                      f"ep {st['ep']:.2e} ed {st['ed']:.2e}", file=sys.stderr, flush=True)
            return False

        s.set_iter_callback(progress)
        t0 = time.perf_counter()
        st = s.compute_energy(guess=False, max_seconds=args.converge)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        keep = ("oiter", "cg_iters", "converged", "ep", "ed", "mu", "t_cg", "t_xz",
                "precond_active", "fallback", "cg_iters_train", "train_outer", "mlp_loss")
        rec = dict(mode="time-to-converge", workload=w["name"], positivity=args.positivity,
                   precond=args.precond, dimx=s.dimx, ncon=s.nconstraints, seconds=round(secs, 3),
                   energy=st["energy_primal"] + w["enuc"], energy_dual=st["energy_dual"] + w["enuc"],
                   **{k: st[k] for k in keep if k in st})
        if args.cpu_iters > 0:
            # the CPU oracle (the reference's plain CG, LAPACK eigensolver) on the
            # same guess for a few outer iterations: seconds per outer iteration,
            # and the time-to-converge it implies at the GPU run's iteration count
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py as orc
            Lc = orc.Layout(w["amopi"], w["na"], w["nb"], None, args.positivity)
            cc = Lc.build_c(w["h1"], w["eri"])
            xc, zc, yc = Lc.guess_random()
            t0 = time.perf_counter()
            rc = Lc.bpsdp(cc, Lc.build_b(), xc, zc, yc, maxiter=args.cpu_iters)
            cpu_it = (time.perf_counter() - t0) / max(1, rc["oiter"])
            rec["cpu_baseline"] = dict(seconds_per_outer=round(cpu_it, 4), cores=orc.num_threads(),
                                       kind="port", sample=f"{rc['oiter']} oracle outer iterations",
                                       projected_seconds=round(cpu_it * st["oiter"], 1))
        print(json.dumps(rec))
        return

    def timed_leg(solver):
        """W untimed outer iterations from the RANDOM guess, then K timed ones."""
        solver.options["MAXITER"] = args.warmup
        st_w = solver.compute_energy(guess=False)
        solver.options["MAXITER"] = args.steps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.compute_energy(guess=False)
        torch.cuda.synchronize()
        return st_w, st, time.perf_counter() - t0

    legs = {}
    for leg in [l for l in args.legs.split(",") if l and l != args.precond]:
        o = dict(s.options, PRECONDITIONER=leg.upper(), MLP_TRAIN_ITERS=max(1, args.warmup - 2))
        sl = V2RDMSolver(w["amopi"], w["na"], w["nb"], 1, args.positivity, device=dev, options=o)
        sl.GetIntegrals(w["h1"], w["eri"])
        sl.BuildConstraints()
        sl.Guess()
        _, stl, el = timed_leg(sl)
        legs[leg] = dict(value=round(stl["oiter"] / el, 4), ms_per_step=round(1e3 * el / stl["oiter"], 3),
                         cg_iters_per_outer=round(stl["cg_iters"] / max(1, stl["oiter"]), 2),
                         ep=stl["ep"], ed=stl["ed"], fallback=stl.get("fallback", 0),
                         precond_active=stl.get("precond_active", 0))
        del sl
        torch.cuda.empty_cache()

    if args.precond == "learned":
        s.options["MLP_TRAIN_ITERS"] = max(1, args.warmup - 2)
    st_w, st, elapsed = timed_leg(s)
    nsteps = st["oiter"]
    ms_per_step = 1e3 * elapsed / nsteps
    legs[args.precond] = dict(value=round(nsteps / elapsed, 4), ms_per_step=round(ms_per_step, 3),
                              cg_iters_per_outer=round(st["cg_iters"] / max(1, nsteps), 2),
                              ep=st["ep"], ed=st["ed"], fallback=st.get("fallback", 0),
                              precond_active=st.get("precond_active", 0))

    # per-kernel timing with HIP events on the engine's stream
    ncon, dimx = s.nconstraints, s.dimx
    yv = s.y.clone()
    out = torch.empty(ncon, dtype=torch.float64, device=dev)
    tmp = torch.empty(dimx, dtype=torch.float64, device=dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        s.bpsdp_ATu(tmp, yv)
        ev0.record(stream)
        for _ in range(args.map_reps):
            s.lib.v2g_ATu(s.h, __import__("ctypes").c_void_p(yv.data_ptr()),
                          __import__("ctypes").c_void_p(tmp.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_atu = ev0.elapsed_time(ev1) / 1e3 / args.map_reps
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(args.map_reps):
            s.lib.v2g_Au(s.h, __import__("ctypes").c_void_p(tmp.data_ptr()),
                         __import__("ctypes").c_void_p(out.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_au = ev0.elapsed_time(ev1) / 1e3 / args.map_reps
    bytes_map = 8.0 * (dimx + ncon)
    xs, zs = s.x.clone(), s.z.clone()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        s.lib.v2g_update_xz(s.h, __import__("ctypes").c_double(st["mu"]),
                            __import__("ctypes").c_void_p(yv.data_ptr()),
                            __import__("ctypes").c_void_p(s.c.data_ptr()),
                            __import__("ctypes").c_void_p(xs.data_ptr()),
                            __import__("ctypes").c_void_p(zs.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_xz = ev0.elapsed_time(ev1) / 1e3
    eflops = _eig_flops([n for n in s.dimensions if n > 0])

    cg_per_outer = st["cg_iters"] / max(1, nsteps)
    roofline = dict(bound="hbm", kernel="A^T.y (k_atu) + A.x (k_au)",
                    achieved=round(2 * bytes_map / (t_atu + t_au) / 1e9, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", traffic=None, t_atu_ms=round(1e3 * t_atu, 4),
                    t_au_ms=round(1e3 * t_au, 4), bytes_per_launch=bytes_map)
    roofline["frac"] = round(roofline["achieved"] / HBM_PEAK_GBS, 4)
    tr = _map_traffic(args.workload)
    if tr and "k_atu" in tr and "k_au" in tr:
        roofline["traffic"] = int((tr["k_atu"] + tr["k_au"]) / 2)
        roofline["traffic_k_atu"], roofline["traffic_k_au"] = tr["k_atu"], tr["k_au"]
    # the eigen split as the loop runs it (warm, tracked): its wall time per outer
    # iteration on the engine stream, against the flops of a full eigendecomposition;
    # the converged warm call measured with events above is reported beside it
    t_xz_loop = st["t_xz"] / max(1, nsteps)
    eig = dict(bound="mfma", achieved=round(eflops / t_xz_loop / 1e12, 3), peak=FP64_PEAK_TFLOPS,
               unit="TFLOP/s", t_update_xz_loop_ms=round(1e3 * t_xz_loop, 3), flops=eflops,
               warm_sweeps_per_call=args.eig_sweeps, inner_sweeps=args.inner_sweeps,
               t_update_xz_converged_ms=round(1e3 * t_xz, 3))
    eig["frac"] = round(eig["achieved"] / FP64_PEAK_TFLOPS, 4)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, cg_per_outer, args.positivity)
        except Exception as e:  # reported, not fatal
            cpu = dict(error=str(e))
    value = nsteps / elapsed
    line = dict(metric=METRIC,
                value=round(value, 4), unit="outer iter/s", n_gpus=1, steps=nsteps,
                warmup=args.warmup, ms_per_step=round(ms_per_step, 3), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload=w["name"], dimx=dimx, ncon=ncon,
                            precond=args.precond, cg_iters_per_outer=round(cg_per_outer, 2),
                            energy=st["energy_primal"] + w["enuc"], ep=st["ep"], ed=st["ed"],
                            mu=st["mu"], t_cg_s=st["t_cg"], t_xz_s=st["t_xz"],
                            eig_warm_sweeps=args.eig_sweeps, eig_inner_sweeps=args.inner_sweeps,
                            warmup_s=st_w["t_total"]),
                legs=legs, roofline=roofline, eig=eig, cpu_baseline=cpu)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
