"""Benchmark: BPSDP outer iterations on one MI355X.

Workload (default): linear H50 / STO-3G, spacing 1.8 bohr, D2h, (50e,50o)
full active space, DQG + spin constraints, FP64 -- the configuration the
headline metric names.  A step is one outer BPSDP iteration (rhs build, CG on
A A^T, Update_xz eigen split, primal/dual errors), exactly as
v2RDMSolver::compute_energy does it; `value` is outer iterations per second
of the plain-CG leg (the reference's CG) with every vector resident in HBM.

Each leg runs W untimed outer iterations from the RANDOM guess, then K timed
ones that continue the same solve (same mu, iteration count and errors: the
timed call resumes the warm-up call, it does not restart the iteration).

The Update_xz eigen split runs warm: the eigenvectors of the previous outer
iteration seed one block-Jacobi sweep (one cyclic-Jacobi sweep per 64 x 64 pair
sub-problem), which tracks the invariant subspaces from iteration to iteration;
before the stopping rule accepts an iteration its split is redone to
convergence, so a converged solve ends at the reference's fixed point.

Also reported on the same JSON line:
  legs          the same timing for each CG variant (plain CG as in the
                reference, Jacobi-preconditioned CG, learned preconditioner;
                the learned leg trains for --train-iters outer iterations and
                its warm-up covers them); a leg whose learned preconditioner
                fell back to Jacobi is marked invalid;
  roofline      the A^T.y + A.x map kernels (the CG operator), timed with HIP
                events on the engine's stream: algorithmic bytes per launch
                8 (dimx + ncon) / mean launch time, against 8 TB/s HBM;
  eig           the Update_xz eigen split as the loop runs it (wall time per
                outer iteration on the engine stream) and one converged warm
                call (HIP events), each against the algorithmic FP64 flops of
                a full eigendecomposition, sum (16/3) N^3, and 78.6 TFLOP/s;
  peaks         HBM copy bandwidth and FP64 GEMM rate measured on this GPU
                (device copy of 4 GiB, hipBLAS DGEMM 8192^3 through torch);
  cpu_baseline  the CPU oracle (OpenMP maps, MKL dsyev) running the same
                solve's outer iterations on the host, continued from the
                GPU plain leg's state after its timed steps (--cpu-passes of
                them), with the host's CPU model and core counts;
  time_to_converge  the recorded runs to R/E_CONVERGENCE 1e-6 of
                profiles/bpsdp_<workload>_converge.md (bench.py --converge),
                when that file exists.
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


def host_info():
    """CPU model, physical cores and logical CPUs of this host (lscpu)."""
    import subprocess
    info = dict(logical_cpus=os.cpu_count())
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {k.strip(): v.strip() for k, v in
              (line.split(":", 1) for line in out.splitlines() if ":" in line)}
        info["cpu_model"] = kv.get("Model name")
        cores = int(kv.get("Core(s) per socket", 0)) * int(kv.get("Socket(s)", 0))
        info["physical_cores"] = cores or None
    except Exception as e:  # reported, not fatal
        info["lscpu_error"] = str(e)
    return info


def cpu_baseline(w, state, passes, positivity="DQG"):
    """The oracle's own BPSDP loop (plain CG as in the reference, MKL dsyev in
    Update_xz) on the host, continuing the GPU solve from `state` (x, z, y,
    mu, oiter, ep, ed after the GPU's timed steps) for `passes` outer
    iterations: the same iterations the GPU timed, on the CPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    L = orc.Layout(w["amopi"], w["na"], w["nb"], None, positivity)
    c, b = L.build_c(w["h1"], w["eri"]), L.build_b()
    x, z, y = (np.ascontiguousarray(state[k]) for k in ("x", "z", "y"))
    resume = {k: state[k] for k in ("mu", "oiter", "ep", "ed")}
    t0 = time.perf_counter()
    r = L.bpsdp(c, b, x, z, y, maxiter=state["oiter"] + passes, resume=resume)
    dt = time.perf_counter() - t0
    done = r["oiter"] - state["oiter"]
    return dict(value=round(done / dt, 5), unit="outer iter/s", cores=orc.num_threads(),
                kind="port", lapack="MKL dsyev" if orc.lapack_backend() else "built-in Jacobi",
                seconds_per_outer=round(dt / max(1, done), 3),
                cg_iters_per_outer=round(r["cg_iters"] / max(1, done), 2),
                t_cg_s=round(r["t_cg"], 3), t_update_xz_s=round(r["t_xz"], 3), **host_info(),
                sample=(f"{done} outer iterations of the oracle's BPSDP loop (plain CG, MKL "
                        f"dsyev) on the same workload, continuing the GPU plain-CG solve from "
                        f"outer iteration {state['oiter']}; {dt:.1f} s"))


def measure_peaks(dev):
    """HBM copy bandwidth and FP64 GEMM rate on this GPU, outside any profiler:
    a device-to-device copy of 4 GiB (read + write bytes counted) and an
    8192^3 DGEMM through torch (hipBLAS), median of 5 timed repeats each."""
    import torch
    n = (4 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, reps=5):
        fn()
        ts = []
        for _ in range(reps):
            ev[0].record()
            fn()
            ev[1].record()
            ev[1].synchronize()
            ts.append(ev[0].elapsed_time(ev[1]) / 1e3)
        return sorted(ts)[len(ts) // 2]

    t_copy = timed(lambda: dst.copy_(src))
    del src, dst
    m = 8192
    a = torch.randn(m, m, dtype=torch.float64, device=dev)
    bm = torch.randn(m, m, dtype=torch.float64, device=dev)
    t_gemm = timed(lambda: torch.matmul(a, bm))
    del a, bm
    torch.cuda.empty_cache()
    return dict(copy_gbs=round(2 * 8 * n / t_copy / 1e9, 1), dgemm_tflops=round(2 * m ** 3 / t_gemm / 1e12, 2),
                spec_hbm_gbs=HBM_PEAK_GBS, spec_fp64_tflops=FP64_PEAK_TFLOPS,
                how="device copy of 4 GiB (bytes read + written), torch.matmul FP64 8192^3, "
                    "median of 5")


def recorded_converge(workload):
    """Runs to R/E_CONVERGENCE 1e-6 committed under profiles/ (one JSON record
    per line starting with 'record ')."""
    path = os.path.join(ROOT, "profiles", f"bpsdp_{workload}_converge.md")
    if not os.path.exists(path):
        return None
    recs = [json.loads(line[len("record "):]) for line in open(path) if line.startswith("record ")]
    return dict(source=os.path.relpath(path, ROOT), runs=recs) if recs else None


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
                EIG_INNER_SWEEPS=args.inner_sweeps, EIG_PAIR_ORDER=args.eig_order, MAXITER=10 ** 6)

    def progress(i, r):
        if i < 0:
            print(f"# {r['stage']}", file=sys.stderr, flush=True)
            return
        print(f"# point {i}: H{r['n']} r={r['spacing_angstrom']:.3f} E={r['energy']:.8f} "
              f"oiter={r['oiter']} converged={r['converged']}", file=sys.stderr, flush=True)

    if args.scan_threads:
        from v2rdm_casscf_amd.bondscan import solve_scan_threads
        res, wall = solve_scan_threads(pts, positivity=args.positivity, workers=len(pts),
                                       options=opts, max_seconds=args.converge, progress=progress)
    else:
        opts.pop("PRECONDITIONER", None)
        res, wall = solve_scan(pts, positivity=args.positivity, options=opts,
                               max_seconds=args.converge, progress=progress)
    conv = sum(int(r["converged"]) for r in res)
    tot = sum(r["oiter"] for r in res)
    print(json.dumps(dict(mode="bond-scan", engine="threads" if args.scan_threads else "batch",
                          points=len(pts), converged=conv, seconds=round(wall, 2),
                          outer_iterations=tot, outer_iterations_per_s=round(tot / wall, 1),
                          max_outer=max(r["oiter"] for r in res),
                          per_point=[dict(n=r["n"], r=round(r["spacing_angstrom"], 4),
                                          energy=round(r["energy"], 8), oiter=r["oiter"],
                                          converged=r["converged"]) for r in res])))


PC_STUDY_SETTINGS = [(2e-3, 10, 0.0), (5e-4, 10, 0.0), (2e-3, 10, 1.0), (2e-3, 10, 10.0),
                     (5e-4, 3, 1.0), (2e-3, 10, 100.0)]


def _pc_study(args):
    """Learned preconditioner against plain CG to convergence on H8 and H12
    (STO-3G, 1.8 bohr, DQG): one JSON line per run, plain CG first, then the
    learned diagonal for each (learning rate, Adam steps per outer iteration,
    penalty) of PC_STUDY_SETTINGS, trained over the first --train-iters outer
    iterations."""
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd.hamiltonians import hydrogen_chain
    for n in (8, 12):
        d = hydrogen_chain(n, 1.8)

        def run(pc, **kw):
            o = dict(PRECONDITIONER=pc, MLP_TRAIN_ITERS=args.train_iters, MAXITER=100000,
                     EIG_WARM_SWEEPS=args.eig_sweeps, EIG_INNER_SWEEPS=args.inner_sweeps,
                     EIG_PAIR_ORDER=args.eig_order)
            o.update(kw)
            s = V2RDMSolver(d["amopi"], n // 2, n // 2, 1, "DQG", options=o)
            t0 = time.perf_counter()
            r = s.compute_energy(d["h1"], d["eri"], enuc=d["enuc"])
            return dict(n=n, pc=pc, **kw, oiter=r["oiter"], cg=r["cg_iters"],
                        cg_per_outer=round(r["cg_iters"] / r["oiter"], 3),
                        fallback=r["fallback"], seconds=round(time.perf_counter() - t0, 2),
                        objective=r["mlp_loss"], energy=r["energy"])

        print(json.dumps(run("NONE")), flush=True)
        for lr, steps, pen in PC_STUDY_SETTINGS:
            print(json.dumps(run("LEARNED", MLP_TRAIN_LR=lr, MLP_TRAIN_STEPS=steps,
                                 MLP_TRAIN_PENALTY=pen)), flush=True)


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
    ap.add_argument("--no-peaks", action="store_true")
    ap.add_argument("--cpu-passes", type=int, default=2,
                    help="outer iterations of the CPU oracle, continuing the GPU solve")
    ap.add_argument("--train-iters", type=int, default=30,
                    help="outer iterations of online training of the learned preconditioner "
                         "(the learned leg's warm-up covers them)")
    ap.add_argument("--map-reps", type=int, default=20)
    ap.add_argument("--eig-sweeps", type=int, default=1,
                    help="block-Jacobi sweeps per warm Update_xz inside the loop (0: converge); "
                         "one sweep tracks the eigenvectors from iteration to iteration with the "
                         "same outer convergence as fully converged splits")
    ap.add_argument("--eig-order", type=int, default=1,
                    help="pair order of a tracking sweep: 1 cyclic in its first round, then only "
                         "the cross-block pairs (32 steps); 0 cyclic in every round (63 steps)")
    ap.add_argument("--inner-sweeps", type=int, default=1,
                    help="cyclic-Jacobi sweeps per 64x64 pair sub-problem of a warm call (0: 12)")
    ap.add_argument("--positivity", default="DQG", help="D, DQ, DG, DQG, DQGT1, DQGT2, DQGT")
    ap.add_argument("--converge", type=float, default=0.0,
                    help="instead of timing K steps: run from the RANDOM guess to R/E_CONVERGENCE "
                         "1e-6 (stopping after this many seconds) and report time-to-converge")
    ap.add_argument("--scan", type=int, default=0,
                    help="batched bond scan: this many H-chains (n in 8..16, spacing 0.7-3.0 A, "
                         "seed 0) solved concurrently to R/E_CONVERGENCE 1e-6; --converge caps it")
    ap.add_argument("--scan-threads", action="store_true",
                    help="with --scan: one engine handle and stream per point on host threads "
                         "instead of one batch")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="with --converge: time this many CPU-oracle outer iterations too")
    ap.add_argument("--maps-only", action="store_true",
                    help="only the A^T.y / A.x kernels (map-reps calls each) and a device copy "
                         "of 8 (dimx + ncon) bytes as the traffic reference, for counter runs")
    ap.add_argument("--pc-study", action="store_true",
                    help="learned preconditioner vs plain CG to convergence on H8 and H12 over "
                         "a sweep of training settings (one JSON line per run)")
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
    if args.pc_study:
        _pc_study(args)
        return

    dev = torch.device("cuda:0")
    w = _workload(args.workload)
    s = V2RDMSolver(w["amopi"], w["na"], w["nb"], 1, args.positivity, device=dev,
                    options=dict(PRECONDITIONER=args.precond.upper(),
                                 EIG_WARM_SWEEPS=args.eig_sweeps,
                                 EIG_INNER_SWEEPS=args.inner_sweeps,
                                 EIG_PAIR_ORDER=args.eig_order))
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
        s.options["MLP_TRAIN_ITERS"] = args.train_iters
        t_last = [time.perf_counter()]

        def progress(oiter, st):
            # one line every ~20 s, so a long solve shows it is alive
            if time.perf_counter() - t_last[0] > 20.0:
                t_last[0] = time.perf_counter()
                gap = abs(st['energy_primal'] - st['energy_dual'])
                print(f"# oiter {oiter} cg {st['cg_iters']} E {st['energy_primal']:.8f} "
                      f"ep {st['ep']:.2e} ed {st['ed']:.2e} gap {gap:.2e}", file=sys.stderr,
                      flush=True)
            return False

        s.set_iter_callback(progress)
        t0 = time.perf_counter()
        st = s.compute_energy(guess=False, max_seconds=args.converge)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        keep = ("oiter", "cg_iters", "converged", "ep", "ed", "mu", "t_cg", "t_xz",
                "precond_active", "fallback", "cg_iters_train", "train_outer", "mlp_loss",
                "exact_splits")
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

    def timed_leg(solver, warmup):
        """`warmup` untimed outer iterations from the RANDOM guess, then K timed
        ones continuing the same solve (resume: no restart of mu or the count)."""
        solver.options["MAXITER"] = warmup
        st_w = solver.compute_energy(guess=False)
        solver.options["MAXITER"] = st_w["oiter"] + args.steps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.compute_energy(resume=True)
        torch.cuda.synchronize()
        return st_w, st, time.perf_counter() - t0

    def leg_record(st_w, st, el):
        n = st["oiter"] - st_w["oiter"]
        rec = dict(value=round(n / el, 4), ms_per_step=round(1e3 * el / n, 3), steps=n,
                   warmup=st_w["oiter"], cg_iters_per_outer=round(st["cg_iters"] / max(1, n), 2),
                   ep=st["ep"], ed=st["ed"], fallback=st.get("fallback", 0),
                   precond_active=st.get("precond_active", 0))
        if st.get("fallback", 0):
            # the learned preconditioner lost its A/B test and the leg ran
            # Jacobi-CG: not a measurement of the learned method
            rec["valid"] = False
            rec["value"] = None
        return rec

    def leg_warmup(leg):
        return max(args.warmup, args.train_iters + 2) if leg == "learned" else args.warmup

    legs = {}
    for leg in [l for l in args.legs.split(",") if l and l != args.precond]:
        o = dict(s.options, PRECONDITIONER=leg.upper(), MLP_TRAIN_ITERS=args.train_iters)
        sl = V2RDMSolver(w["amopi"], w["na"], w["nb"], 1, args.positivity, device=dev, options=o)
        sl.GetIntegrals(w["h1"], w["eri"])
        sl.BuildConstraints()
        sl.Guess()
        legs[leg] = leg_record(*timed_leg(sl, leg_warmup(leg)))
        print(f"# leg {leg}: {legs[leg]}", file=sys.stderr, flush=True)
        del sl
        torch.cuda.empty_cache()

    s.options["MLP_TRAIN_ITERS"] = args.train_iters
    st_w, st, elapsed = timed_leg(s, leg_warmup(args.precond))
    nsteps = st["oiter"] - st_w["oiter"]
    ms_per_step = 1e3 * elapsed / nsteps
    legs[args.precond] = leg_record(st_w, st, elapsed)
    print(f"# leg {args.precond}: {legs[args.precond]}", file=sys.stderr, flush=True)
    # the solve's state after the timed steps, for the CPU baseline
    state = dict(x=s.x.cpu().numpy(), z=s.z.cpu().numpy(), y=s.y.cpu().numpy(), mu=st["mu"],
                 oiter=st["oiter"], ep=st["ep"], ed=st["ed"])

    # per-kernel timing with HIP events on the engine's stream
    import ctypes
    vp = ctypes.c_void_p
    ncon, dimx = s.nconstraints, s.dimx
    yv = s.y.clone()
    out = torch.empty(ncon, dtype=torch.float64, device=dev)
    tmp = torch.empty(dimx, dtype=torch.float64, device=dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        s.bpsdp_ATu(tmp, yv)
        ev0.record(stream)
        for _ in range(args.map_reps):
            s.lib.v2g_ATu(s.h, vp(yv.data_ptr()), vp(tmp.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_atu = ev0.elapsed_time(ev1) / 1e3 / args.map_reps
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(args.map_reps):
            s.lib.v2g_Au(s.h, vp(tmp.data_ptr()), vp(out.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_au = ev0.elapsed_time(ev1) / 1e3 / args.map_reps
    bytes_map = 8.0 * (dimx + ncon)
    xs, zs = s.x.clone(), s.z.clone()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        s.lib.v2g_update_xz(s.h, ctypes.c_double(st["mu"]), vp(yv.data_ptr()), vp(s.c.data_ptr()),
                            vp(xs.data_ptr()), vp(zs.data_ptr()))
        ev1.record(stream)
    ev1.synchronize()
    t_xz = ev0.elapsed_time(ev1) / 1e3
    del xs, zs, out, tmp, yv
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
    # the eigen split as the loop runs it (warm, tracked: one sweep) and one
    # converged warm call, each against the flops of a full eigendecomposition
    t_xz_loop = st["t_xz"] / max(1, nsteps)
    eig = dict(bound="mfma", achieved=round(eflops / t_xz_loop / 1e12, 3), peak=FP64_PEAK_TFLOPS,
               unit="TFLOP/s", t_update_xz_loop_ms=round(1e3 * t_xz_loop, 3), flops=eflops,
               warm_sweeps_per_call=args.eig_sweeps, inner_sweeps=args.inner_sweeps,
               pair_order=args.eig_order,
               t_update_xz_converged_ms=round(1e3 * t_xz, 3),
               achieved_converged=round(eflops / t_xz / 1e12, 3),
               frac_converged=round(eflops / t_xz / 1e12 / FP64_PEAK_TFLOPS, 4),
               launch_shape=dict(zip(("lanes", "pair_threads"), s.eig_shape())))
    eig["frac"] = round(eig["achieved"] / FP64_PEAK_TFLOPS, 4)
    peaks = None
    if not args.no_peaks:
        try:
            peaks = measure_peaks(dev)
        except Exception as e:  # reported, not fatal
            peaks = dict(error=str(e))
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, state, args.cpu_passes, args.positivity)
        except Exception as e:  # reported, not fatal
            cpu = dict(error=str(e))
    value = nsteps / elapsed
    line = dict(metric=METRIC,
                value=round(value, 4), unit="outer iter/s", n_gpus=1, steps=nsteps,
                warmup=st_w["oiter"], ms_per_step=round(ms_per_step, 3), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload=w["name"], dimx=dimx, ncon=ncon,
                            precond=args.precond, cg_iters_per_outer=round(cg_per_outer, 2),
                            energy=st["energy_primal"] + w["enuc"], ep=st["ep"], ed=st["ed"],
                            mu=st["mu"], t_cg_s=st["t_cg"], t_xz_s=st["t_xz"],
                            eig_warm_sweeps=args.eig_sweeps, eig_inner_sweeps=args.inner_sweeps,
                            warmup_s=st_w["t_total"]),
                legs=legs, roofline=roofline, eig=eig, peaks=peaks, cpu_baseline=cpu,
                time_to_converge=recorded_converge(args.workload))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
