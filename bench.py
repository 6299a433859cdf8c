"""Benchmark of the MI355X Rocflu hot path (one GPU).

Workload (default, ``curtain_spec``): the two-way coupled shock-particle
curtain -- 400x80x80 hexes (2,560,000 cells), a Mach 2 shock into quiescent
air, 4,000,000 particles (lognormal diameters, median 115 um, sigma_ln 0.2,
rho_p 2500) uniform in a slab at 20% volume fraction; second-order AUSM+
with LSQ gradients and the Venkatakrishnan limiter; trilinear interpolation,
Schiller-Naumann quasi-steady drag + added mass + pressure gradient,
Gaussian-filter projection of width 2 dx; Wray RK3, FP64.  A step is one
RFLU_TimeStepping iteration: dt evaluation + 3 coupled RK stages.

Prints one JSON line (see the driver contract): value = Mcell-updates/s;
roofline = the dominant kernel (the one with the most time in the timed
region) with its algorithmic bytes per launch over its mean launch
duration (HIP events on the library's stream, over the timed region);
cpu_baseline = the CPU oracle (a C restatement of the reference loops, its
independent loops on all host cores, and on one core) on a bounded sample of
the same workload; secondary = the second-order tet workload (64^3 x 5 tets,
Roe) measured the same way.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nC, nF, nB, mbar, order):
    """Compulsory FP64 HBM bytes per launch of each kernel.

    Partition of the reference-dataflow byte model: per step
    B_dt = 48 Nc + 40 Nf + 36 Nb; per stage (first order)
    141.33 Nc + 40 Nf + 36 Nb, (second order) adds (24 + 48 + 4 mbar + 4) Nc
    + 24 Nf + 24 Nb.  Each stage's share is attributed to the kernel that
    streams it: k_grad the stencil geometry (cofg, R-factor, CSR), k_flux the
    face geometry and connectivity, k_update the conserved state, volume and
    RK register.
    """
    b = {"k_dtmin": 48.0 * nC + 40.0 * nF + 36.0 * nB}
    if order == 2:
        b["k_grad"] = (24.0 + 48.0 + 4.0 * mbar + 4.0) * nC
        b["k_flux"] = 64.0 * nF + 60.0 * nB
    else:
        b["k_grad"] = 0.0
        b["k_flux"] = 40.0 * nF + 36.0 * nB
    b["k_update"] = (40.0 + 40.0 + 8.0 + 160.0 / 3.0) * nC
    b["step"] = b["k_dtmin"] + 3.0 * (b["k_grad"] + b["k_flux"] + b["k_update"])
    return b


def picl_bytes(nP, nCp, nVg=0):
    """Compulsory bytes per stage of the particle coupling (reference
    dataflow): per particle x, v, T and the RK register read and written,
    d_p, rho_p, V_p read, the element id read and written (256 B); per
    particle-occupied cell the interpolation inputs (cv 40, limited
    gradients 120, rhs 40, vol 8) and the 5 projected fields read and
    written (288 B).  With the Gaussian filter, k_gauss reads each
    particle's position and 5 mapped values once (64 B) and each vertex that
    receives a deposit its coordinates and writes its 5 values (64 B, nVg
    vertices); those bytes are this kernel's share, not part of B_step."""
    b = {"k_particles": 256.0 * nP, "k_picl_fields": 208.0 * nCp, "k_deposit": 80.0 * nCp}
    if nVg:
        b["k_gauss"] = 64.0 * nP + 64.0 * nVg
    return b


# ppiclF node n = i + 2j + 4k of a hex is its c2v entry NODE2V[n]
NODE2V = np.array([0, 3, 4, 7, 1, 2, 5, 6])


def deposit_vertices(reg, mesh):
    """Vertices that receive a Gaussian deposit (nonzero PHIP)."""
    nodes = reg.picl_nodes()  # (nCells, 5, 8)
    cell, n = np.nonzero(nodes[:, 0, :])
    return int(np.unique(np.asarray(mesh["c2v"])[cell, NODE2V[n]]).size)


VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4  # wave-level VALU instructions/s: 1024 SIMDs, one per 4 cycles


PROFILES = {("tets64_roe", False): "rflu_tets64_roe_kernels.md", ("tets64_roe", True): "rflu_tets64_roe_fp32_kernels.md",
            ("curtain_spec", False): "picl_curtain_spec_kernels.md"}


def _profile_rows(workload, fp32, ncols):
    name = PROFILES.get((workload, fp32))
    if name is None:
        return []
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return []
    rows = []
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == ncols:
            rows.append(cells)
    return rows


def profiled_traffic(kernel, workload, fp32=False):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by the copy-kernel
    calibration recorded there), or None when the summary lacks the kernel."""
    for cells in _profile_rows(workload, fp32, 5):
        if cells[0].startswith(kernel):
            try:
                return (float(cells[2]) + float(cells[4])) * 1e6
            except ValueError:
                return None
    return None


def profiled_valu(kernel, workload, fp32=False):
    """Wave-level VALU instructions per launch of `kernel` (the SQ_INSTS_VALU
    column of the committed counter summary's shader-sequencer table), or
    None."""
    name = PROFILES.get((workload, fp32))
    if name is None or not os.path.exists(os.path.join(ROOT, "profiles", name)):
        return None
    col = None
    for line in open(os.path.join(ROOT, "profiles", name)):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if "SQ_INSTS_VALU" in cells:
            col = cells.index("SQ_INSTS_VALU")
            continue
        if col is not None and len(cells) > col and cells[0].startswith(kernel):
            try:
                return float(cells[col])
            except ValueError:
                return None
    return None


def build_workload(name):
    from rocflupicl_amd import configs
    # limiterGuard=1: the reference's Venkatakrishnan denominator guard
    # (denom + 1e-12) flips and amplifies gradients when denom is in
    # (-1e-12, 0); on this mesh that drives the run to NaN within a few steps
    # (the oracle and the GPU agree bit for bit on that failure).  The bench
    # therefore adds 1e-12 with the sign of denom; the kernel work is the same.
    if name == "tets64_roe":
        return (configs.shock_tets(n=64, discr=1, limiterGuard=1),
                dict(model="RFLU 2nd-order Roe + Venkat(sign-safe), 64^3x5 tets", n=64))
    if name == "tets64_hllc":
        return (configs.shock_tets(n=64, discr=3, limiterGuard=1),
                dict(model="RFLU 2nd-order HLLC + Venkat(sign-safe), 64^3x5 tets", n=64))
    if name == "sod":
        return configs.sod(), dict(model="RFLU 1st-order AUSM+ Sod 200x4x4 hex", n=None)
    if name == "curtain":
        # limiterGuard=1 for the same reason as above: with the reference's
        # denominator guard the Mach 2 shock on this hex box produces NaN in
        # the first step (oracle and GPU alike).
        mesh, params, ic, parts, picl = configs.curtain(limiterGuard=1)
        return (mesh, params, ic), dict(model="RFLU 2nd-order AUSM+ + ppiclF (IDW-27, QS/AM Parmar + PG, "
                                              "box projection), 400x80x80 hex, 4M particles", n=None,
                                        particles=parts, picl=picl)
    if name == "curtain_spec":
        # the shock-particle-curtain as specified for this benchmark rather than
        # as Rocflu wires ppiclF: trilinear interpolation, Schiller-Naumann
        # quasi-steady drag + pressure gradient + added mass, Gaussian-filter
        # projection of width 2 dx
        mesh, params, ic, parts, picl = configs.curtain(limiterGuard=1)
        picl.update(CURTAIN_SPEC)
        return (mesh, params, ic), dict(model="RFLU 2nd-order AUSM+ + ppiclF (trilinear, QS Schiller-Naumann + "
                                              "AM Parmar + PG, Gaussian projection 2dx), 400x80x80 hex, "
                                              "4M particles", n=None, particles=parts, picl=picl)
    if name == "curtain_deck":
        # the shock-particle-curtain with the particle physics of the AirShock
        # input deck (input_files/AirShock/rectshktb.inp:97-116): Osnes drag,
        # Briney added mass, pressure gradient, collisions level 4 with torques
        # and lift, viscous-unsteady force, Osnes fluctuations, periodic y and
        # z.  Briney's model holds for equal spheres, so the cloud is
        # monodisperse (d_p = 115 um), and, like a particle file's cloud, no
        # two particles touch at the start (the deck's file is not shipped)
        mesh, params, ic, parts, picl = configs.curtain(limiterGuard=1, sigma_ln=0.0, placement="lattice")
        picl.update(configs.AIRSHOCK_PICL)
        return (mesh, params, ic), dict(model="RFLU 2nd-order AUSM+ + ppiclF with the AirShock deck's particle "
                                              "physics (IDW-27, QS Osnes + fluctuations, AM Briney, PG, viscous "
                                              "unsteady, collisions level 4, periodic y/z, box projection), "
                                              "400x80x80 hex, 4M monodisperse particles", n=None, particles=parts,
                                        picl=picl)
    raise SystemExit(f"unknown workload {name}")


# h = 0.0005 is the curtain's spacing
CURTAIN_SPEC = dict(interpMode=1, qsFlag=5, projMode=1, filterWidth=2 * 0.0005, filterAlpha=1e-2)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _oracle_sample(name):
    """A bounded sample of the workload for the CPU oracle: (region,
    particles or 0, description)."""
    from oracle import oracle as orc
    from rocflupicl_amd import configs
    if name.startswith("curtain"):
        deck = name == "curtain_deck"
        mesh, params, ic, (y, pc), picl = configs.curtain(nx=40, ny=8, nz=8, npart=4000, limiterGuard=1,
                                                          sigma_ln=0.0 if deck else 0.2,
                                                          placement="lattice" if deck else "random")
        if name == "curtain_spec":
            picl.update(CURTAIN_SPEC)
        if name == "curtain_deck":
            picl.update(configs.AIRSHOCK_PICL)
        o = orc.Region(mesh, **params)
        o.set_cv(ic(o.cells()[1]))
        o.picl_init(y, pc, **picl)
        return o, 4000, f"{name} at 40x8x8 hexes (2,560 cells) + 4,000 particles, same physics"
    if name.startswith("tets64"):
        mesh, params, ic = configs.shock_tets(n=32, discr=1 if name.endswith("roe") else 3, limiterGuard=1)
        desc = "shock_tets n=32 (163,840 tets), same physics"
    else:
        mesh, params, ic = configs.sod()
        desc = "sod 200x4x4 (3,200 hexes), full workload"
    o = orc.Region(mesh, **params)
    o.set_cv(ic(o.cells()[1]))
    return o, 0, desc


def _time_oracle(name, threads, budget_s):
    from oracle import oracle as orc
    orc.set_threads(threads)
    o, nP, desc = _oracle_sample(name)
    o.step(1)  # warm-up
    steps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < budget_s:
        o.step(1)
        steps += 1
        t = time.perf_counter() - t0
    return o.nCells * steps / t / 1e6, nP * steps / t, steps, t, desc


def cpu_baseline(name, budget_s=10.0):
    """The CPU oracle timed on a bounded sample of the same workload, with its
    independent loops on all host cores (OpenMP; threads = OMP_NUM_THREADS or
    the host's CPU count) and on one core."""
    from oracle import oracle as orc
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    v_all, p_all, st_all, t_all, desc = _time_oracle(name, cores, budget_s)
    v_one, p_one, st_one, t_one, _ = _time_oracle(name, 1, budget_s)
    orc.set_threads(cores)
    out = dict(value=v_all, unit="Mcell-updates/s", cores=cores, kind="port",
               sample=f"{desc}; {st_all} steps in {t_all:.1f} s on {cores} threads, {st_one} steps in "
                      f"{t_one:.1f} s on 1 core; C oracle (OpenMP over its per-cell, per-element and "
                      f"per-particle loops; the face scatters stay serial)",
               single_core=dict(value=v_one, unit="Mcell-updates/s", cores=1),
               cpu_model=cpu_model(), nproc=os.cpu_count())
    if p_all:
        out["particle_updates_per_s"] = p_all
        out["single_core"]["particle_updates_per_s"] = p_one
    return out


def run_batch(args):
    """Batched ensemble: instances grouped by (cell type, gamma) into batch
    handles, all advanced together (each batch on its own stream)."""
    from rocflupicl_amd import configs, rflu
    t0 = time.perf_counter()
    ens = configs.ensemble(args.instances)
    groups = {}
    for key, mesh, params, ic in ens:
        groups.setdefault(key, []).append((mesh, params, ic))
    prec = 1 if args.fp32 else 0
    batches = []
    for key, items in sorted(groups.items()):
        b = rflu.Batch([it[0] for it in items], **dict(items[0][1], fluxPrecision=prec))
        for k, (mesh, params, ic) in enumerate(items):
            b.set_state(k, ic(b.cofg(k)))
        b.autotune(reps=3)  # gradient kernel and stage fusion, both bit-identical choices
        batches.append((key, b, items))
    fused = [(key, b.fused) for key, b, _ in batches]
    cells = sum(b.nCells for _, b, _ in batches)
    setup_s = time.perf_counter() - t0
    print(f"# ensemble: {args.instances} instances, {cells} cells in {len(batches)} batches, setup {setup_s:.1f}s",
          file=sys.stderr, flush=True)
    for _, b, _ in batches:
        b.step_async(args.warmup)
    for _, b, _ in batches:
        b.wait()
    t0 = time.perf_counter()
    for _, b, _ in batches:
        b.step_async(args.steps)
    for _, b, _ in batches:
        b.wait()
    dt_s = time.perf_counter() - t0
    batches_cfg = [dict(cell_type=k[0], gamma=k[1], instances=b.nInst, cells=b.nCells) for k, b, _ in batches]
    err = None
    if args.fp32:
        # the FP64 path twice: from the same initial states (the reference),
        # and from the initial states rounded to FP32 with every operation
        # still FP64, which shows how far each instance's own dynamics carry
        # a perturbation of FP32 size over the same steps
        def rel(a, r):
            return max(np.linalg.norm(a[:, g] - r[:, g]) / np.linalg.norm(r[:, g]) for g in ([0], [1, 2, 3], [4]))
        errs, sens, keys = [], [], []
        for key, b, items in batches:
            ref = rflu.Batch([it[0] for it in items], **items[0][1])
            per = rflu.Batch([it[0] for it in items], **items[0][1])
            for k, (mesh, params, ic) in enumerate(items):
                cv0 = ic(ref.cofg(k))
                ref.set_state(k, cv0)
                per.set_state(k, np.asarray(cv0, dtype=np.float32).astype(np.float64))
            ref.RFLU_TimeStepping(args.warmup + args.steps, use_graph=True)
            per.RFLU_TimeStepping(args.warmup + args.steps, use_graph=True)
            for k in range(b.nInst):
                r = ref.get_state(k)[0]
                errs.append(rel(b.get_state(k)[0], r))
                sens.append(rel(per.get_state(k)[0], r))
                keys.append(key)
            ref.close()
            per.close()
        errs, sens = np.asarray(errs), np.asarray(sens)
        miss = np.flatnonzero(errs > 1e-6)
        err = dict(max=float(np.max(errs)), median=float(np.median(errs)),
                   within_1e6=f"{int(np.sum(errs <= 1e-6))}/{len(errs)}",
                   fp64_from_fp32_rounded_start=dict(max=float(np.max(sens)), median=float(np.median(sens)),
                                                     within_1e6=f"{int(np.sum(sens <= 1e-6))}/{len(sens)}"),
                   above_1e6=[dict(cell_type=keys[i][0], gamma=keys[i][1], fp32_mode=float(errs[i]),
                                   fp64_from_fp32_rounded_start=float(sens[i])) for i in miss[:32]])
    # the plain FP64 path on the same ensemble: a sample of the instances, each
    # a region of its own stepped alone (graph replays), no batching
    plain = None
    if args.plain_sample > 0:
        for _, b, _ in batches:
            b.close()
        pick = ens[::max(1, len(ens) // args.plain_sample)][:args.plain_sample]
        cu, secs, pc = 0.0, 0.0, 0
        for key, mesh, params, ic in pick:
            r = rflu.Region(mesh, **params)
            r.set_state(ic(r.mesh_arrays()["cofg"]))
            r.RFLU_TimeStepping(args.warmup, use_graph=True)
            r.sync()
            t1 = time.perf_counter()
            r.RFLU_TimeStepping(args.steps, use_graph=True)
            r.sync()
            secs += time.perf_counter() - t1
            cu += r.nCells * args.steps
            pc += r.nCells
            r.close()
        pv = cu / secs / 1e6
        plain = dict(value=round(pv, 2), unit="Mcell-updates/s", instances=len(pick), cells=pc,
                     note="each sampled instance alone in the plain FP64 path (one region, graph replays), "
                          "aggregate = sampled cell-updates / summed wall time")
    out = {
        "metric": "Mcell-updates/s/GPU & % HBM roofline",
        "value": round(cells * args.steps / dt_s / 1e6, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt_s / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-flux/f64-accumulate" if args.fp32 else "f64",
        "data": "synthetic (seeded meshes, shock initial states)",
        "config": {"workload": "batch_ensemble", "instances": args.instances, "cells": cells,
                   "batches": batches_cfg,
                   "model": "RFLU 2nd-order AUSM+ + Venkat(sign-safe), per-instance dt", "parallelism": "single GPU"},
        "fp32_rel_l2_vs_fp64": err,
        "fused_stages": [dict(cell_type=k[0], gamma=k[1], fused=f) for k, f in fused],
    }
    if plain:
        out["plain_fp64_path"] = plain
        out["vs_plain_fp64_path"] = round(out["value"] / plain["value"], 3)
    print(json.dumps(out))


def run_serve_sweep(args):
    """The serving trace at each arrival rate of --rates (the same seeded
    requests, Poisson arrival times scaled to the rate): p50/p99 latency and
    aggregate throughput per rate, one JSON line."""
    from rocflupicl_amd import serve
    rows = []
    for rate in [float(r) for r in args.rates.split(",")]:
        reqs = serve.trace(args.requests, rate, seed=args.seed)
        srv = serve.Server(max_active=args.max_active, chunk=args.chunk, builders=args.builders)
        results, wall = srv.run(reqs)
        summ = serve.summarize(results, wall)
        row = dict(rate_per_s=rate, requests=summ["requests"], failed=summ["failed"], wall_s=round(wall, 2),
                   p50_latency_s=round(summ["p50_latency_s"], 3), p99_latency_s=round(summ["p99_latency_s"], 3),
                   mean_latency_s=round(summ["mean_latency_s"], 3),
                   mcell_updates_per_s=round(summ["mcell_updates_per_s"], 2),
                   particle_updates_per_s=round(summ["particle_updates_per_s"], 1))
        print(f"# rate {rate}: {json.dumps(row)}", file=sys.stderr, flush=True)
        rows.append(row)
    out = {
        "metric": "Mcell-updates/s/GPU & % HBM roofline; served p50/p99 latency at fixed accuracy",
        "unit": "Mcell-updates/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f64",
        "data": "synthetic trace (seeded meshes, particle clouds and shock states)",
        "config": {"workload": "serve_trace_sweep", "requests_per_rate": args.requests, "seed": args.seed,
                   "cells": "log-uniform 10k-2M", "particles": "0 or uniform 0-500k (half the requests)",
                   "steps": "uniform 100-2000", "max_active": args.max_active, "chunk": args.chunk,
                   "model": "RFLU 2nd-order AUSM+ (+ ppiclF QS/AM Parmar + PG)", "parallelism": "single GPU"},
        "sweep": rows,
    }
    print(json.dumps(out))


def run_serve(args):
    """Serving trace: Poisson arrivals, continuous batching (rocflupicl_amd.serve)."""
    from rocflupicl_amd import serve
    if args.rates:
        return run_serve_sweep(args)
    reqs = serve.trace(args.requests, args.rate, seed=args.seed)
    srv = serve.Server(max_active=args.max_active, chunk=args.chunk, builders=args.builders)
    results, wall = srv.run(reqs)
    summ = serve.summarize(results, wall)
    out = {
        "metric": "Mcell-updates/s/GPU & % HBM roofline; served p50/p99 latency at fixed accuracy",
        "value": round(summ["mcell_updates_per_s"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic trace (seeded meshes, particle clouds and shock states)",
        "config": {"workload": "serve_trace", "requests": args.requests, "rate_per_s": args.rate,
                   "cells": "log-uniform 10k-2M", "particles": "0 or uniform 0-500k (half the requests)",
                   "steps": "uniform 100-2000", "max_active": args.max_active, "chunk": args.chunk,
                   "model": "RFLU 2nd-order AUSM+ (+ ppiclF QS/AM Parmar + PG)", "parallelism": "single GPU"},
        "latency_s": {"p50": round(summ["p50_latency_s"], 3), "p99": round(summ["p99_latency_s"], 3),
                      "mean": round(summ["mean_latency_s"], 3)},
        "wall_s": round(wall, 2), "failed": summ["failed"],
        "particle_updates_per_s": round(summ["particle_updates_per_s"], 1),
        "requests": [dict(rid=r.rid, cells=r.cells, particles=r.particles, steps=r.steps,
                          arrival=round(r.arrival, 3), started=round(r.started, 3), done=round(r.done, 3),
                          error=r.error) for r in results],
    }
    print(json.dumps(out))


KERNELS = ("k_dtmin", "k_grad", "k_rhog", "k_flux", "k_gather", "k_update", "k_fb_grad", "k_picl_fields",
           "k_contacts", "k_interp", "k_pairs", "k_particles", "k_scan", "k_fill", "k_deposit", "k_gsel", "k_gauss",
           "k_permute", "k_feedback")


def measure(workload, args, fp32=False):
    """Run one workload: warm-up, then the timed steps with per-kernel HIP
    events on the library's stream.  Returns (result dict, region, inputs)."""
    from rocflupicl_amd import rflu
    (mesh, params, ic), info = build_workload(workload)
    t0 = time.perf_counter()
    reg = rflu.Region(mesh, **params, fluxPrecision=1 if fp32 else 0)
    m = reg.mesh_arrays()
    reg.set_state(ic(m["cofg"]))
    nP = 0
    if "particles" in info:
        y, pc = info["particles"]
        reg.picl_init(y, pc, **info["picl"])
        nP = y.shape[0]
    setup_s = time.perf_counter() - t0
    nC, nF, nB = reg.nCells, reg.nFaces, reg.nBFaces
    order = params["spaceOrder"]
    mbar = reg.nStencil / nC if order == 2 else 0.0
    byt = algorithmic_bytes(nC, nF, nB, mbar, order)

    reg.RFLU_TimeStepping(args.warmup, use_graph=args.graph)
    reg.sync()
    reg.set_profiling(not args.graph)
    t0 = time.perf_counter()
    reg.RFLU_TimeStepping(args.steps, use_graph=args.graph)
    reg.sync()
    dt_s = time.perf_counter() - t0
    ms_step = dt_s / args.steps * 1e3
    value = nC * args.steps / dt_s / 1e6

    nCp = nVg = 0
    if nP:
        nCp = int(np.count_nonzero(reg.picl_cells()["phi"]))
        nVg = deposit_vertices(reg, mesh) if info["picl"].get("projMode") == 1 else 0
        pb = picl_bytes(nP, nCp, nVg)
        byt.update(pb)
        # B_step of the reference dataflow: B_dt + 3 (B_stage + B_PICL), B_PICL = 256 Np + 288 Ncp
        byt["step"] += 3.0 * (256.0 * nP + 288.0 * nCp)
    print(f"# workload {workload}: {nC} cells, {nF} faces, {nB} bfaces, mbar {mbar:.2f}, {nP} particles, "
          f"{nCp} occupied cells, {nVg} deposit vertices, setup {setup_s:.1f}s, B_step {byt['step']/1e9:.4f} GB",
          file=sys.stderr, flush=True)

    kern = {}
    if not args.graph:
        for k in KERNELS:
            ms, n = reg.kernel_time(k)
            if n:
                kern[k] = dict(ms=ms, launches=n, total_ms=ms * n)
    reg.set_profiling(False)  # (clears the per-kernel times)
    roof = None
    if kern:
        # the dominant kernel: the most time in the timed region
        dom = max(kern, key=lambda k: kern[k]["total_ms"])
        if dom == "k_grad":  # the gradient routine runs as the tile kernel (default) or k_grad2
            if os.environ.get("ROCFLUPICL_GRAD_VARIANT", "4") != "4":
                pname = "k_grad2"
            elif nP:
                pname = "k_grad_tile<1>"
            else:
                pname = "k_grad_tile_f32" if fp32 else "k_grad_tile_gas"
        elif dom == "k_gauss":
            pname = "k_gauss_w"
        else:
            pname = dom
        traffic = profiled_traffic(pname, workload, fp32)
        share = kern[dom]["total_ms"] / (dt_s * 1e3)
        roof = dict(bound="hbm", kernel=dom, achieved=None, peak=HBM_PEAK_GBS, unit="GB/s", frac=None,
                    traffic=None if traffic is None else int(traffic), algorithmic_bytes=None,
                    ms_per_launch=round(kern[dom]["ms"], 4), share_of_step=round(share, 4))
        if dom in byt:
            ach = byt[dom] / (kern[dom]["ms"] * 1e-3) / 1e9
            roof.update(achieved=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4), algorithmic_bytes=int(byt[dom]))
        vi = profiled_valu(pname, workload, fp32)
        if vi:
            rate = vi / (kern[dom]["ms"] * 1e-3)
            roof["valu"] = dict(bound="VALU issue", achieved=round(rate / 1e9, 1), peak=VALU_ISSUE_PEAK / 1e9,
                                unit="G wave-instructions/s", frac=round(rate / VALU_ISSUE_PEAK, 4),
                                instructions_per_launch=int(vi))
    step_gbs = byt["step"] / (ms_step * 1e-3) / 1e9
    t_end, dt_end = reg.time()
    res = {
        "value": round(value, 2), "ms_per_step": round(ms_step, 4),
        "config": {"workload": workload, "model": info["model"], "cells": nC, "faces": nF, "bfaces": nB,
                   "stencil_mean": round(mbar, 3), "particles": nP or None, "parallelism": "single GPU"},
        "roofline": roof,
        "step_roofline": {"algorithmic_bytes_per_step": int(byt["step"]), "achieved_GBs": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v["ms"], 4) for k, v in kern.items()},
        "particle_updates_per_s": round(nP * args.steps / dt_s, 1) if nP else None,
        "sim_time": t_end,
    }
    return res, reg, (mesh, params, ic)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="curtain_spec")
    ap.add_argument("--secondary", default="tets64_roe",
                    help="a second workload measured the same way and reported under 'secondary' ('none': skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true", help="time hipGraph replays (no per-kernel events)")
    ap.add_argument("--fp32", action="store_true",
                    help="FP32 gradients and face fluxes, FP64 accumulation and update; also reports the "
                         "relative L2 error against the FP64 path after the same steps")
    ap.add_argument("--instances", type=int, default=256, help="batch_ensemble: number of instances")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="batch_ensemble: instances run alone in the plain FP64 path for comparison (0: skip)")
    ap.add_argument("--requests", type=int, default=24, help="serve_trace: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve_trace: arrivals per second")
    ap.add_argument("--rates", default="", help="serve_trace: comma-separated arrival rates to sweep")
    ap.add_argument("--seed", type=int, default=5, help="serve_trace: trace seed")
    ap.add_argument("--max-active", type=int, default=8, help="serve_trace: requests in flight")
    ap.add_argument("--chunk", type=int, default=10, help="serve_trace: steps per launch batch")
    ap.add_argument("--builders", type=int, default=4, help="serve_trace: set-up threads")
    ap.add_argument("--copy-probe-mib", type=int, default=512,
                    help="size of the plain-copy calibration kernel run after the timed region (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    if args.workload == "batch_ensemble":
        return run_batch(args)
    if args.workload == "serve_trace":
        return run_serve(args)
    from rocflupicl_amd import rflu
    res, reg, (mesh, params, ic) = measure(args.workload, args, args.fp32)
    copy_gbs = None
    if args.copy_probe_mib > 0:
        import ctypes
        from rocflupicl_amd import _lib
        g = ctypes.c_double()
        _lib.check(_lib.load().rflu_gpu_bandwidth_probe(args.copy_probe_mib << 20, 5, ctypes.byref(g)))
        copy_gbs = round(g.value, 1)
    fp32_err = None
    if args.fp32:
        ref = rflu.Region(mesh, **params)
        ref.set_state(ic(ref.mesh_arrays()["cofg"]))
        ref.RFLU_TimeStepping(args.warmup + args.steps, use_graph=True)
        a, b = reg.get_state()[0], ref.get_state()[0]
        # relative L2 error of rho, the momentum vector and rho*E (the largest)
        fp32_err = float(max(np.linalg.norm(a[:, g] - b[:, g]) / np.linalg.norm(b[:, g])
                             for g in ([0], [1, 2, 3], [4])))
        ref.close()
    reg.close()
    secondary = None
    if args.secondary != "none" and args.secondary != args.workload:
        sres, sreg, _ = measure(args.secondary, args, False)
        sreg.close()
        secondary = {k: sres[k] for k in ("value", "ms_per_step", "config", "roofline", "step_roofline", "kernels_ms")}
        secondary["unit"] = "Mcell-updates/s"
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.workload)
    out = {
        "metric": "Mcell-updates/s/GPU & % HBM roofline",
        "value": res["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-accumulate" if args.fp32 else "f64",
        "data": "synthetic (seeded mesh, shock initial state and particle cloud)",
        "config": res["config"],
        "roofline": res["roofline"],
        "step_roofline": res["step_roofline"],
        "kernels_ms": res["kernels_ms"],
        "particle_updates_per_s": res["particle_updates_per_s"],
        "copy_probe_GBs": copy_gbs,
        "fp32_rel_l2_vs_fp64": fp32_err,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "sim_time": res["sim_time"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
