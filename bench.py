"""Throughput benchmark: MPIDForce energy+force evaluations per second on one
MI355X for the 98,304-atom mutual-polarisation water box (water32768: 32,768
waters, 9.932 nm box, 128^3 PME grid, 0.9 nm cutoff, octopoles, anisotropic
polarisability, induced dipoles converged to 1e-6 Debye).

One step = one full evaluation (lab-frame rotation, permanent field, mutual
induced-dipole solve, real-space and reciprocal energy/forces/torques, torque
to force) with positions and forces already resident in device memory.  The
solver starts from mu0 = alpha E every step, as the reference does.

Prints ONE JSON line (rank 0).  Extra fields: the all-FP64 path's rate on the
same box, mixed-vs-FP64 accuracy, per-kernel device times from HIP events on
the evaluator's stream, the dominant kernel's roofline and the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic floating-point operations of the pair formulas, as counted by
# mpidopenmmplugin_amd/csrc/opcount.cpp (make -C mpidopenmmplugin_amd/csrc
# opcount): (both sites polarisable, so the Thole terms are evaluated; a site
# without polarisability).  The field formulas are one-sided: the radial
# factors of a pair, and the field one site's moments produce at the other;
# elec_pair is the quasi-internal-frame interaction without its charge-charge
# channel, which charge_pair evaluates (in FP64) for every pair; elec_pair_bare
# is elec_pair for a pair whose second site is a bare charge, and
# elec_pair_bare_cart its Cartesian form, which the tile kernel runs.
PAIR_FLOPS = {"fixed_field_radial": (73, 47), "fixed_field_from": (105, 105), "induced_field_radial": (36, 23),
              "induced_field_from": (15, 15), "elec_pair": (2106, 2104), "elec_pair_bare": (570, 568),
              "elec_pair_bare_cart": (194, 194), "charge_pair": (19, 19)}


def pair_kernel_flops(cat, pc):
    """Algorithmic FLOPs of one launch of a pair-kernel family: the work the
    result needs, each pair's radial factors once.  pc: pair counts."""
    F = PAIR_FLOPS
    if cat == "fixed_pairs":  # permanent field at every polarisable site
        return (pc["pol"] * (F["fixed_field_radial"][0] + 2 * F["fixed_field_from"][0]) +
                pc["mixed"] * (F["fixed_field_radial"][1] + F["fixed_field_from"][0]))
    if cat == "induced_pairs":  # solver operator: pairs of polarisable sites, both sides
        return pc["pol"] * (F["induced_field_radial"][0] + 2 * F["induced_field_from"][0])
    if cat == "electrostatic_pairs":  # the charge-charge channel of every pair, the rest where moments exist
        general = pc["all"] - pc["cc"] - pc["bare1"]
        return (pc["all"] * F["charge_pair"][0] + pc["pol"] * F["elec_pair"][0] +
                (general - pc["pol"]) * F["elec_pair"][1] + pc["bare1"] * F["elec_pair_bare_cart"][1])
    return None


MI355X_PEAK = {"hbm_GBps": 8000.0, "fp32_TFLOPs": 157.3, "fp64_TFLOPs": 78.6}


def count_pairs(pos, box, cutoff, flat):
    """Unordered pairs within the cutoff that the pair kernels evaluate (the
    scaled covalent partner pairs, which run in the FP64 covalent kernels,
    excluded), split by kind: all, both sites polarisable ("pol"), exactly one
    ("mixed"), both bare charges without polarisability ("cc"), exactly one
    bare charge ("bare1")."""
    from scipy.spatial import cKDTree
    L = np.asarray(box, dtype=np.float64)
    p = np.mod(pos, L)
    p = np.minimum(p, np.nextafter(L, 0))
    t = cKDTree(p, boxsize=L)
    pairs = t.query_pairs(cutoff, output_type="ndarray")
    n = len(p)
    pol = np.asarray(flat["polarity"]).reshape(-1, 3).sum(axis=1) != 0
    moments = np.concatenate([np.asarray(flat[k]).reshape(n, -1) for k in ("dipoles", "quadrupoles", "octopoles")], 1)
    bare = (~pol) & np.all(moments == 0.0, axis=1)
    off = np.asarray(flat["cov_offsets"])
    atoms = np.asarray(flat["cov_atoms"])
    # Covalent12..14 entries (types 0-2) carry scales other than 1
    cnt = off[3::4][:n] - off[0::4][:n]
    if cnt.sum():
        a = np.repeat(np.arange(n), cnt)
        b = np.concatenate([atoms[off[4 * i]:off[4 * i + 3]] for i in range(n)])
        ckey = np.unique(np.minimum(a, b).astype(np.int64) * n + np.maximum(a, b))
        pkey = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int64) * n + np.maximum(pairs[:, 0], pairs[:, 1])
        pairs = pairs[~np.isin(pkey, ckey)]
    pi, pj = pairs[:, 0], pairs[:, 1]
    return {"all": int(len(pairs)), "pol": int((pol[pi] & pol[pj]).sum()), "mixed": int((pol[pi] ^ pol[pj]).sum()),
            "cc": int((bare[pi] & bare[pj]).sum()), "bare1": int((bare[pi] ^ bare[pj]).sum())}


def grid_bytes(cat, cfg, n, launches, mixed):
    """Compulsory HBM bytes of one evaluation's launches of a grid-kernel
    family: inputs read once, outputs written once.  The first launch of
    each family is the permanent-multipole cycle (FP64 in both modes)."""
    g = cfg["grid"]
    G = g ** 3
    Gh = g * g * (g // 2 + 1)
    sz = 4 if mixed else 8
    m = max(0, launches - 1)  # induced-dipole cycles in the working precision
    if cat == "spread":
        return n * (32 + 20 * 8) + G * 8 + m * (n * 7 * sz + G * sz)
    if cat == "fft":
        return 2 * (G * 8 + Gh * 16) + m * 2 * (G * sz + Gh * 2 * sz)
    if cat == "gather":
        # permanent phi (35 per atom, FP64), m-1 field gathers (3 per atom), one final 35-term gather
        return (G * 8 + n * (32 + 35 * 8)) + max(0, m - 1) * (G * sz + n * 7 * sz) + (G * sz + n * 39 * sz)
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(full=False):
    """The FP64 oracle on one host thread: a cell-list restatement of the
    reference's MPIDReferencePmeForce (same formulas and DIIS solver; pairs
    from cell lists instead of the reference's all-pairs loop), timed on the
    216-water and 4096-water boxes, median of 3 evaluations each.  value is
    the 98,304-atom box's rate scaled from the 4096-water box by the atom and
    grid-point ratio (8x), or, with full, one evaluation of the 98,304-atom
    box itself (about two minutes)."""
    from mpidopenmmplugin_amd.waterbox import make_waterbox
    from oracle import oracle

    # the oracle prints a line per solver iteration (a large box takes minutes)
    os.environ["MPID_ORACLE_PROGRESS"] = "1"

    def median_time(name, reps):
        f, pos, box = make_waterbox(name)
        flat = f.flatten()
        times, iters = [], None
        for _ in range(reps):
            t = time.perf_counter()
            r = oracle.evaluate(flat, pos, box)
            times.append(time.perf_counter() - t)
            iters = r["iterations"]
            print(f"[bench] cpu baseline {name}: {times[-1]:.2f} s", flush=True)
        return float(np.median(times)), times, iters

    c0, c0_all, _ = median_time("water216", 3)
    c1, c1_all, it1 = median_time("water4096", 3)
    out = {"cores": 1, "kind": "port", "unit": "evals/s", "cpu_model": _cpu_model(),
           "host_cpus": len(os.sched_getaffinity(0)), "water216_s": c0_all, "water4096_s": c1_all}
    if full:
        c2, c2_all, it2 = median_time("water32768", 1)
        out.update(value=1.0 / c2, water32768_s=c2_all, sample_seconds=c2,
                   sample=(f"the FP64 oracle (a cell-list restatement of the reference's FP64 path with its DIIS "
                           f"solver, not its all-pairs loop) on 1 host thread: one evaluation of the 98,304-atom "
                           f"box took {c2:.1f} s ({it2} DIIS iterations); medians of 3 on the 648- and 12,288-atom "
                           f"boxes {c0:.2f} s and {c1:.1f} s"))
    else:
        out.update(value=(1.0 / c1) / 8.0, sample_seconds=float(sum(c0_all) + sum(c1_all)),
                   sample=(f"the FP64 oracle (a cell-list restatement of the reference's FP64 path with its DIIS "
                           f"solver, not its all-pairs loop) on 1 host thread: median of 3 evaluations of the "
                           f"12,288-atom box {c1:.1f} s ({it1} DIIS iterations; 648-atom box {c0:.2f} s); value = "
                           f"1/(8 x that), the 98,304-atom box having 8x the atoms and grid points"))
    return out


def run(args):
    import torch
    from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel
    from mpidopenmmplugin_amd.waterbox import CONFIGS, make_waterbox

    t0 = time.perf_counter()
    force, pos, box = make_waterbox(args.workload)
    n = force.getNumMultipoles()
    cfg = CONFIGS[args.workload]
    dev = torch.device("cuda:0")
    d_pos = torch.tensor(pos, dtype=torch.float64, device=dev).contiguous()
    d_f = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0

    def make(precision, eps=None):
        k = HipCalcMPIDForceKernel(precision, single_stream=args.single_stream)
        f = force
        if eps is not None:
            f, _, _ = make_waterbox(args.workload, target_eps=eps)
        k.initialize(n, f)
        return k

    def timed_rate(k, steps, warmup):
        for _ in range(warmup):
            k.execute_device(d_pos.data_ptr(), box, d_f.data_ptr())
        torch.cuda.synchronize()
        k.profile(reset=True)
        t = time.perf_counter()
        e = None
        for _ in range(steps):
            e = k.execute_device(d_pos.data_ptr(), box, d_f.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return dt / steps, e

    km = make(args.precision)
    # live HIP events on the dominant kernel (the real-space electrostatics,
    # outside the solver's graph replays) over the timed evaluations
    km.setProfileFamilies(["electrostatic_pairs"])
    ms_step, energy = timed_rate(km, args.steps, args.warmup)
    live_ms, live_n = km.profile(reset=True)["electrostatic_pairs"]
    km.setProfiling(False)
    st = km.stats()
    print(f"[bench] {args.workload} {args.precision}: {1000 * ms_step:.2f} ms/step, solver {st}", file=sys.stderr,
          flush=True)

    # per-kernel device time (HIP events on the evaluator's stream), separate pass
    def kernel_times(k, nprof):
        k.setProfiling(True)
        k.profile(reset=True)
        for _ in range(nprof):
            k.execute_device(d_pos.data_ptr(), box, d_f.data_ptr())
        prof = k.profile(reset=True)
        k.setProfiling(False)
        return {name: {"ms_per_step": v[0] / nprof, "launches_per_step": v[1] / nprof} for name, v in prof.items()}

    kernels = kernel_times(km, max(1, min(3, args.steps)))

    result = {}
    # FP64 path on the same box and its accuracy reference
    fp64 = None
    if not args.skip_fp64:
        kd = make("double")
        ms64, e64 = timed_rate(kd, max(1, args.steps // 2), 1)
        kt64 = kernel_times(kd, 2)
        fp64 = {"ms_per_step": 1000 * ms64, "evals_per_s": 1.0 / ms64, "stats": kd.stats(),
                "kernels": {k: round(v["ms_per_step"], 4) for k, v in kt64.items()}, "kernel_times": kt64}
        print(f"[bench] {args.workload} double: {1000 * ms64:.2f} ms/step", file=sys.stderr, flush=True)
        del kd
        # accuracy of the timed path against the FP64 path converged to 1e-10 Debye
        kref = make("double", eps=1e-10)
        e_ref, f_ref = kref.execute(pos, box)
        mu_ref = kref.getInducedDipoles(pos, box)
        del kref
        kchk = make(args.precision)
        e_m, f_m = kchk.execute(pos, box)
        mu_m = kchk.getInducedDipoles(pos, box)
        del kchk
        result["accuracy_vs_fp64"] = {
            "energy_rel": abs(e_m - e_ref) / abs(e_ref),
            "force_rel_rms": float(np.sqrt(((f_m - f_ref) ** 2).sum() / (f_ref ** 2).sum())),
            "induced_dipole_rel_rms": float(np.sqrt(((mu_m - mu_ref) ** 2).sum() / (mu_ref ** 2).sum())),
            "reference": "FP64 GPU path converged to 1e-10 Debye",
        }

    # Extrapolated (OPT) polarization on the same box: both precisions timed
    # like the headline path, and the mixed path's accuracy against FP64
    if not args.skip_fp64:
        fx, _, _ = make_waterbox(args.workload, polarization=2)
        ext, outs = {}, {}
        for prec in ("mixed", "double"):
            kx = HipCalcMPIDForceKernel(prec)
            kx.initialize(n, fx)
            msx, _ = timed_rate(kx, max(1, args.steps // 2), 1)
            outs[prec] = kx.execute(pos, box)
            ext[prec + "_ms_per_step"] = 1000 * msx
            del kx
        (e_x, f_x), (e_xd, f_xd) = outs["mixed"], outs["double"]
        ext["speedup_mixed_vs_fp64"] = ext["double_ms_per_step"] / ext["mixed_ms_per_step"]
        ext["accuracy_mixed_vs_fp64"] = {
            "energy_rel": abs(e_x - e_xd) / abs(e_xd),
            "force_rel_rms": float(np.sqrt(((f_x - f_xd) ** 2).sum() / (f_xd ** 2).sum())),
        }
        print(f"[bench] {args.workload} extrapolated: mixed {ext['mixed_ms_per_step']:.2f} ms/step, "
              f"double {ext['double_ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
        result["extrapolated"] = ext

    # roofline of every kernel family; the dominant one is reported
    flat = force.flatten()
    pc = count_pairs(pos, box, cfg["cutoff"], flat)
    mixed = args.precision == "mixed"
    rooflines = {}
    traffic = profiled_traffic(args.workload, args.precision)
    for cat, kv in kernels.items():
        ms, nl = kv["ms_per_step"], kv["launches_per_step"]
        if ms <= 0 or nl <= 0:
            continue
        avg_ms = ms / nl
        flops = pair_kernel_flops(cat, pc)  # per launch
        if flops is not None:
            achieved = flops / (avg_ms * 1e-3) / 1e12
            peak = MI355X_PEAK["fp32_TFLOPs"] if mixed else MI355X_PEAK["fp64_TFLOPs"]
            rooflines[cat] = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                              "frac": achieved / peak, "traffic": traffic[cat] / nl if cat in traffic else None,
                              "avg_launch_ms": avg_ms,
                              "flops_per_launch": flops,
                              "note": "vector ALU arithmetic (no matrix instructions); peak is MI355X's published "
                                      "dense FP32/FP64 rate"}
        else:
            b = grid_bytes(cat, cfg, n, nl, mixed)
            if b is None:
                continue
            achieved = b / (ms * 1e-3) / 1e9
            rooflines[cat] = {"bound": "hbm", "achieved": achieved, "peak": MI355X_PEAK["hbm_GBps"], "unit": "GB/s",
                              "frac": achieved / MI355X_PEAK["hbm_GBps"],
                              "traffic": traffic[cat] / nl if cat in traffic else None, "avg_launch_ms": avg_ms,
                              "bytes_per_step": b}
    # the dominant single kernel: the electrostatics, timed live in the timed
    # region; HBM traffic per launch from the committed rocprofv3 counters
    e_flops = pair_kernel_flops("electrostatic_pairs", pc)
    e_avg = live_ms / max(1, live_n)
    e_peak = MI355X_PEAK["fp32_TFLOPs"] if mixed else MI355X_PEAK["fp64_TFLOPs"]
    e_ach = e_flops / (e_avg * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": e_ach, "peak": e_peak, "unit": "TFLOP/s", "frac": e_ach / e_peak,
            "traffic": traffic.get("electrostatic_pairs"), "kernel": "k_electrostatics_" +
            ("f32" if mixed else "f64"), "avg_launch_ms": e_avg, "launches_timed": live_n,
            "flops_per_launch": e_flops,
            "note": "vector ALU arithmetic, no matrix instructions; peak is MI355X's published dense FP32/FP64 "
                    "rate; traffic in bytes per launch" +
                    ("; the mixed path's FP64 check of the induced dipoles runs on another stream during this "
                     "kernel and shares its CUs, and the live duration includes that" if mixed else "")}

    # the hand-written FFT convolution against rocFFT (through torch.fft,
    # which calls hipFFT/rocFFT on ROCm) on the same grid: R2C, a real
    # influence-function multiply, C2R, in the working precision
    fft_check = None
    if "fft" in kernels and kernels["fft"]["launches_per_step"] > 0:
        g = cfg["grid"]

        def rocfft_ms(rdt):
            q = torch.randn((g, g, g), dtype=rdt, device=dev)
            infl = torch.rand((g, g, g // 2 + 1), dtype=rdt, device=dev)
            for _ in range(3):
                torch.fft.irfftn(torch.fft.rfftn(q) * infl, s=(g, g, g))
            torch.cuda.synchronize()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 20
            ev0.record()
            for _ in range(reps):
                torch.fft.irfftn(torch.fft.rfftn(q) * infl, s=(g, g, g))
            ev1.record()
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1) / reps

        fft_check = {"grid": [g, g, g], "rocfft_fp32_ms_per_convolution": rocfft_ms(torch.float32),
                     "rocfft_fp64_ms_per_convolution": rocfft_ms(torch.float64),
                     "note": "rocFFT through torch.fft (R2C, elementwise multiply by a real array, C2R; torch's "
                             "own kernel for the multiply), HIP events on torch's stream; the hand-written passes "
                             "apply the influence function inside the transform.  The hand-written times are "
                             "per-launch averages of the fft family: FP32 in the mixed solver iterations (the "
                             "mixed path's one FP64 permanent-multipole convolution included), FP64 in the "
                             "FP64 path"}
        kf = kernels["fft"]
        if mixed:
            fft_check["hand_written_mixed_ms_per_convolution"] = kf["ms_per_step"] / kf["launches_per_step"]
        if fp64 and "fft" in fp64["kernel_times"] and fp64["kernel_times"]["fft"]["launches_per_step"] > 0:
            kf = fp64["kernel_times"]["fft"]
            fft_check["hand_written_fp64_ms_per_convolution"] = kf["ms_per_step"] / kf["launches_per_step"]
        print(f"[bench] fft check {g}^3: {fft_check}", file=sys.stderr, flush=True)

    # host-buffer rate (positions and forces crossing PCIe each evaluation)
    t = time.perf_counter()
    for _ in range(3):
        km.execute(pos, box)
    pcie_ms = 1000 * (time.perf_counter() - t) / 3

    out = {
        "metric": "energy+force evals/s, 98k-atom mutual-pol water box; force RMS err vs FP64",
        "value": 1.0 / ms_step,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair/reciprocal arithmetic, fp64 solver reductions" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded water box generator)",
        "config": {"workload": args.workload, "atoms": n, "box_nm": float(box[0]), "pme_grid": cfg["grid"],
                   "cutoff_nm": cfg["cutoff"], "bspline_order": 6, "polarization": "mutual",
                   "target_epsilon_debye": 1e-6, "precision": args.precision, "pairs": pc,
                   "streams": 1 if args.single_stream else 3},
        "energy_kJ_mol": energy,
        "solver": st,
        "kernels": kernels,
        "roofline": roof,
        "rooflines": rooflines,
        "fft_check": fft_check,
        "host_buffer_ms_per_step": pcie_ms,
        "setup_s": setup_s,
    }
    if fp64:
        fp64.pop("kernel_times", None)
        out["fp64_path"] = {"evals_per_s": fp64["evals_per_s"], "ms_per_step": fp64["ms_per_step"],
                            "kernels_ms_per_step": fp64["kernels"]}
        out["speedup_vs_fp64"] = fp64["ms_per_step"] / (1000 * ms_step)
    out.update(result)
    if not args.skip_gradients:
        out["gradients"] = adjoint_leg(args)
    if args.batch_systems > 0:
        out["batch"] = batch_leg(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(full=args.cpu_full)
    return out


def batch_leg(args):
    """The batched workload: independent periodic water boxes of U(32, 1024)
    waters, one single-stream context each, evaluated together from a pool
    of host threads; also run one system after another (one thread) for
    comparison."""
    from mpidopenmmplugin_amd.batch import BatchEvaluator
    from mpidopenmmplugin_amd.waterbox import make_batch

    t0 = time.perf_counter()
    systems = make_batch(count=args.batch_systems)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ev = BatchEvaluator(systems, precision=args.precision, threads=args.batch_threads)
    init_s = time.perf_counter() - t0
    atoms = ev.num_atoms
    grids = sorted({int(f.getPMEParameters()[1]) for f, _, _ in systems})
    del systems
    print(f"[bench] batch: {len(ev)} systems, {atoms} atoms, generated in {gen_s:.1f} s, contexts in {init_s:.1f} s",
          file=sys.stderr, flush=True)
    ev.execute()  # first evaluation builds lists and solver graphs
    reps = max(1, args.batch_reps)
    t = time.perf_counter()
    for _ in range(reps):
        res = ev.execute()
    dt = (time.perf_counter() - t) / reps
    e_sum = float(sum(e for e, _ in res))
    t = time.perf_counter()
    ev.execute(sequential=True)
    dt_seq = time.perf_counter() - t
    ev.close()
    print(f"[bench] batch: {1000 * dt:.1f} ms per batch with {args.batch_threads} threads, "
          f"{1000 * dt_seq:.1f} ms one system after another", file=sys.stderr, flush=True)
    return {"workload": "independent water boxes, U(32, 1024) waters each (seeded), 1.0 g/cm^3, "
                        "PME grid ~0.08 nm (2,3,5-smooth), cutoff min(0.8, L/2 - 1e-3), mutual to 1e-6 Debye",
            "systems": args.batch_systems, "atoms": atoms, "pme_grids": grids, "precision": args.precision,
            "spread_tilings": ("autotuned (mpidopenmmplugin_amd/spread_tilings.json)"
                               if os.path.exists(os.path.join(ROOT, "mpidopenmmplugin_amd", "spread_tilings.json"))
                               else "the library's built-in choice"),
            "host_threads": args.batch_threads, "ms_per_batch": 1000 * dt, "systems_per_s": args.batch_systems / dt,
            "atoms_per_s": atoms / dt, "ms_per_batch_sequential": 1000 * dt_seq,
            "speedup_vs_sequential": dt_seq / dt, "energy_sum_kJ_mol": e_sum, "reps": reps,
            "generate_s": gen_s, "contexts_s": init_s}


def adjoint_leg(args):
    """Energy gradients of the 12,288-atom box (water4096: octopoles,
    anisotropic polarisability, 64^3 grid, mutual to 1e-6 Debye) with respect
    to the positions and the per-particle parameters, through
    mpid_energy_gradients (forward solve, adjoint solve, gradient kernels),
    timed against a plain evaluation on the same host buffers; the same
    through the torch.autograd.Function with the per-type parameterisation
    (float32 leaves); and the per-type gradients against the committed
    central differences of the FP64 oracle (tests/golden).  Also the
    gradient/forward ratio on the 98,304-atom box."""
    import torch
    from mpidopenmmplugin_amd.autograd import MPIDEnergy, expand_water_parameters, water_type_parameters
    from mpidopenmmplugin_amd.kernel import HipCalcMPIDForceKernel
    from mpidopenmmplugin_amd.waterbox import make_waterbox

    def ratio(workload, prec, reps):
        force, pos, box = make_waterbox(workload)
        k = HipCalcMPIDForceKernel(prec)
        k.initialize(force.getNumMultipoles(), force)
        for _ in range(2):
            k.execute(pos, box)
            g = k.energyGradients(pos, box)
        t = time.perf_counter()
        for _ in range(reps):
            k.execute(pos, box)
        fwd = (time.perf_counter() - t) / reps
        t = time.perf_counter()
        for _ in range(reps):
            g = k.energyGradients(pos, box)
        grd = (time.perf_counter() - t) / reps
        return {"forward_ms": 1000 * fwd, "gradients_ms": 1000 * grd, "gradients_over_forward": grd / fwd,
                "adjoint_iterations": g["adjoint_iterations"]}

    out = {"workload": "water4096: 12,288 atoms, 64^3 grid, mutual to 1e-6 Debye; gradients of the energy with "
                       "respect to positions and every per-particle parameter (host buffers both ways)"}
    for prec in ("mixed", "double"):
        out[prec] = ratio("water4096", prec, args.steps)
        print(f"[bench] gradients water4096 {prec}: {out[prec]}", file=sys.stderr, flush=True)
    out["water32768_mixed"] = ratio("water32768", "mixed", max(2, args.steps // 2))
    print(f"[bench] gradients water32768 mixed: {out['water32768_mixed']}", file=sys.stderr, flush=True)

    # the autograd path with per-type parameters
    force, pos, box = make_waterbox("water4096")
    nw = force.getNumMultipoles() // 3
    model = MPIDEnergy(force, box, "mixed")
    p = water_type_parameters(force, dtype=torch.float32)
    x = torch.tensor(pos, dtype=torch.float32, requires_grad=True)

    def step():
        for v in list(p.values()) + [x]:
            v.grad = None
        e = model(x, *expand_water_parameters(p, nw), p["default_thole"])
        e.backward()
        return e

    step()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    out["autograd_ms"] = 1000 * (time.perf_counter() - t) / args.steps
    got = np.concatenate([np.atleast_1d(p[k].grad.double().numpy()) for k in
                          ("q_O", "q_H", "d_O", "Q_O", "O_O", "alpha_O", "thole_O", "thole_H", "default_thole")])
    fx = os.path.join(ROOT, "tests", "golden", "water4096_gradients_fd.npz")
    if os.path.exists(fx):
        ref = np.load(fx)
        fd = ref["gradient"]
        big = np.abs(fd) > 1e-9 * np.abs(fd).max()
        rel = np.abs(got - fd)[big] / np.abs(fd)[big]
        out["per_type_vs_fd"] = {"max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel)),
                                 "params_compared": int(big.sum()), "params": int(fd.size),
                                 "reference": "central differences of the FP64 oracle's energy re-converged to "
                                              "1e-10 Debye (steps 1e-2 relative for the charges and multipoles, "
                                              "in which the energy is quadratic; four-point, 1e-4, for the "
                                              "polarisabilities and Thole widths); float32 parameter leaves"}
    print(f"[bench] gradients autograd: {out['autograd_ms']:.2f} ms, {out.get('per_type_vs_fd')}", file=sys.stderr,
          flush=True)
    return out


def profiled_traffic(workload, precision):
    """HBM bytes per evaluation of each kernel family from the committed
    rocprofv3 summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the
    copy-kernel calibration recorded there, times the dispatches per
    evaluation): {family: bytes}, empty if there is none."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}_{precision}.md")
    out = {}
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_evaluation["):
                fam = line[line.index("[") + 1:line.index("]")]
                out[fam] = float(line.split(":", 1)[1])
    except OSError:
        pass
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of mib MiB (read mib, write mib), the
    reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(mib * 262144, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": mib * 1048576, "copy_bytes_written": mib * 1048576, "copies": 2}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "double"])
    ap.add_argument("--workload", default="water32768", choices=["water216", "water4096", "water32768"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-gradients", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU baseline on the 98,304-atom box itself (about two minutes)")
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="systems in the batched-workload leg (0 skips it)")
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--single-stream", action="store_true",
                    help="issue every kernel of the evaluation on one stream (serialised kernel times)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build evaluates one system on one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    out = run(args)
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
