"""MBD@rsSCS throughput on one MI355X: solves/s at 3N = 12288.

One step = one full E_MBD evaluation (the span pmbd.x times around
MBD_at_rsSCS, MBD_rsSCS.f90:220-222): SCS at 21 imaginary frequencies, C6 and
omega sums, coupled-QHO matrix, all eigenvalues, energy.  The workload is the
4096-atom random periodic cell (38 A cube, C:H:N:O = 40:40:10:10, seeded; SCS
cutoff 40 A, supercell 25 A, CFDM 100 A), so n_SCS = n_MBD = 12288.  The
geometry is resident on the device before the timed region.

Forms timed in the same run: plain FP64 (FP64 tensors, FP64 MFMA Cholesky);
mixed (FP32 tensors, FP32 MFMA Cholesky with FP64 refinement); learned
(FP32 tensors, batched block-PCG over all 21 frequencies with the network's
warm start when weights are loaded, gated on the FP64 residual with Cholesky
fallback).  The eigen stage is FP64 in every form.  The headline value is the
learned form; each form's errors against the FP64 form are reported.

Prints one JSON line on stdout; progress goes to stderr.
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

PEAK = {"hbm_gbs": 8000.0, "fp64_tflops": 78.6, "fp32_tflops": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_solves(eng, mode, steps, warmup):
    for _ in range(warmup):
        eng.run_energy_only(mode)
    t0 = time.perf_counter()
    e = None
    diag = None
    for _ in range(steps):
        e, diag = eng.run_energy_only(mode)
    dt = time.perf_counter() - t0
    return dt, e, diag


def roofline_entry(prof, cls, bound):
    k = prof[cls]
    ms = k["ms"]
    if bound == "hbm":
        achieved = k["bytes"] / (ms * 1e-3) / 1e9
        peak = PEAK["hbm_gbs"]
        unit = "GB/s"
    else:
        achieved = k["flops"] / (ms * 1e-3) / 1e12
        peak = PEAK[bound]
        unit = "TFLOP/s"
    return {"kernel": cls, "bound": "hbm" if bound == "hbm" else "mfma", "achieved": achieved,
            "peak": peak, "unit": unit, "frac": achieved / peak, "traffic": None,
            "launches": k["launches"], "avg_launch_ms": ms / max(k["launches"], 1),
            "algorithmic_per_launch": (k["bytes"] if bound == "hbm" else k["flops"])
            / max(k["launches"], 1)}


def measured_traffic_ratio(kernel):
    """HBM bytes per launch of profiler class `kernel` over its algorithmic
    bytes, from the committed rocprofv3 --pmc summaries (profiles/*traffic*.md:
    FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a 1 GiB
    copy), or None when no summary covers the class."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for line in open(path):
            t = line.strip().strip("`").split()
            if len(t) == 3 and t[0] == f"traffic_ratio[{kernel}]" and t[1] == "=":
                return float(t[2])
    return None


def traffic_probe(args):
    """Workload of the hardware-counter runs: one 1 GiB copy with the plain
    copy kernel (the known-size reference), then one learned-form solve of the
    benchmark cell with every kernel launched eagerly (no graph replay), its
    SCS stage, CFDM build and two-stage eigen reduction at n = 12288.  Meant to
    run under rocprofv3 --pmc; prints progress and, per profiler class, the
    launches and algorithmic bytes of the solve (stderr) for the summary."""
    os.environ["MBD_TRD_GRAPH"] = "0"  # eager launches: one counter record per dispatch
    from mbd_amd import engine, generators
    ms = engine.device_copy(1 << 30)
    log(f"traffic-probe: copy 1 GiB in {ms:.3f} ms")
    geo, st = generators.CONFIGS[args.config]()
    eng = engine.MbdEngine(0)
    eng.set_geometry(geo, st)
    eng.profile_enable(True)
    r = eng.run("learned")
    prof = eng.profile_read()
    eng.close()
    log(f"traffic-probe: {args.config} learned solve E={r.e_mbd:.10f} iters={r.diag['refine_iters']}")
    for k, v in prof.items():
        if v["launches"]:
            log(f"traffic-probe: class {k} launches {v['launches']} "
                f"bytes_per_launch {v['bytes'] / v['launches']:.6e}")


def batched_line(steps, warmup):
    """Batched form: 512 chain molecules (8-256 atoms) per launch, mixed
    precision, default launch shape, hipGraph replays; the FP64 batch and a
    sample of the single-system path alongside."""
    from mbd_amd import engine, generators
    mols, st = generators.chain_molecules()
    eng = engine.MbdEngine(0)
    eng.batch_set(mols, st)
    out = {"workload": "c4_chains_512: 512 non-periodic chains, 8-256 atoms, all 21 points",
           "molecules": len(mols), "atoms": int(sum(m.n_atoms for m in mols))}
    res = {}
    for mode in ("mixed", "fp64"):
        for _ in range(warmup):
            eng.batch_run(mode, full=False)
        t0 = time.perf_counter()
        dev = 0.0
        for _ in range(steps):
            r = eng.batch_run(mode, full=False)
            dev += r["diag"]["t_ms"]
        dt = (time.perf_counter() - t0) / steps
        res[mode] = eng.batch_run(mode)
        out["launch_shape"] = {"index": res[mode]["diag"]["config"],
                               **engine.batch_config(res[mode]["diag"]["config"])}
        out[mode] = {"ms_per_batch": 1e3 * dt, "device_ms_per_batch": dev / steps,
                     "molecules_per_s": len(mols) / dt,
                     "workgroup_ms_by_phase": {k: round(v, 2) for k, v in
                                               res[mode]["diag"]["phase_ms"].items()},
                     "fallbacks": res[mode]["diag"]["fallbacks"],
                     "max_iters": res[mode]["diag"]["max_iters"]}
        log(f"bench: batched {mode} {1e3 * dt:.2f} ms per 512 molecules")
    out["error_vs_fp64"] = {
        "e_mbd_ha_max": float(np.max(np.abs(res["mixed"]["energy"] - res["fp64"]["energy"]))),
        "c6_rel_max": float(np.max(np.abs(res["mixed"]["c6_eff"] / res["fp64"]["c6_eff"] - 1)))}
    eng.close()
    # single-system path on every 8th molecule, scaled to the batch
    single = engine.MbdEngine(0)
    t_single = 0.0
    for g in mols[::8]:
        single.set_geometry(g, st)
        single.run_energy_only("mixed")
        t0 = time.perf_counter()
        single.run_energy_only("mixed")
        t_single += time.perf_counter() - t0
    single.close()
    out["single_system_path_s_per_batch_est"] = 8 * t_single
    out["speedup_vs_single_system_path"] = 8 * t_single / (out["mixed"]["ms_per_batch"] * 1e-3)
    return out


def cpu_baseline(geo, st, sample_atoms=1200):
    """The oracle on the host cores: one complete solve (SCS at all 21
    frequencies, CFDM, eigenvalues, energy: or_mbd_run) of a smaller cell
    from the same generator at the same density, large enough that the
    supercell stays 1 x 1 x 1 as at the full size, scaled by the cube of the
    order ratio; and, as a cross-check at the full size, one SCS frequency
    (tensor build, Cholesky, 3-RHS solve) timed and multiplied by 21."""
    from mbd_amd import generators
    from oracle import oracle as O
    n = 3 * geo.n_atoms
    box = geo.lattice_ang[0][0] * (sample_atoms / geo.n_atoms) ** (1.0 / 3.0)
    small, sst = generators.random_periodic(sample_atoms, box, seed=4096)
    xyz, lat = small.to_bohr()
    sy = O.system_from(small.names, xyz, small.hirshfeld, lat, small.periodic, sst)
    t0 = time.perf_counter()
    r = O.mbd_run(sy)
    t_solve = time.perf_counter() - t0
    ns = 3 * sample_atoms
    scale = (n / ns) ** 3
    s_per_solve = scale * t_solve
    xyz, lat = geo.to_bohr()
    t_f = O.time_scs_frequency(O.system_from(geo.names, xyz, geo.hirshfeld, lat, geo.periodic, st), 10)
    return {"value": 1.0 / s_per_solve, "unit": "solves/s", "s_per_solve": s_per_solve,
            "cores": O.num_threads(), "kind": "port",
            "sample": (f"oracle (C, OpenMP): one complete solve (21-point SCS, CFDM, eigenvalues, "
                       f"energy) of a {sample_atoms}-atom cell of the same generator and density "
                       f"(box {box:.2f} A, supercell {list(r.supercell)}, n={ns}) timed "
                       f"{t_solve:.2f} s and scaled by (n/{ns})^3 = {scale:.1f}; cross-check at "
                       f"n={n}: one SCS frequency {t_f:.2f} s, x21 = {21 * t_f:.1f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3_random_4096")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--headline", default="learned", choices=["mixed", "learned"])
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="workload for rocprofv3 --pmc runs (no JSON line)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding "
                                   "(by frequency) is a design note only", "n_gpus": args.gpus}))
        sys.exit(2)

    from mbd_amd import generators
    from mbd_amd.engine import MbdEngine

    geo, st = generators.CONFIGS[args.config]()
    n = 3 * geo.n_atoms
    log(f"bench: {args.config} N={geo.n_atoms} n={n}")
    eng = MbdEngine(0)
    eng.set_geometry(geo, st)
    from mbd_amd import learned
    have_net = os.path.exists(learned.DEFAULT_WEIGHTS)
    if have_net:
        eng.net_set_weights(learned.load())

    out = {"metric": "MBD@rsSCS solves/s at 3N=12288 (mixed form; FP64 form alongside)",
           "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32+fp64", "data": "synthetic"}

    # plain FP64 form (reference form)
    fp64 = None
    if not args.no_fp64:
        dt64, e64, d64 = timed_solves(eng, "fp64", args.steps, args.warmup)
        r64 = eng.run("fp64")
        fp64 = {"solves_per_s": args.steps / dt64, "s_per_solve": dt64 / args.steps,
                "e_mbd_ha": r64.e_mbd, "info_scs": r64.diag["info_scs"],
                "n_negative": r64.diag["n_negative"]}
        log(f"bench: fp64 {dt64 / args.steps:.3f} s/solve E={r64.e_mbd:.10f}")

    def errors(r, ref):
        return {"e_mbd_ha": abs(r.e_mbd - ref.e_mbd),
                "c6_rel_max": float(np.max(np.abs(r.c6_eff / ref.c6_eff - 1.0))),
                "alpha_scs_rel_max": float(np.max(np.abs(r.coupled_pol / ref.coupled_pol - 1.0))),
                "c6_mol_rel": abs(r.c6_mol / ref.c6_mol - 1.0)}

    forms = {}
    for form in ("mixed", "learned"):
        dt, e, dg = timed_solves(eng, form, args.steps, args.warmup)
        eng.profile_enable(True)
        r = eng.run(form)
        prof = eng.profile_read()
        eng.profile_enable(False)
        forms[form] = {"dt": dt, "r": r, "prof": prof}
        log(f"bench: {form} {dt / args.steps:.3f} s/solve E={r.e_mbd:.10f} diag={r.diag}")
    head = args.headline
    hf = forms[head]
    rm, prof_m = hf["r"], hf["prof"]
    out["value"] = args.steps / hf["dt"]
    out["ms_per_step"] = 1e3 * hf["dt"] / args.steps
    out["s_per_solve"] = hf["dt"] / args.steps
    out["network_warm_start"] = have_net
    out["config"] = {"workload": f"{args.config}: {geo.n_atoms}-atom periodic cell, "
                                 f"n_SCS={n}, n_MBD={n * int(np.prod(rm.diag['supercell']))}",
                     "form": head, "scs_cutoff_A": st.mbd_scs_dip_cutoff,
                     "supercell_cutoff_A": st.mbd_supercell_cutoff,
                     "cfdm_cutoff_A": st.mbd_cfdm_dip_cutoff}
    out["metric"] = f"MBD@rsSCS solves/s at 3N=12288 ({head} form; other forms alongside)"
    for form, f in forms.items():
        rec = {"solves_per_s": args.steps / f["dt"], "s_per_solve": f["dt"] / args.steps,
               "e_mbd_ha": f["r"].e_mbd, "iters": f["r"].diag["refine_iters"],
               "residual": f["r"].diag["residual"],
               "fallback_used": f["r"].diag["fallback_used"],
               "kernels_ms": {k: round(v["ms"], 3) for k, v in f["prof"].items() if v["launches"]}}
        if fp64 is not None:
            rec["speedup_vs_fp64"] = fp64["s_per_solve"] / rec["s_per_solve"]
            rec["error_vs_fp64"] = errors(f["r"], r64)
        out[form] = rec
    if fp64 is not None:
        out["fp64"] = fp64
        out["speedup_vs_fp64"] = out[head]["speedup_vs_fp64"]
        out["error_vs_fp64"] = out[head]["error_vs_fp64"]
    dom = max(prof_m, key=lambda k: prof_m[k]["ms"])
    bound = {"trd_symv": "hbm", "scs_build": "hbm", "cfdm_build": "hbm", "trsv": "hbm",
             "refine_residual": "hbm", "pcg_symv": "hbm", "chol_syrk_gemm": "fp32_tflops",
             "chol_trsm_gemm": "fp32_tflops", "trd_syr2k_gemm": "fp64_tflops",
             "sbr_symm": "fp64_tflops", "sbr_panel_qr": "hbm", "sbr_chase": "hbm"}.get(dom, "hbm")
    out["roofline"] = roofline_entry(prof_m, dom, bound)
    ratio = measured_traffic_ratio(dom)
    if ratio is not None:
        out["roofline"]["traffic"] = ratio * out["roofline"]["algorithmic_per_launch"]
        out["roofline"]["traffic_basis"] = (
            f"counter ratio {ratio} (FETCH_SIZE+WRITE_SIZE over algorithmic bytes, measured on "
            f"the same workload) times the algorithmic bytes per launch of this run")
    others = {}
    for k in ("scs_build", "cfdm_build", "pcg_symv", "refine_residual", "trd_symv", "chol_syrk_gemm",
              "sbr_symm", "trd_syr2k_gemm", "sbr_chase"):
        if k in prof_m and prof_m[k]["launches"] and k != dom:
            others[k] = roofline_entry(prof_m, k, {"chol_syrk_gemm": "fp32_tflops",
                                                   "sbr_symm": "fp64_tflops",
                                                   "trd_syr2k_gemm": "fp64_tflops"}.get(k, "hbm"))
            r = measured_traffic_ratio(k)
            if r is not None and prof_m[k]["bytes"] > 0:
                others[k]["traffic"] = r * prof_m[k]["bytes"] / max(prof_m[k]["launches"], 1)
                others[k]["traffic_basis"] = (
                    f"counter ratio {r} (FETCH_SIZE+WRITE_SIZE over algorithmic bytes, measured on "
                    f"the same workload) times the algorithmic bytes per launch of this run")
    for k in ("scs_build", "cfdm_build"):
        if k in others:
            others[k]["note"] = ("bound by the vector ALU (exp, erf and a division per image term "
                                 "inside the cutoff sphere), not by HBM: the GB/s figure counts "
                                 "the output blocks written once")
    out["roofline_other"] = others
    eng.close()
    if not args.no_batch:
        out["batched"] = batched_line(args.steps, args.warmup)
    if not args.no_cpu:
        log("bench: CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(geo, st)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
