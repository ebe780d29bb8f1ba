#!/usr/bin/env python
"""Benchmark of the MI355X energy-and-sampling core.

Workload: the seeded synthetic 300-residue chain (same generator and seed as
the tests), torsion Metropolis MC at 300 K with single-torsion +-30 deg
uniform moves, every dE evaluated by the device-resident MC kernel.  One
bench step = one batch: `--chains` independent chains (one workgroup each,
chain c seeded mt19937(seed + c)) advancing `--mc-steps` MC steps each from
the loaded structure.  value = dE evaluations per second over the timed
steps (chains x MC steps / wall time), inputs resident in HBM.

Also reported: the single-chain rate (one chain, latency-bound), the
roofline of the MC kernel (algorithmic FP32 work per launch over its
HIP-event-timed duration), and the host-CPU rate of the FP64 oracle (the
reference CPU path restated, timed in a subprocess on the same structure,
mode and seeds), plus the reference module's own generate_ensemble rate when
its build is present.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "protein-structure-analysis_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

KB = 0.0019872041          # kcal/(mol K)
FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32
# algorithmic FP32 operation census per unit of work (counted from the
# reference formulas): pair energy 40, HCT change of a cross pair 2 x 22,
# SASA fold entry 12, Rodrigues rotation 38 per moved atom, boundary
# dihedral record old + new 2 x 47
CENSUS = {"pair_evals": 40, "cross_pairs": 44, "sasa_entries": 12, "moved_atoms": 38, "dih_records": 94}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--mc-steps", type=int, default=200)
    ap.add_argument("--single-steps", type=int, default=1000)
    ap.add_argument("--residues", type=int, default=300)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--force-reps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the path has no multi-GPU form in this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import protein_physics as PP
    from alma import _native as N
    from alma import synth

    kt = KB * 300.0
    t0 = time.time()
    s = synth.make_chain(args.residues, args.seed)
    topo = PP.BondTopology()
    topo.build(s["resnames"], s["atomnames"], [int(v) for v in s["res_idx"]])
    csr = _csr(topo)
    n = len(s["xyz"])
    log("system: %d residues, %d atoms, %d rotatable bonds (%.1fs)" % (args.residues, n, len(csr["rb_i"]), time.time() - t0))

    ctx = N.Context(0)
    ctx.set_system(s["xyz"], s["charge"], s["rstar"], s["eps"], csr)
    dev = ctx.device_name()

    def batch(k):
        return ctx.mc_run(N.MC_TORSION30, args.mc_steps, n_chains=args.chains, temperature=kt,
                          seed=1000 + k * args.chains)

    for k in range(args.warmup):
        batch(k)
        log("warmup %d done" % k)
    work = {key: 0.0 for key in CENSUS}
    kernel_ms = 0.0
    launches = 0
    rebuild_ms = 0.0
    acc = 0
    t_start = time.perf_counter()
    for k in range(args.steps):
        r = batch(args.warmup + k)
        tm = ctx.mc_timing()
        kernel_ms += tm["mc_kernel_ms"]
        launches += int(tm["mc_launches"])
        rebuild_ms += tm["rebuild_ms"]
        for key in CENSUS:
            work[key] += tm[key]
        acc += sum(x["accepted"] for x in r["stats"])
        log("step %d done" % k)
    elapsed = time.perf_counter() - t_start
    evals = args.steps * args.chains * args.mc_steps
    value = evals / elapsed

    # single chain, same mode (latency-bound)
    t1 = time.perf_counter()
    ctx.mc_run(N.MC_TORSION30, args.single_steps, n_chains=1, temperature=kt, seed=77)
    single = args.single_steps / (time.perf_counter() - t1)
    log("single chain: %.1f evals/s" % single)

    # full evaluation (calculate_potential: list + Born + all terms) against
    # the same evaluation plus the backward kernels (forces and torsion
    # gradients), same coordinates, host round trip included in both
    forces = _forces_rate(ctx, args.force_reps)
    log("energy %.1f evals/s, forces %.1f evals/s" % (forces["energy_evals_per_s"], forces["force_evals_per_s"]))

    flops = sum(CENSUS[k] * work[k] for k in CENSUS)
    achieved = flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    roofline = {"bound": "mfma", "roof": "FP32 vector ALU peak (no MFMA: the pair sums are not a dense contraction)", "achieved": achieved,
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": None,
                "kernel": "k_mc", "kernel_ms_per_launch": kernel_ms / max(launches, 1),
                "algorithmic_flop_per_launch": flops / max(launches, 1),
                "census": {k: work[k] / max(launches, 1) for k in CENSUS},
                "rebuild_ms_total": rebuild_ms, "kernel_ms_total": kernel_ms}

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(s, csr, args, kt)

    out = {
        "metric": "ΔE evals/s + force evals/s, 300-res protein, 1 GPU vs CPU at ≤1e-5 rel err",
        "value": value,
        "unit": "dE evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 reductions and Metropolis dE",
        "data": "synthetic: seeded %d-residue chain (%d atoms, with H), ff14SB parameters" % (args.residues, n),
        "config": {"workload": "torsion Metropolis MC dE, single-torsion +-30 deg moves, 300 K",
                   "residues": args.residues, "atoms": n, "chains": args.chains,
                   "mc_steps_per_chain_per_step": args.mc_steps, "seed": args.seed,
                   "device": dev},
        "single_chain_evals_per_s": single,
        "accept_rate": acc / max(evals, 1),
        "forces": forces,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _forces_rate(ctx, reps):
    ctx.energy()
    ctx.forces()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.energy()
    te = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.forces()
    tf = (time.perf_counter() - t0) / reps
    return {"energy_evals_per_s": 1.0 / te, "force_evals_per_s": 1.0 / tf, "force_over_energy_cost": tf / te,
            "note": "full calculate_potential evaluation vs the same plus dE/dx (Born chain rule) and "
                    "dE/dtheta, one system, host round trip included"}


def _csr(t):
    import numpy as np
    excl = t.excl
    sides = t.rot_bond_sides
    return {
        "excl_off": np.concatenate([[0], np.cumsum([len(v) for v in excl])]).astype(np.int32),
        "excl_idx": np.array([j for v in excl for j in v], np.int32),
        "rb_i": np.array(t.rb_atom_i, np.int32), "rb_j": np.array(t.rb_atom_j, np.int32),
        "rb_kind": np.array(t.rb_kind, np.int32), "rb_scale": np.array(t.rot_bond_scale),
        "side_off": np.concatenate([[0], np.cumsum([len(v) for v in sides])]).astype(np.int32),
        "side_idx": np.array([j for v in sides for j in v], np.int32),
        "dih_a": np.array(t.dih_a, np.int32), "dih_b": np.array(t.dih_b, np.int32),
        "dih_c": np.array(t.dih_c, np.int32), "dih_d": np.array(t.dih_d, np.int32),
        "dih_toff": np.array(t.dih_term_offsets, np.int32), "dih_v2": np.array(t.dih_term_v2),
        "dih_n": np.array(t.dih_term_n, np.int32), "dih_gamma": np.array(t.dih_term_gamma),
        "ss_i": np.array([p[0] for p in t.disulfide_pairs], np.int32),
        "ss_j": np.array([p[1] for p in t.disulfide_pairs], np.int32),
        "cp_phi": np.array([p[0] for p in t.concerted_pairs], np.int32),
        "cp_psi": np.array([p[1] for p in t.concerted_pairs], np.int32),
        "csp_a": np.zeros(0, np.int32), "csp_b": np.zeros(0, np.int32),
    }


def _cpu_baseline(s, csr, args, kt):
    """Host-CPU leg: the oracle in a subprocess, bounded to tens of seconds."""
    import numpy as np
    from alma import _native as N
    cores = min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "system.npz")
        np.savez(path, xyz=s["xyz"], charge=s["charge"], rstar=s["rstar"], eps=s["eps"],
                 resnames=np.array(s["resnames"]), atomnames=np.array(s["atomnames"]),
                 res_idx=s["res_idx"], **{"csr_" + k: v for k, v in csr.items()})
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), path,
               str(args.cpu_steps), str(cores), "1000", repr(kt), str(N.MC_TORSION30)]
        try:
            r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                               text=True, timeout=600)
        except subprocess.TimeoutExpired:
            log("cpu baseline timed out")
            return None
        if r.returncode != 0:
            log("cpu baseline failed:", r.stderr[-2000:])
            return None
        res = json.loads(r.stdout.strip().splitlines()[-1])
    out = {"value": res["port_multi_evals_per_s"], "unit": "dE evals/s", "cores": cores,
           "kind": "port",
           "sample": "%d chains x %d MC steps (single-torsion +-30 deg, 300 K) of the same "
                     "structure and seeds, FP64 oracle restating the reference CPU path, "
                     "OpenMP over chains" % (cores, args.cpu_steps),
           "single_chain_value": res["port_single_evals_per_s"]}
    if "reference_multi_evals_per_s" in res:
        out["reference_generate_ensemble"] = {
            "value": res["reference_multi_evals_per_s"], "single_chain_value": res["reference_single_evals_per_s"],
            "threads": res.get("reference_threads"),
            "note": "reference module built from its own sources, its own move set"}
    return out


if __name__ == "__main__":
    main()
