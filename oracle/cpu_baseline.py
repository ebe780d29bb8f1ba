"""CPU baseline timing for the benchmark (TEST INFRASTRUCTURE: the
benchmark's CPU-baseline leg runs this in a subprocess, so the reference
module's -ffast-math floating-point mode never touches the benchmark process).

  python oracle/cpu_baseline.py <system.npz> <steps> <chains> <seed> <kT> <mode>

Times, on the host cores:
  * the oracle (FP64 restatement of the reference CPU path) in the same MC
    mode and seeds as the GPU run: one chain, then `chains` chains with
    OpenMP over chains (generate_ensemble's parallel pattern);
  * when oracle/_ref holds the reference module built from its own sources:
    its generate_ensemble (reference move set) single chain and `chains`
    chains, on the same structure.
Prints one JSON object.
"""
import glob
import importlib.machinery
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import oracle as O  # noqa: E402  (oracle/oracle.py: this script runs from oracle/)


def main():
    path, steps, chains, seed, kt, mode = sys.argv[1:7]
    steps, chains, seed, kt, mode = int(steps), int(chains), int(seed), float(kt), int(mode)
    z = np.load(path, allow_pickle=False)
    sysd = {k: z[k] for k in ("xyz", "charge", "rstar", "eps")}
    csr = {k[4:]: z[k] for k in z.files if k.startswith("csr_")}
    out = {"threads": O.num_threads(), "steps": steps, "chains": chains}
    t0 = time.perf_counter()
    O.mc_run(sysd, csr, steps, kt, 0.12, seed=seed, mode=mode)
    t1 = time.perf_counter()
    out["port_single_evals_per_s"] = steps / (t1 - t0)
    O.mc_multi(sysd, csr, chains, steps, kt, 0.12, seed=seed, mode=mode)
    t2 = time.perf_counter()
    out["port_multi_evals_per_s"] = chains * steps / (t2 - t1)
    out["port_multi_seconds"] = t2 - t1
    hits = glob.glob(os.path.join(HERE, "_ref", "protein_physics*.so"))
    if hits:
        loader = importlib.machinery.ExtensionFileLoader("protein_physics", hits[0])
        spec = importlib.util.spec_from_loader("protein_physics", loader)
        R = importlib.util.module_from_spec(spec)
        loader.exec_module(R)
        x = sysd["xyz"]
        parts = [R.Particle(float(x[k, 0]), float(x[k, 1]), float(x[k, 2]), float(sysd["charge"][k]),
                            float(sysd["rstar"][k]), float(sysd["eps"][k])) for k in range(len(x))]
        t = R.BondTopology()
        t.build([str(v) for v in z["resnames"]], [str(v) for v in z["atomnames"]],
                [int(v) for v in z["res_idx"]])
        eng = R.PhysicsEngine()
        rs = max(steps // 2, 10)
        t3 = time.perf_counter()
        eng.generate_ensemble(parts, t, 1, rs, kt, 0.12)
        t4 = time.perf_counter()
        eng.generate_ensemble(parts, t, chains, rs, kt, 0.12)
        t5 = time.perf_counter()
        out["reference_single_evals_per_s"] = rs / (t4 - t3)
        out["reference_multi_evals_per_s"] = chains * rs / (t5 - t4)
        out["reference_threads"] = eng.num_threads()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
