"""Benchmark of the MI355X stream-collide step (one JSON line on stdout).

Default workload (value): D3Q19 BGK flow, FP64, 256^3, tau = 0.8, seeded
overlapping spheres (radius U[6,14], seed 202) to porosity 0.40 between
4-plane solid-free buffers, Zou-He rho = 1.001 at z = 0 and 1.000 at
z = 255, periodic in x and y, started from rest.  A step is one LBM time
step (alternately the odd stream-collide-stream and the even collide
half-step, each with its boundary kernels) over the whole lattice, with all
populations resident in HBM.

Also measured in the same run and reported under "extra": D3Q19 flow +
solute at 256^3 in FP64, FP32-shifted (FP32 or FP64 collision) and
FP16-scaled storage (FP32 collision), a device copy bandwidth, and the CPU
oracle on a bounded sample of the default workload (cpu_baseline).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "lbgpu-prime-dev_amd")
METRIC = "D3Q19 MLUPS + % HBM roofline (FP64 vs FP32-shifted); adjoint/forward time"
HBM_PEAK = 8.0e12


def load_pkg():
    name = "lbgpu_prime_dev_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n_fluid, n_nodes, q, nsubs, s):
    """bytes one time step must move: every population of every fluid node
    read and written once, plus the 1-byte solid mask"""
    return nsubs * n_fluid * q * 2 * s + n_nodes


def run_workload(L, mask, kw, steps, warmup, label):
    lat = L.Lattice(mask=mask, **kw)
    info = lat.info()
    lat.init_rest()
    t0 = time.perf_counter()
    lat.run(warmup, write_last=False)
    log(f"[{label}] warmup {warmup} steps {time.perf_counter() - t0:.2f}s")
    lat.profile(True)
    lat.profile_read()
    t0 = time.perf_counter()
    lat.run(steps, write_last=True)
    wall = time.perf_counter() - t0
    prof = lat.profile_read()
    lat.profile(False)
    sums = lat.reduce_sums()
    lat.close()
    log(f"[{label}] {steps} steps {wall:.3f}s kernels {prof}")
    return info, wall, prof, sums


def cpu_baseline(mask, kw, budget_s):
    sys.path.insert(0, ROOT)
    from oracle import pyoracle
    o = pyoracle.Oracle(**kw)
    o.set_solids(mask)
    o.init_rest()
    o.run(2, write_last=False)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.run(2, write_last=False)
        n += 2
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return n, el, pyoracle.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    L = load_pkg()
    copy_gbps = L.copy_bandwidth(2 << 30, 20)
    log(f"device copy bandwidth {copy_gbps:.1f} GB/s")
    n = args.size
    t0 = time.perf_counter()
    mask, phi = L.spheres3d(n, n, n, 202, 6.0, 14.0, 0.40, zbuf=4)
    log(f"geometry {n}^3 porosity {phi:.4f} in {time.perf_counter() - t0:.1f}s")
    base = dict(dims=3, ni=n, nj=n, nk=n, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001, rho_top=1.0)

    kw = dict(base, nsubs=1)
    info, wall, prof, sums = run_workload(L, mask, kw, args.steps, args.warmup, "flow-fp64")
    N, Nf = info["nodes"], info["fluid"]
    ms_step = wall / args.steps * 1e3
    mlups = N * args.steps / wall / 1e6
    mflups = Nf * args.steps / wall / 1e6
    B = algorithmic_bytes(Nf, N, 19, 1, 8)
    odd_ms = prof["odd"][0] / max(prof["odd"][1], 1)
    even_ms = prof["even"][0] / max(prof["even"][1], 1)
    dom = "odd" if odd_ms >= even_ms else "even"
    dom_ms = max(odd_ms, even_ms)
    achieved = B / (dom_ms * 1e-3)
    step_bw = B * args.steps / wall
    result = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"d3q19_flow_fp64_{n}cubed_spheres_phi0.40_zouhe",
            "lattice": [n, n, n], "porosity": round(phi, 5), "fluid_nodes": Nf, "tau": 0.8,
            "rho_in": 1.001, "rho_out": 1.0, "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_half_step<19,1,double,double,{dom == 'odd'}>".replace("True", "true")
            .replace("False", "false"),
            "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
            "bytes_per_launch": B, "avg_launch_ms": round(dom_ms, 5),
        },
        "extra": {
            "mflups": round(mflups, 1),
            "odd_avg_ms": round(odd_ms, 5), "even_avg_ms": round(even_ms, 5),
            "odd_frac": round(B / (odd_ms * 1e-3) / HBM_PEAK, 4) if odd_ms else None,
            "even_frac": round(B / (even_ms * 1e-3) / HBM_PEAK, 4) if even_ms else None,
            "step_frac_wall": round(step_bw / HBM_PEAK, 4),
            "copy_gbps": round(copy_gbps, 1),
            "dominant_frac_of_copy": round(achieved / (copy_gbps * 1e9), 4),
            "mass_sum": sums[0],
        },
    }

    if not args.no_extra:
        variants = [("flow_solute_fp64", 0, 0, 8), ("flow_solute_fp32_shifted", 1, 1, 4),
                    ("flow_solute_fp32_shifted_f64_collision", 1, 0, 4), ("flow_solute_fp16_scaled", 2, 1, 2)]
        ext = {}
        for name, st, ct, s in variants:
            kw2 = dict(base, nsubs=2, zface_solute=1, c_bot=1.0, storage=st, compute=ct,
                       store_shift_c=0.5 if st == 2 else 0.0)
            steps2 = max(args.steps // 2, 10)
            info2, wall2, prof2, _ = run_workload(L, mask, kw2, steps2, 10, name)
            B2 = algorithmic_bytes(info2["fluid"], info2["nodes"], 19, 2, s)
            o_ms = prof2["odd"][0] / max(prof2["odd"][1], 1)
            e_ms = prof2["even"][0] / max(prof2["even"][1], 1)
            ext[name] = {
                "mlups": round(info2["nodes"] * steps2 / wall2 / 1e6, 1),
                "mflups": round(info2["fluid"] * steps2 / wall2 / 1e6, 1),
                "ms_per_step": round(wall2 / steps2 * 1e3, 4),
                "odd_frac": round(B2 / (o_ms * 1e-3) / HBM_PEAK, 4),
                "even_frac": round(B2 / (e_ms * 1e-3) / HBM_PEAK, 4),
                "step_frac_wall": round(B2 * steps2 / wall2 / HBM_PEAK, 4),
            }
        if "flow_solute_fp64" in ext and "flow_solute_fp32_shifted" in ext:
            ext["fp32_over_fp64_speed"] = round(ext["flow_solute_fp32_shifted"]["mflups"]
                                                / ext["flow_solute_fp64"]["mflups"], 3)
        result["extra"].update(ext)

    if not args.no_cpu:
        n_cpu, el, threads = cpu_baseline(mask, dict(base, nsubs=1), args.cpu_seconds)
        result["cpu_baseline"] = {
            "value": round(N * n_cpu / el / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C, OpenMP {threads} threads), same {n}^3 flow FP64 lattice, "
                      f"{n_cpu} steps after 2 warm-up steps, {el:.1f} s",
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
