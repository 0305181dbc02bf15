"""Benchmark of the MI355X stream-collide step (one JSON line on stdout).

Default workload (value): D3Q19 BGK flow plus the D3Q19 solute distribution,
FP64, 256^3, tau = 0.8, tau_s = 0.75, seeded overlapping spheres (radius
U[6,14], seed 202) to porosity 0.40 between 4-plane solid-free buffers,
Zou-He rho = 1.001 at z = 0 and 1.000 at z = 255, solute c = 1 at z = 0 and
zero gradient at z = 255, periodic in x and y, started from rest.  A step is
one LBM time step (alternately the odd stream-collide-stream and the even
collide half-step, each with its boundary conditions) over the whole
lattice, with all populations resident in HBM.

Also measured in the same run and reported under "extra": the flow-only
lattice of the same geometry (FP64, its permeability after 5,000 steps),
flow + solute at 256^3 in FP32-shifted (FP32 or FP64 collision) and
FP16-scaled storage, the Walsh variant, the discrete adjoint at 128^3, the
192^3 flow + solute and 128^2 D2Q9 configurations, the precision table, the
448^3 slab, a device copy bandwidth, and the CPU oracle on a bounded sample
of the default workload (cpu_baseline).
"""
import argparse
import hashlib
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


def run_workload(L, mask, kw, steps, warmup, label, ns=None, total_steps=0):
    """warmup steps, then `steps` timed steps (kernel times from HIP events
    on the library's stream); with total_steps above warmup + steps the run
    continues, untimed, to total_steps from rest before the sums are taken"""
    lat = L.Lattice(mask=mask, **kw)
    if ns is not None:
        lat.set_ns(ns)
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
    if total_steps > warmup + steps:
        lat.run(total_steps - warmup - steps, write_last=True)
    info["steps_from_rest"] = max(total_steps, warmup + steps)
    sums = lat.reduce_sums()
    lat.close()
    log(f"[{label}] {steps} steps {wall:.3f}s kernels {prof}")
    return info, wall, prof, sums


class KeepSkipped(Exception):
    """the kept-state part of the adjoint workload was not requested"""


def adjoint_timing(L, n=128, steps=2000, window=200, every=100, keep_gb=285.0):
    """discrete-adjoint workload: D3Q19 flow + solute 128^3, spheres to
    porosity 0.40 (seed 404), inlet map U[0.8, 1.2] (seed 4040); time of the
    objective (forward) run and of the adjoint sweep, replaying every segment
    from its checkpoint and with up to keep_gb of HBM of kept states"""
    mask, phi = L.spheres3d(n, n, n, 404, 6.0, 14.0, 0.40, zbuf=4)
    cmap = np.random.default_rng(4040).uniform(0.8, 1.2, (n, n))
    lat = L.Lattice(mask=mask, dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75),
                    gaccel=(0.0, 0.0, 1e-6, 0, 0, 0), zface_pressure=1, rho_bot=1.001, rho_top=1.0,
                    zface_solute=1, c_bot_map=cmap, flux_plane=n - 2)
    def start():
        lat.init_rest()
        lat.time = 0

    # warm-up pair: allocates the checkpoints and the adjoint workspace,
    # which the library keeps for the next run
    start()
    lat.objective(steps, window, every)
    lat.adjoint()
    # a plain forward run of the same length, for reference
    start()
    t0 = time.perf_counter()
    lat.run(steps, write_last=True)
    t_plain = time.perf_counter() - t0
    # the timed pair
    start()
    t0 = time.perf_counter()
    J = lat.objective(steps, window, every)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad, gmap = lat.adjoint()
    t_adj = time.perf_counter() - t0
    # where the adjoint sweep spends its time (events on every launch)
    lat.profile(True)
    lat.profile_read()
    lat.adjoint()
    prof = lat.profile_read()
    lat.profile(False)
    # the same with the last segments' states kept in HBM by the objective
    # run (lbm_objective_keep_states): those segments are not replayed
    kept = {"budget_gb": keep_gb}
    try:
        if keep_gb <= 0:
            raise KeepSkipped()
        lat.keep_states(int(keep_gb * 1e9))
        start()
        lat.objective(steps, window, every)
        lat.adjoint()
        start()
        t0 = time.perf_counter()
        J_k = lat.objective(steps, window, every)
        t_fwd_k = time.perf_counter() - t0
        kept["kept_steps"] = lat.objective_info()["kept_steps"]
        t0 = time.perf_counter()
        grad_k, gmap_k = lat.adjoint()
        t_adj_k = time.perf_counter() - t0
        same = bool(J_k == J and np.array_equal(grad_k, grad) and np.array_equal(gmap_k, gmap))
        kept.update({"forward_s": round(t_fwd_k, 4), "adjoint_s": round(t_adj_k, 4),
                     "adjoint_over_plain_forward": round(t_adj_k / t_plain, 3),
                     "gradient_over_plain_forward": round((t_fwd_k + t_adj_k) / t_plain, 3),
                     "same_gradient_bits": same})
        log(f"[adjoint] kept states ({keep_gb:.0f} GB): forward {t_fwd_k:.3f}s adjoint {t_adj_k:.3f}s, "
            f"same gradient {same}")
    except KeepSkipped:
        kept["skipped"] = "no kept-state budget (--keep-gb 0)"
    except L.LbmError as err:
        # device memory the budget assumed was not there: reported, not fatal
        kept["error"] = str(err)
        log(f"[adjoint] kept states ({keep_gb:.0f} GB) failed: {err}")
    lat.close()
    # the same objective at a plane the solute front reaches within the run
    # (z = 8; at the outlet plane J is at round-off after 2000 steps)
    lat = L.Lattice(mask=mask, dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75),
                    gaccel=(0.0, 0.0, 1e-6, 0, 0, 0), zface_pressure=1, rho_bot=1.001, rho_top=1.0,
                    zface_solute=1, c_bot_map=cmap, flux_plane=8)
    lat.init_rest()
    lat.time = 0
    J8 = lat.objective(steps, window, every)
    grad8, gmap8 = lat.adjoint()
    lat.close()
    reach = {"flux_plane": 8, "J": J8, "dJ_dgz": grad8[2], "dJ_dtau": grad8[3], "dJ_dtau_s": grad8[4],
             "dJ_dcmap_sum": float(gmap8.sum())}
    log(f"[adjoint] plane 8: J {J8:.6e} grad {grad8}")
    replay_ms = prof["odd"][0] + prof["even"][0]
    adj_ms = prof["adjoint"][0]
    log(f"[adjoint] J {J:.6e} forward {t_fwd:.3f}s (plain run {t_plain:.3f}s) adjoint {t_adj:.3f}s "
        f"(replay {replay_ms:.1f} ms, adjoint kernels {adj_ms:.1f} ms, bc {prof['bc'][0]:.1f} ms) grad {grad}")
    return {"lattice": [n, n, n], "porosity": round(phi, 5), "geometry_sha256": sha(mask), "steps": steps,
            "window": window,
            "kept_states": kept,
            "checkpoint_every": every,
            "outlet_plane_at_roundoff": {"flux_plane": n - 2, "J": J, "dJ_dgz": grad[2], "dJ_dtau": grad[3],
                                         "dJ_dtau_s": grad[4], "dJ_dcmap_sum": float(gmap.sum())},
            "reachable_plane": reach,
            "forward_s": round(t_fwd, 4), "plain_forward_s": round(t_plain, 4),
            "adjoint_s": round(t_adj, 4), "adjoint_replay_ms": round(replay_ms, 2),
            "adjoint_kernels_ms": round(adj_ms, 2),
            "adjoint_over_plain_forward": round(t_adj / t_plain, 3),
            "gradient_over_plain_forward": round((t_fwd + t_adj) / t_plain, 3)}


def recorded_traffic(kernel_sig, name="hbm_traffic_d3q19_flow_solute_fp64_256.md"):
    """HBM bytes per launch of a kernel as recorded by the two --pmc runs
    summarised in profiles/<name> (FETCH_SIZE and WRITE_SIZE scaled by the
    same counters on a plain copy kernel)"""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if kernel_sig in line and line.startswith("|"):
                return float(line.strip().strip("|").split("|")[-1]) * 1e9, os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def permeability(sums, n_nodes, tau, rho_in, rho_out, length):
    """Darcy permeability of a pressure-driven run from LBM_REDUCE_SUMS:
    k = nu <u_z> / (dp / L), nu = (tau - 1/2)/3, dp = (rho_in - rho_out)/3,
    <u_z> the superficial average (solid nodes count as zero velocity)"""
    nu = (tau - 0.5) / 3.0
    return nu * (sums[3] / n_nodes) / ((rho_in - rho_out) / 3.0 / length)


def sha(mask):
    return hashlib.sha256(np.ascontiguousarray(mask).tobytes()).hexdigest()


def throughput_c3(L, n=192, steps=1000, cpu_s=0.0):
    """configuration 3 as a throughput workload: D3Q19 flow + solute 192^3,
    porosity 0.45 (seed 303), z-face pressure and solute conditions, in FP64,
    FP32-shifted (FP32 collision) and FP16-scaled storage; with cpu_s > 0 the
    CPU oracle (FP64) timed beside it on the same lattice for about cpu_s"""
    mask, phi = L.spheres3d(n, n, n, 303, 6.0, 14.0, 0.45, zbuf=4)
    base = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=n - 2)
    out = {"lattice": [n, n, n], "porosity": round(phi, 5), "geometry_sha256": sha(mask), "steps": steps}
    for name, st, ct, s in (("fp64", 0, 0, 8), ("fp32_shifted", 1, 1, 4), ("fp16_scaled", 2, 1, 2)):
        kw = dict(base, storage=st, compute=ct)
        info, wall, prof, _ = run_workload(L, mask, kw, steps, 10, f"c3 {name}")
        B = algorithmic_bytes(info["fluid"], info["nodes"], 19, 2, s)
        o_ms = prof["odd"][0] / max(prof["odd"][1], 1)
        e_ms = prof["even"][0] / max(prof["even"][1], 1)
        out[name] = {"mflups": round(info["fluid"] * steps / wall / 1e6, 1),
                     "ms_per_step": round(wall / steps * 1e3, 4),
                     "odd_frac": round(B / (o_ms * 1e-3) / HBM_PEAK, 4),
                     "even_frac": round(B / (e_ms * 1e-3) / HBM_PEAK, 4),
                     "step_frac_wall": round(B * steps / wall / HBM_PEAK, 4)}
    out["fp32_over_fp64_speed"] = round(out["fp32_shifted"]["mflups"] / out["fp64"]["mflups"], 3)
    if cpu_s > 0:
        nf = int((mask.reshape(-1) != 1).sum())
        samples, threads = cpu_baseline(mask, base, cpu_s)
        rates = sorted(nf * k / el / 1e6 for k, el in samples)
        out["cpu_oracle_fp64"] = {"mflups": round(rates[len(rates) // 2], 2), "threads": threads,
                                  "samples": [f"{k} steps in {el:.1f} s" for k, el in samples]}
    return out


def throughput_c1(L, n=128, steps=10000, cpu=False):
    """configuration 1: D2Q9 FP64 128^2, discs to porosity 0.70 (seed 101)
    between walls at y = 0 and 127, tau = 1, g_x = 1e-6, the full 10,000
    steps from rest (the lattice fits in cache; reported as MFLUPS only);
    with cpu the CPU oracle runs the same 10,000 steps"""
    mask, phi = L.discs2d(n, n, 101, 4.0, 10.0, 0.70, walls=True)
    kw = dict(dims=2, ni=n, nj=n, tau=(1.0, 1.0), gaccel=(1e-6, 0, 0, 0, 0, 0))
    lat = L.Lattice(mask=mask, **kw)
    info = lat.info()
    lat.init_rest()
    lat.run(100, write_last=False)
    lat.init_rest()
    t0 = time.perf_counter()
    lat.run(steps, write_last=True)
    wall = time.perf_counter() - t0
    lat.close()
    log(f"[c1] {steps} steps {wall:.3f}s")
    out = {"lattice": [n, n], "porosity": round(phi, 5), "geometry_sha256": sha(mask), "fluid_nodes": info["fluid"],
           "steps": steps, "mflups": round(info["fluid"] * steps / wall / 1e6, 1),
           "us_per_step": round(wall / steps * 1e6, 2)}
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle import pyoracle
        o = pyoracle.Oracle(**kw)
        o.set_solids(mask)
        o.init_rest()
        t0 = time.perf_counter()
        o.run(steps, write_last=True)
        el = time.perf_counter() - t0
        out["cpu_oracle_fp64"] = {"mflups": round(info["fluid"] * steps / el / 1e6, 2),
                                  "threads": pyoracle.num_threads(), "seconds": round(el, 3)}
    return out


def precision_table(L, n=192, steps=4000, plane=16):
    """configuration 3: D3Q19 flow + solute 192^3, porosity 0.45 (seed 303),
    solute c = 1 at the inlet, zero gradient at the outlet, 4000 steps in
    each storage precision; errors against the FP64 run, which the tests
    show equal to the FP64 CPU oracle bit for bit.  The solute flux is taken
    at plane z = `plane`, which the front passes within the run; the
    specified outlet plane (z = n - 2) is reported from the FP64 run, where
    after 4000 steps the flux is still at round-off (the front is 40-60
    planes from the inlet), so its relative errors would mean nothing"""
    mask, phi = L.spheres3d(n, n, n, 303, 6.0, 14.0, 0.45, zbuf=4)
    base = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                rho_top=1.0, zface_solute=1, c_bot=1.0)
    fl = mask.reshape(-1) != 1
    runs = {}
    for name, st, ct in (("fp64", 0, 0), ("fp64_store_fp32_collision", 0, 1), ("fp32_shifted", 1, 1),
                         ("fp32_shifted_fp64_collision", 1, 0), ("fp16_scaled", 2, 1)):
        lat = L.Lattice(mask=mask, storage=st, compute=ct, flux_plane=plane, **base)
        lat.init_rest()
        t0 = time.perf_counter()
        lat.run(steps, write_last=True)
        wall = time.perf_counter() - t0
        rho, ux, uy, uz = lat.mv(0)
        c = lat.mv(1)[0]
        flux = lat.last_flux()[0]
        sums = lat.reduce_sums()
        k = permeability(sums, n ** 3, 0.8, 1.001, 1.0, n - 1)
        # the FP64 GPU reductions: flow mass sum(rho - rho_A), solute mass sum(c)
        runs[name] = (rho[fl] - 1.0, np.stack([ux[fl], uy[fl], uz[fl]]), c[fl], flux, k, wall, sums[0], sums[5])
        lat.close()
        log(f"[precision] {name} {wall:.2f}s flux(z={plane}) {flux:.6e} k {k:.6e}")
    lat = L.Lattice(mask=mask, flux_plane=n - 2, **base)
    lat.init_rest()
    lat.run(steps, write_last=True)
    outlet = lat.last_flux()[0]
    lat.close()
    ref = runs["fp64"]
    out = {"lattice": [n, n, n], "porosity": round(phi, 5), "steps": steps,
           "permeability_fp64": ref[4], "flux_plane": plane, "flux_fp64": ref[3], "flow_mass_fp64": ref[6],
           "solute_mass_fp64": ref[7],
           "outlet_flux_fp64_at_roundoff": {"plane": n - 2, "flux": outlet}}

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    def spread(a, b):
        """99th percentile and mean of the node errors, relative to the
        field's largest FP64 value (the maxima can sit at a few nodes)"""
        d = np.abs(a - b).ravel()
        s = max(np.max(np.abs(b)), 1e-300)
        return [float(np.quantile(d, 0.99) / s), float(d.mean() / s)]
    for name, r in runs.items():
        if name == "fp64":
            continue
        out[name] = {"drho": rel(r[0], ref[0]), "u": rel(r[1], ref[1]), "c": rel(r[2], ref[2]),
                     "drho_p99_mean": spread(r[0], ref[0]), "u_p99_mean": spread(r[1], ref[1]),
                     "c_p99_mean": spread(r[2], ref[2]),
                     "flux": abs(r[3] - ref[3]) / abs(ref[3]), "permeability": abs(r[4] - ref[4]) / abs(ref[4]),
                     "flow_mass": abs(r[6] - ref[6]) / abs(ref[6]), "solute_mass": abs(r[7] - ref[7]) / abs(ref[7]),
                     "seconds": round(r[5], 3)}
    return out


def slab_c5(L, n=448, steps=2000):
    """configuration 5: one GPU's z-slab, D3Q19 flow + solute 448^3,
    FP32-shifted storage, porosity 0.35 (r 8-20, seed 505), pressure driven,
    halo planes through the in-device periodic exchange; rate and drift
    from the FP64 run of the same slab"""
    mask, phi = L.spheres3d(n, n, n, 505, 8.0, 20.0, 0.35, zbuf=4)
    base = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                rho_top=1.0, zface_solute=1, c_bot=1.0, halo_mode=1, flux_plane=n - 2)
    fl = mask.reshape(-1) != 1
    res = {}
    for name, st, ct, s in (("fp32_shifted", 1, 0, 4), ("fp64", 0, 0, 8)):
        lat = L.Lattice(mask=mask, storage=st, compute=ct, **base)
        info = lat.info()
        lat.init_rest()
        lat.profile(True)
        lat.profile_read()
        t0 = time.perf_counter()
        lat.run(steps, write_last=True)
        wall = time.perf_counter() - t0
        prof = lat.profile_read()
        B = algorithmic_bytes(info["fluid"], info["nodes"], 19, 2, s)
        o_ms = prof["odd"][0] / max(prof["odd"][1], 1)
        e_ms = prof["even"][0] / max(prof["even"][1], 1)
        rho, ux, uy, uz = lat.mv(0)
        c = lat.mv(1)[0]
        k = permeability(lat.reduce_sums(), n ** 3, 0.8, 1.001, 1.0, n - 1)
        res[name] = {"mlups": round(info["nodes"] * steps / wall / 1e6, 1),
                     "mflups": round(info["fluid"] * steps / wall / 1e6, 1),
                     "ms_per_step": round(wall / steps * 1e3, 4),
                     "odd_frac": round(B / (o_ms * 1e-3) / HBM_PEAK, 4),
                     "even_frac": round(B / (e_ms * 1e-3) / HBM_PEAK, 4),
                     "halo_ms_per_step": round(prof["halo"][0] / steps, 5),
                     "permeability": k, "outlet_flux": lat.last_flux()[0],
                     "_fields": (rho[fl] - 1.0, np.stack([ux[fl], uy[fl], uz[fl]]), c[fl])}
        lat.close()
        log(f"[slab] {name} {wall:.2f}s {res[name]['mflups']} MFLUPS")

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    a, b = res["fp32_shifted"].pop("_fields"), res["fp64"].pop("_fields")
    res["drift_fp32_vs_fp64"] = {"drho": rel(a[0], b[0]), "u": rel(a[1], b[1]), "c": rel(a[2], b[2]),
                                 "permeability": abs(res["fp32_shifted"]["permeability"] - res["fp64"]["permeability"])
                                 / abs(res["fp64"]["permeability"])}
    res["lattice"] = [n, n, n]
    res["geometry_sha256"] = sha(mask)
    res["porosity"] = round(phi, 5)
    res["steps"] = steps
    return res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mask, kw, budget_s, repeats=3):
    """the CPU oracle on the same lattice: `repeats` timed samples of whole
    step pairs, each about budget_s / repeats long; returns (steps, seconds)
    of each sample and the thread count"""
    sys.path.insert(0, ROOT)
    from oracle import pyoracle
    o = pyoracle.Oracle(**kw)
    o.set_solids(mask)
    o.init_rest()
    o.run(2, write_last=False)
    samples = []
    for r in range(repeats):
        n = 0
        t0 = time.perf_counter()
        while True:
            o.run(2, write_last=False)
            n += 2
            el = time.perf_counter() - t0
            if el >= budget_s / repeats:
                break
        samples.append((n, el))
    return samples, pyoracle.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants-only", action="store_true",
                    help="of the extras, run only the 256^3 flow + solute storage and Walsh variants")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="of the extras, run only the 128^3 discrete-adjoint workload")
    ap.add_argument("--keep-gb", type=float, default=285.0,
                    help="device memory (GB) the adjoint workload's objective run may keep states in; 0: none")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    L = load_pkg()
    copy_gbps = L.copy_bandwidth(2 << 30, 20)
    log(f"device copy bandwidth {copy_gbps:.1f} GB/s")
    n = args.size
    t0 = time.perf_counter()
    mask, phi = L.spheres3d(n, n, n, 202, 6.0, 14.0, 0.40, zbuf=4)
    log(f"geometry {n}^3 porosity {phi:.4f} in {time.perf_counter() - t0:.1f}s")
    base = dict(dims=3, ni=n, nj=n, nk=n, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001, rho_top=1.0)
    head_kw = dict(base, nsubs=2, zface_solute=1, c_bot=1.0)

    # the headline: flow + solute, FP64 storage and collision
    info, wall, prof, sums = run_workload(L, mask, head_kw, args.steps, args.warmup, "flow-solute-fp64")
    N, Nf = info["nodes"], info["fluid"]
    ms_step = wall / args.steps * 1e3
    mlups = N * args.steps / wall / 1e6
    mflups = Nf * args.steps / wall / 1e6
    B = algorithmic_bytes(Nf, N, 19, 2, 8)
    odd_ms = prof["odd"][0] / max(prof["odd"][1], 1)
    even_ms = prof["even"][0] / max(prof["even"][1], 1)
    dom = "odd" if odd_ms >= even_ms else "even"
    dom_ms = max(odd_ms, even_ms)
    achieved = B / (dom_ms * 1e-3)
    step_bw = B * args.steps / wall
    odd_t = "true" if dom == "odd" else "false"
    traffic, traffic_src = recorded_traffic(f"k_half_step<19, 2, double, double, {odd_t}, true, false>")
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
            "workload": f"d3q19_flow_solute_fp64_{n}cubed_spheres_phi0.40_zouhe",
            "lattice": [n, n, n], "porosity": round(phi, 5), "fluid_nodes": Nf, "tau": 0.8, "tau_s": 0.75,
            "rho_in": 1.001, "rho_out": 1.0, "c_in": 1.0, "parallelism": "single-gpu",
            "geometry_sha256": sha(mask),
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_half_step<19,2,double,double,{odd_t},true,false>",
            "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic, "traffic_source": traffic_src,
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
            "flow_mass_sum": sums[0], "solute_mass_sum": sums[5],
        },
    }
    ext = result["extra"]

    # configuration 2: the flow alone on the same lattice, with its
    # permeability after the configuration's 5,000 steps from rest (an
    # untimed continuation when warmup + steps is less)
    kw_flow = dict(base, nsubs=1)
    info2, wall2, prof2, sums2 = run_workload(L, mask, kw_flow, args.steps, args.warmup, "flow-fp64",
                                              total_steps=5000)
    B2 = algorithmic_bytes(Nf, N, 19, 1, 8)
    o_ms = prof2["odd"][0] / max(prof2["odd"][1], 1)
    e_ms = prof2["even"][0] / max(prof2["even"][1], 1)
    ext["flow_fp64_c2"] = {
        "mlups": round(N * args.steps / wall2 / 1e6, 1), "mflups": round(Nf * args.steps / wall2 / 1e6, 1),
        "ms_per_step": round(wall2 / args.steps * 1e3, 4),
        "odd_frac": round(B2 / (o_ms * 1e-3) / HBM_PEAK, 4), "even_frac": round(B2 / (e_ms * 1e-3) / HBM_PEAK, 4),
        "step_frac_wall": round(B2 * args.steps / wall2 / HBM_PEAK, 4),
        "mass_sum": sums2[0],
        # Darcy permeability, L = LZ - 1 between the face planes
        "permeability": permeability(sums2, N, 0.8, 1.001, 1.0, n - 1),
        "permeability_after_steps": info2["steps_from_rest"],
    }

    if args.adjoint_only:
        ext["adjoint_c4"] = adjoint_timing(L, keep_gb=args.keep_gb)
    elif not args.no_extra:
        variants = [("flow_solute_fp32_shifted", 1, 1, 4), ("flow_solute_fp32_shifted_f64_collision", 1, 0, 4),
                    ("flow_solute_fp16_scaled", 2, 1, 2)]
        for name, st, ct, s in variants:
            kw2 = dict(head_kw, storage=st, compute=ct)
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
        # Walsh partial bounce-back as the reference's Walsh build runs it:
        # every node computed, the sphere voxels carrying ns = 0.5; the
        # kernel reads the 8-byte ns of each node besides the populations
        steps2 = max(args.steps // 2, 10)
        info2, wall2, prof2, _ = run_workload(L, np.zeros_like(mask), head_kw, steps2, 10, "walsh",
                                              ns=np.where(mask == 1, 0.5, 0.0))
        B2 = algorithmic_bytes(info2["fluid"], info2["nodes"], 19, 2, 8) + 8 * info2["fluid"]
        o_ms = prof2["odd"][0] / max(prof2["odd"][1], 1)
        e_ms = prof2["even"][0] / max(prof2["even"][1], 1)
        ext["flow_solute_fp64_walsh_all_nodes"] = {
            "mlups": round(info2["nodes"] * steps2 / wall2 / 1e6, 1),
            "ms_per_step": round(wall2 / steps2 * 1e3, 4),
            "odd_frac": round(B2 / (o_ms * 1e-3) / HBM_PEAK, 4),
            "even_frac": round(B2 / (e_ms * 1e-3) / HBM_PEAK, 4),
            "step_frac_wall": round(B2 * steps2 / wall2 / HBM_PEAK, 4),
        }
        if "flow_solute_fp32_shifted" in ext:
            ext["fp32_over_fp64_speed"] = round(ext["flow_solute_fp32_shifted"]["mflups"] / mflups, 3)
        if not args.variants_only:
            ext["adjoint_c4"] = adjoint_timing(L, keep_gb=args.keep_gb)
            ext["throughput_c3"] = throughput_c3(L, cpu_s=0.0 if args.no_cpu else 9.0)
            ext["throughput_c1"] = throughput_c1(L, cpu=not args.no_cpu)
            ext["precision_c3"] = precision_table(L)
            ext["slab_c5"] = slab_c5(L)

    if not args.no_cpu:
        samples, threads = cpu_baseline(mask, head_kw, args.cpu_seconds)
        rates = sorted(N * n_cpu / el / 1e6 for n_cpu, el in samples)
        med = rates[len(rates) // 2]
        result["cpu_baseline"] = {
            "value": round(med, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (the parity build: C, FP64, -O3 -march=x86-64-v3 -ffp-contract=off, OpenMP, "
                      f"{threads} threads on {os.cpu_count()} visible CPUs, "
                      f"{cpu_model()}, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}), "
                      f"same {n}^3 flow + solute FP64 lattice; median of {len(samples)} samples of "
                      + ", ".join(f"{n_cpu} steps in {el:.1f} s" for n_cpu, el in samples),
            "samples_mlups": [round(r, 2) for r in rates],
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
