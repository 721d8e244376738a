"""Benchmark: forward photon transport on one MI355X.

Workload (one "step"): the heterogeneous 60^3 cube with a radius-15 sphere,
1e8 photons, 10 time gates over 0-5 ns, four 1 mm disk detectors with partial
path lengths and exit positions saved, float32 fluence output.

value = photons per millisecond of the transport kernel (HIP events on the
kernel's stream; inputs resident in HBM, fluence normalisation included in the
following finalize kernel).  The JSON line also carries:
  roofline      VALU lane-instruction roofline of the transport kernel (frozen
                per-event operation model x oracle event counts, see
                ROOFLINE_MODEL below) and the float64-atomic deposit rate
  cpu_baseline  the CPU oracle (C restatement of the reference kernel, OpenMP)
                timed on this host on a bounded sample of the same workload

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons N]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/ms (fwd, frac of roofline); adjoint gradient/s vs finite-diff at ≤2% err"

# Frozen per-event operation model (single-precision/integer VALU lane
# instructions of the reference algorithm, transcendentals counted once here and
# weighted x4 below because they issue at quarter rate):
#   step      one main-loop iteration: property fetch, 3-way voxel DDA with 3
#             divides, step clipping, move, Beer-Lambert exp, time update,
#             voxel index and boundary tests       ~65 ops, 5 transcendental
#   deposit   voxel-exit fluence weight and gate  ~17 ops, 1 transcendental (rcp)
#   scatter   scattering length (log), azimuth sincos, Henyey-Greenstein
#             inversion, acos/sin, rotatevector with 3 rsqrt, 3 RNG draws
#                                                  ~135 ops, 10 transcendental
#   boundary  Fresnel coefficient, RNG draw, reflect/transmit ~55 ops, 2 transcendental
ROOFLINE_MODEL = {
    "step": (65, 5),
    "deposit": (17, 1),
    "scatter": (135, 10),
    "boundary": (55, 2),
}
# MI355X published FP32 vector peak 157.3 TFLOP/s counts 2 flops per FMA and
# 2-wide packed FP32: lane-instruction issue peak = 157.3 / 4 = 39.3 T/s.
VALU_PEAK_TOPS = 157.3 / 4.0


def ops_per_photon(counters):
    n = counters["nlaunch"]
    tot = 0.0
    for key, cnt in (("step", counters["nstep"]), ("deposit", counters["ndep"]), ("scatter", counters["nscat"]),
                     ("boundary", counters["nbnd"])):
        ops, trans = ROOFLINE_MODEL[key]
        tot += cnt * (ops + 3 * trans)
    return tot / n, counters["ndep"] / n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--mathmode", type=int, default=0, help="0: hardware transcendentals, 1: bit-reproducible")
    ap.add_argument("--cpu-photons", type=float, default=3e5, help="photons in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import mcxcl_amd
    from mcxcl_amd import workloads

    nphoton = int(args.photons)
    cfg = workloads.sphere_cube_workload(nphoton=nphoton, mathmode=args.mathmode)

    for _ in range(args.warmup):
        mcxcl_amd.run(cfg)

    kernel_ms = []
    wall = []
    absorbed = []
    ndet = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = mcxcl_amd.run(cfg)
        wall.append((time.perf_counter() - t0) * 1e3)
        kernel_ms.append(out["stat"]["kernelms"])
        absorbed.append(out["stat"]["energyabs"] / out["stat"]["energytot"])
        ndet.append(out["detp"].shape[1] if "detp" in out else 0)

    ms_per_step = float(np.mean(kernel_ms))
    value = nphoton / ms_per_step

    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "photons/ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "60^3 label cube + r=15 sphere (label 2), 1e8 photons, 10 gates 0-5 ns, "
                        "4 disk detectors r=1 mm, savedetflag DPX, seed 1648335518",
            "photons": nphoton,
            "gates": 10,
            "mathmode": "fast" if args.mathmode == 0 else "parity",
            "parallelism": "single GPU, 2560 persistent lanes per CU",
        },
        "wall_ms_per_step": round(float(np.mean(wall)), 3),
        "absorbed_fraction": round(float(np.mean(absorbed)), 6),
        "detected_per_step": int(np.mean(ndet)),
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle

        ncpu = os.cpu_count() or 1
        cpu_n = int(args.cpu_photons)
        ccfg = workloads.sphere_cube_workload(nphoton=cpu_n, autopilot=0, nthread=65536, nblocksize=64)
        ref = oracle.run(ccfg, nthread=65536, nomp=ncpu)
        cpu_rate = cpu_n / (ref["seconds"] * 1e3)
        opp, depp = ops_per_photon(ref["counters"])
        achieved_tops = value * 1e3 * opp / 1e12
        res["roofline"] = {
            "bound": "valu",
            "achieved": round(achieved_tops, 3),
            "peak": round(VALU_PEAK_TOPS, 3),
            "unit": "Tops/s",
            "frac": round(achieved_tops / VALU_PEAK_TOPS, 4),
            "traffic": None,
            "ops_per_photon": round(opp, 1),
            "atomics_per_photon": round(depp, 2),
            "atomic_rate_per_s": round(value * 1e3 * depp, 0),
        }
        res["cpu_baseline"] = {
            "value": round(cpu_rate, 2),
            "unit": "photons/ms",
            "cores": ncpu,
            "kind": "port",
            "sample": "%d photons of the same workload (65536-thread map), oracle in C with OpenMP" % cpu_n,
            "absorbed_fraction": round(ref["stat"]["energyabs"] / ref["stat"]["energytot"], 6),
        }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
