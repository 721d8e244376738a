"""Benchmark: forward photon transport on one MI355X.

Workload (one "step"): the heterogeneous 60^3 cube with a radius-15 sphere
(label 2), 1e8 photons, 10 time gates over 0-5 ns, four 1 mm disk detectors with
partial path lengths and exit positions saved, float32 fluence output.

value = photons per millisecond of the whole device job: the transport kernel
plus the fluence finalisation kernel, bracketed by HIP events on the stream they
run on, with the volume, media table and RNG seeds already resident in HBM.
The JSON line also carries:
  roofline      the transport kernel against the FP32/INT32 vector-issue peak
                (algorithmic operations per photon from a frozen per-event
                model times the oracle's event counts, divided by the kernel's
                HIP-event duration); traffic = HBM bytes per launch read from the
                committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present;
                "atomic" = the kernel's float64 atomic rate against the ceiling
                measured live by mcx_device_atomic_ceiling
  cpu_baseline  the C restatement of the reference kernel in its CPU-baseline
                build (-O3, libm float32 math, float32 tallies, OpenMP) timed
                on this host on a bounded sample of the same workload, with
                the host's thread count and CPU model
  adjoint       replay-adjoint gradient of detector signals on the 200^3
                four-layer volume against central finite differences
  inverse       16 sources x 16 detectors on 128^3 per-voxel media: forward,
                folded backward, all 256 signals' gradients, finite-difference check
  phiphi        the phi.phi adjoint outputs: 4 sources and 4 detectors on 128^3,
                run time and the device post-processing's bytes per second
  batch         256 independent 64^3 simulations in one persistent launch,
                against the same instances run one call each

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons N]
       python bench.py --copy-probe BYTES   (plain device copy, for counter calibration)
"""

import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/ms (fwd, frac of roofline); adjoint gradient/s vs finite-diff at ≤2% err"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "transport_sphere60_1e8_hbm_traffic.txt")

# Frozen per-event operation model: VALU lane operations of the reference
# algorithm (src/mcx_core.cl mcx_main_loop), counted from its source; divides,
# square roots and transcendentals are listed separately and weighted x4 because
# they issue at quarter rate.
#   step      one main-loop iteration: property fetch, 3-face voxel DDA,
#             step clipping, move, Beer-Lambert attenuation, time update, voxel
#             index, inside test, medium lookup, termination/roulette tests
#                                                   ~65 ops, 5 div/exp
#   deposit   voxel-exit weight, gate index, address  ~17 ops, 1 div
#   scatter   scattering length (log), azimuth sincos, Henyey-Greenstein
#             inversion, cos/sin of the polar angle, rotation with 3 rsqrt,
#             3 xorshift128+ draws                    ~135 ops, 10 div/sqrt/trans
#   boundary  Fresnel coefficient, RNG draw, reflect/transmit  ~55 ops, 2 div/sqrt
ROOFLINE_MODEL = {
    "step": (65, 5),
    "deposit": (17, 1),
    "scatter": (135, 10),
    "boundary": (55, 2),
}
# MI355X: 256 CUs x 64 FP32 lanes x 2.4 GHz = 39.3 T lane operations/s.  The
# published 157.3 TFLOP/s FP32 vector peak is this issue rate x2 (FMA) x2
# (packed FP32); this kernel's work is compares, selects, integer indexing and
# single (unpacked) FP32 operations, so the issue rate is its roofline.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
FP32_PEAK_TFLOPS = 157.3


def ops_per_photon(counters):
    n = counters["nlaunch"]
    tot = 0.0
    for key, cnt in (("step", counters["nstep"]), ("deposit", counters["ndep"]), ("scatter", counters["nscat"]),
                     ("boundary", counters["nbnd"])):
        ops, trans = ROOFLINE_MODEL[key]
        tot += cnt * (ops + 3 * trans)
    return tot / n, counters["ndep"] / n


def read_traffic():
    """per-launch figures of the committed rocprofv3 counter summary (HBM bytes,
    global atomics per photon)"""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    out = {}
    with open(TRAFFIC_SUMMARY) as f:
        for m in re.finditer(r"^([a-z_]+):\s*([0-9.eE+-]+)\s*$", f.read(), re.M):
            out[m.group(1)] = float(m.group(2))
    return out or None


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(n, 1)


def cpu_model():
    """the host CPU's model name (lscpu's 'Model name', from /proc/cpuinfo)"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_probe(nbytes):
    """one plain device-to-device copy of nbytes (FETCH_SIZE/WRITE_SIZE calibration)"""
    import torch

    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}))


def adjoint_leg(photons, nfd, seed=0, hrel=1.0):
    """Gradient of the 8 detector signals of the 200^3 four-layer workload with
    respect to per-voxel mua, by replay (mcxcl_amd.autograd), against central
    finite differences at nfd voxels drawn uniformly at random from the
    sensitive region (|dS/dmua| >= 1e-4 of the largest; the gradient is exactly
    zero where no detected photon passes).  Runs
    with the fixed photon-per-lane map so that every forward run launches the
    same photons (exact common random numbers: absorption does not move
    photons).  The finite-difference time is measured on those nfd voxels; the
    64-voxel figure is that time per voxel x 64 unless nfd >= 64."""
    import torch
    from mcxcl_amd import autograd, workloads

    os.environ["MCXHIP_DYNAMIC"] = "0"
    try:
        cfg = workloads.layered_workload(nphoton=int(photons), mathmode=0)
        vol = cfg["vol"]
        prop = np.asarray(cfg["prop"], np.float32)
        mua0 = prop[vol, 0].astype(np.float32)
        mus0 = prop[vol, 1].astype(np.float32)
        autograd.detector_signals(torch.tensor(mua0), torch.tensor(mus0), cfg)   # warm-up
        t0 = time.perf_counter()
        mua = torch.tensor(mua0, requires_grad=True)
        S = autograd.detector_signals(mua, torch.tensor(mus0), cfg)
        t1 = time.perf_counter()
        S.sum().backward()
        t2 = time.perf_counter()
        g = mua.grad.numpy()
        visited = np.flatnonzero(g.ravel() != 0)
        # the sensitive region: voxels whose |dS/dmua| is at least 1e-4 of the largest.
        # Below that a voxel is grazed by a few photons and the signal change of a
        # central difference falls under the float32 rounding of their exit weights
        # (and the float64 rounding of S), so no finite difference resolves it.
        sensitive = visited[np.abs(g.ravel()[visited]) >= 1e-4 * np.abs(g).max()]
        rng = np.random.default_rng(seed)
        picks = rng.choice(sensitive, min(nfd, sensitive.size), replace=False)
        errs, tf = [], []
        tfd0 = time.perf_counter()
        for flat in picks:
            v = np.unravel_index(flat, g.shape)
            # a step of hrel x mua: the signal change of a voxel that detected photons
            # only graze is a few float32 ulps of their weights at small steps, while
            # the central difference's own error is (h L)^2 / 6 < 0.2% for L < 5 mm
            h = hrel * mua0[v]
            sig = []
            for s in (1, -1):
                m = mua0.copy()
                m[v] += s * h
                ta = time.perf_counter()
                sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg).detach().sum()))
                tf.append(time.perf_counter() - ta)
            fd = (sig[0] - sig[1]) / (2 * h)
            errs.append(abs(fd - g[v]) / abs(g[v]))
            print("adjoint fd voxel %s: fd %.6g grad %.6g" % (tuple(int(i) for i in v), fd, g[v]), file=sys.stderr,
                  flush=True)
        t_fd = time.perf_counter() - tfd0
        t_adj = t2 - t0
        fd64 = t_fd / len(picks) * 64
        errs = np.array(errs)
        return {
            "workload": "200^3 4-layer volume, %d photons, 8 detectors, gradient of sum of detector signals "
                        "w.r.t. per-voxel mua" % int(photons),
            "forward_s": round(t1 - t0, 4),
            "backward_s": round(t2 - t1, 4),
            "gradient_per_s": round(1.0 / t_adj, 4),
            "voxels_checked": int(len(picks)),
            "voxel_sample": "uniform over the %d voxels with |dS/dmua| >= 1e-4 max (of %d with a nonzero gradient, "
                            "%d in all); central differences with steps of +-%g x mua" %
                            (sensitive.size, visited.size, g.size, hrel),
            "fd_measured_s": round(t_fd, 3),
            "fd_64_voxels_s": round(fd64, 3),
            "fd_64_voxels_measured": bool(len(picks) >= 64),
            "speedup_vs_fd_64_voxels": round(fd64 / t_adj, 1),
            "max_rel_err_vs_fd": round(float(errs.max()), 6),
            "median_rel_err_vs_fd": round(float(np.median(errs)), 7),
            "frac_within_2pct": round(float((errs <= 0.02).mean()), 4),
        }
    finally:
        os.environ.pop("MCXHIP_DYNAMIC", None)


def inverse_leg(photons_per_source, nfd, seed=0):
    """The multi-source differentiable step: 128^3 per-voxel mua/mus (two layers
    with smooth +-30% perturbations), 16 sources x 16 detectors in one launch.
    Times the forward run, the folded backward of a scalar loss over the 256
    signals, and the all-pairs gradients (every signal's dS/dmua and dS/dmus
    from two replays with one slice per source and detector); checks single
    signals against central differences (steps of +-mua) at nfd voxels, each
    drawn uniformly from that signal's sensitive region (|dS/dmua| >= 1e-4 of
    its largest).  Fixed photon-per-lane map, so finite differences use common
    random numbers."""
    import torch
    from mcxcl_amd import autograd, workloads

    os.environ["MCXHIP_DYNAMIC"] = "0"
    try:
        cfg, mua0, mus0 = workloads.inverse_workload(photons_per_source=int(photons_per_source))
        wts = torch.tensor(np.random.default_rng(seed).uniform(0.5, 1.5, (16, 16)))
        autograd.detector_signals(torch.tensor(mua0), torch.tensor(mus0), cfg)   # warm-up
        t0 = time.perf_counter()
        mua = torch.tensor(mua0, requires_grad=True)
        mus = torch.tensor(mus0, requires_grad=True)
        S = autograd.detector_signals(mua, mus, cfg)
        t1 = time.perf_counter()
        (S * wts).sum().backward()
        t2 = time.perf_counter()
        _, dmua, _ = autograd.signal_jacobians(mua0, mus0, cfg)
        t3 = time.perf_counter()
        rng = np.random.default_rng(seed)
        errs, tf = [], []
        for k in range(nfd):
            s_, d_ = int(rng.integers(16)), int(rng.integers(16))
            g = dmua[s_, d_]
            sensitive = np.flatnonzero(np.abs(g.ravel()) >= 1e-4 * np.abs(g).max())
            v = np.unravel_index(sensitive[int(rng.integers(0, sensitive.size))], g.shape)
            h = mua0[v]
            sig = []
            for sgn in (1, -1):
                m = mua0.copy()
                m[v] += sgn * h
                ta = time.perf_counter()
                sig.append(float(autograd.detector_signals(torch.tensor(m), torch.tensor(mus0), cfg)[s_, d_]))
                tf.append(time.perf_counter() - ta)
            fd = (sig[0] - sig[1]) / (2 * h)
            errs.append(abs(fd - g[v]) / abs(g[v]))
        t_fwd = float(np.median(tf))
        return {
            "workload": "128^3 per-voxel mua/mus (2 layers, +-30%% smooth perturbations), 16 sources x 16 "
                        "detectors in one launch, %d photons per source" % int(photons_per_source),
            "signals": int(S.numel()),
            "forward_s": round(t1 - t0, 4),
            "backward_scalar_loss_s": round(t2 - t1, 4),
            "all_pairs_gradients_s": round(t3 - t2, 4),
            "fd_one_voxel_one_signal_s": round(2 * t_fwd, 3),
            "voxels_checked": nfd,
            "max_rel_err_vs_fd": round(float(max(errs)), 5) if errs else None,
            "median_rel_err_vs_fd": round(float(np.median(errs)), 6) if errs else None,
        }
    finally:
        os.environ.pop("MCXHIP_DYNAMIC", None)


def phiphi_leg(photons, nrep=3):
    """The phi.phi adjoint outputs (outputtype adjoint_mua_d): a 128^3
    homogeneous medium (mua 0.005/mm, mus 1/mm, g 0.8, n 1.37), 4 pencil
    sources and 4 detectors (r 2 mm, facing into the medium) on the z=0 face,
    all 8 launched in one run (srcid -1).  Times the run and the device work
    after the transport (jobms - kernelms: the float field of the 8 sources and
    the two adjoint kernels), and states that work's algorithmic bytes: the
    field written once and read by each kernel, the 2 x 16 pair products
    written."""
    import mcxcl_amd

    dim = 128
    pos = [[32.5, 32.5], [95.5, 32.5], [32.5, 95.5], [95.5, 95.5]]
    cfg = {
        "nphoton": int(photons), "seed": 1648335518, "vol": np.ones((dim, dim, dim), np.uint8),
        "prop": [[0, 0, 1, 1], [0.005, 1.0, 0.8, 1.37]], "tstart": 0, "tend": 5e-9, "tstep": 5e-9,
        "srcpos": [[x, y, 0] for x, y in pos], "srcdir": [[0, 0, 1, 0]] * 4,
        "detpos": [[x + 16, y + 16, 0, 2] for x, y in pos], "detdir": [[0, 0, 1, 0]] * 4,
        "outputtype": "adjoint_mua_d", "issavedet": 0,
    }
    mcxcl_amd.run(cfg)
    walls, post, kern = [], [], []
    for _ in range(nrep):
        t0 = time.perf_counter()
        out = mcxcl_amd.run(cfg)
        walls.append(time.perf_counter() - t0)
        kern.append(out["stat"]["kernelms"])
        post.append(out["stat"]["jobms"] - out["stat"]["kernelms"])
    nvox, ns, nd = dim ** 3, 4, 4
    nbytes = nvox * (ns + nd) * (8 + 4) + 2 * nvox * (ns + nd) * 4 + 2 * nvox * ns * nd * 4
    post_ms = float(np.median(post))
    return {
        "workload": "128^3 homogeneous, 4 pencil sources + 4 detectors as disk sources in one launch, "
                    "%d photons, outputtype adjoint_mua_d (jmua and jd for 16 pairs)" % int(photons),
        "run_wall_s": round(float(np.median(walls)), 4),
        "transport_ms": round(float(np.median(kern)), 3),
        "post_ms": round(post_ms, 3),
        "post_algorithmic_bytes": nbytes,
        "post_GBps": round(nbytes / (post_ms * 1e-3) / 1e9, 1) if post_ms > 0 else None,
        "outputs": sorted(k for k in out if k.startswith("j")),
    }


def features_leg(photons):
    """Fast-build throughput of the full general path's features on the 60^3
    sphere-in-cube geometry of the headline workload (same properties, 10
    gates, 4 detectors, reflection): the plain label run for reference, packed
    labelplus words, the ball as split voxels (svmc), RF forward (150 MHz), two
    patterns sharing photons, and a polarised cube (0.3 um spheres in water,
    632.8 nm) with the same pencil beam.  kernelms of one run each, after one
    warm-up run."""
    import mcxcl_amd
    from mcxcl_amd import workloads

    n = int(photons)
    base = workloads.sphere_cube_workload(nphoton=n, tstep=5e-10)
    vol = np.asarray(base["vol"])
    prop = np.asarray(base["prop"], np.float32)
    lplus = np.stack([prop[vol, 0], np.zeros(vol.shape, np.float32), vol.astype(np.float32)])
    q = np.moveaxis(np.indices(vol.shape), 0, -1) + 0.5
    d = np.linalg.norm(q - 30.0, axis=-1)
    sv = np.zeros((8,) + vol.shape, np.uint8)
    sv[0] = np.where(d < 15, 2, 1)
    near = np.abs(d - 15) < 0.9
    nhat = (q - 30.0) / np.maximum(d, 1e-9)[..., None]
    off = 30.0 + 15 * nhat - np.moveaxis(np.indices(vol.shape), 0, -1)
    ok = near & np.all((off >= 0) & (off <= 1), axis=-1)
    sv[0][ok], sv[1][ok] = 2, 1                               # lower: inside the ball, upper: outside
    pb = np.clip(np.rint(off * 255), 0, 255).astype(np.uint8)
    nb = np.clip(np.rint((nhat + 1) * 127.5), 0, 255).astype(np.uint8)
    for k in range(3):
        sv[2 + k][ok], sv[5 + k][ok] = pb[..., k][ok], nb[..., k][ok]
    pats = np.random.default_rng(1).uniform(0.2, 1.0, (2, 60, 60)).astype(np.float32)
    cases = {
        "label": {},
        "labelplus": {"vol": lplus, "savedetflag": "DX"},
        "svmc": {"vol": sv},
        "rf_forward": {"omega": 2 * np.pi * 150e6},
        "shared_patterns": {"srctype": "pattern", "srcpos": [0, 0, 0], "srcparam1": [60, 0, 0, 60],
                            "srcparam2": [0, 60, 0, 60], "srcnum": 2, "srcpattern": pats},
        "polarised": {"vol": np.ones(vol.shape, np.uint8), "prop": [[0, 0, 1, 1], [0.005, 0, 0, 1.33]],
                      "polprop": [[0.005, 0.3, 0.002, 1.59, 1.33]], "lambda": 632.8, "srciquv": [1, 1, 0, 0]},
    }
    out = {}
    for name, extra in cases.items():
        cfg = dict(base, **extra)
        mcxcl_amd.run(dict(cfg, nphoton=max(n // 10, 1)))
        r = mcxcl_amd.run(cfg)
        ms = float(r["stat"]["kernelms"])
        out[name] = {"kernel_ms": round(ms, 3), "photons_per_ms": round(n / ms, 1) if ms > 0 else None}
    out["workload"] = ("60^3 sphere-in-cube (r 15, 10 gates, 4 detectors, reflection), %d photons per case, "
                       "fast build; polarised: a 60^3 Mie suspension, same beam" % n)
    return out


def batch_leg(ninst, steps, count_photons=200):
    """Many independent simulations in one persistent launch: ninst 64^3 volumes
    with a random ellipsoid each, per-instance media, log-uniform photon counts
    in [1e4, 1e6], pencil/isotropic/disk sources (mcxcl_amd.workloads.
    batched_workload).  A step simulates every instance; the schedule is
    tuned once (launch shape, per-XCD lists balanced by probed costs) and each
    step replays the captured HIP graph.  The same
    instances run one mcx_run_simulation call each give the sequential rate."""
    from mcxcl_amd import pmcxcl, workloads

    cfgs = workloads.batched_workload(ninst=ninst)
    total = int(sum(c["nphoton"] for c in cfgs))
    t0 = time.perf_counter()
    with pmcxcl.Batch(cfgs) as b:
        setup_s = time.perf_counter() - t0
        b.run(1)
        step_ms = []
        for _ in range(steps):
            step_ms.append(b.run(1)["stepms"])
            print("batch step: %.2f ms" % step_ms[-1], file=sys.stderr, flush=True)
        outs = b.collect()
        info = b.info
    launched_all = all(o["stat"]["energytot"] == c["nphoton"] for o, c in zip(outs, cfgs))
    seq_job = 0.0
    t1 = time.perf_counter()
    for c in cfgs:
        seq_job += pmcxcl.run(c)["stat"]["jobms"]
    seq_wall = (time.perf_counter() - t1) * 1e3
    ms = float(np.mean(step_ms))
    roof = None
    if count_photons > 0:
        # algorithmic work of the step: every instance's event counts per photon from
        # the checker build of the oracle on a sample of its own photons, times its
        # photon count, weighted by the frozen operation model
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle

        ops = 0.0
        for c in cfgs:
            sample = dict(c, nphoton=int(count_photons), autopilot=0, nthread=1024, nblocksize=64)
            cnt = oracle.run(sample, nthread=1024, nomp=host_threads())["counters"]
            ops += c["nphoton"] * ops_per_photon(cnt)[0]
        achieved = ops / (ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 3),
                "unit": "Tops/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                "ops_per_photon": round(ops / total, 1), "kernel": "batch_kernel (step time: prepare + "
                "transport + finalize of one graph replay)",
                "count_sample": "%d photons of each instance through the checker build" % int(count_photons)}
    return {
        "workload": "%d independent 64^3 simulations, random ellipsoid each, per-instance mua/mus/g, "
                    "log-uniform 1e4-1e6 photons, pencil/isotropic/disk sources, 1 gate" % ninst,
        "roofline": roof,
        "photons_per_step": total,
        "ms_per_step": round(ms, 3),
        "photons_per_ms": round(total / ms, 1),
        "every_instance_launched_all_photons": bool(launched_all),
        "schedule": {k: info[k] for k in ("blocksize", "nblock", "nqueue", "probe", "usegraph")},
        "setup_and_autotune_s": round(setup_s, 2),
        "autotune_ms": round(info["tunems"], 1),
        "sequential_device_ms": round(seq_job, 1),
        "sequential_wall_ms": round(seq_wall, 1),
        "speedup_vs_sequential_device": round(seq_job / ms, 2),
        "speedup_vs_sequential_wall": round(seq_wall / ms, 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--mathmode", type=int, default=0, help="0: hardware transcendentals, 1: bit-reproducible")
    ap.add_argument("--cpu-photons", type=float, default=3e6, help="photons in the CPU-baseline sample")
    ap.add_argument("--count-photons", type=float, default=1e6,
                    help="photons of the oracle run that counts events for the operation model")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, help="run one plain device copy of this many bytes and exit")
    ap.add_argument("--adjoint-photons", type=float, default=1e8,
                    help="photons of the replay-adjoint gradient leg (0 skips it)")
    ap.add_argument("--adjoint-fd-voxels", type=int, default=4,
                    help="voxels checked by central differences (64 measures the whole 64-voxel sweep)")
    ap.add_argument("--adjoint-fd-step", type=float, default=1.0,
                    help="central-difference step of the adjoint leg as a multiple of the voxel's mua")
    ap.add_argument("--inverse-photons", type=float, default=1e7,
                    help="photons per source of the 16x16 multi-source gradient leg (0 skips it)")
    ap.add_argument("--inverse-fd-voxels", type=int, default=2)
    ap.add_argument("--phiphi-photons", type=float, default=1e7,
                    help="photons of the phi.phi adjoint-output leg (0 skips it)")
    ap.add_argument("--batch-instances", type=int, default=256,
                    help="simulations of the batched-launch leg (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--feature-photons", type=float, default=1e7,
                    help="photons per case of the full-path feature throughput leg (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on several GPUs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import mcxcl_amd
    from mcxcl_amd import workloads

    nphoton = int(args.photons)
    cfg = workloads.sphere_cube_workload(nphoton=nphoton, mathmode=args.mathmode)

    for _ in range(args.warmup):
        mcxcl_amd.run(cfg)

    job_ms, kernel_ms, wall, absorbed, ndet = [], [], [], [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        out = mcxcl_amd.run(cfg)
        wall.append((time.perf_counter() - t0) * 1e3)
        st = out["stat"]
        job_ms.append(st["jobms"])
        kernel_ms.append(st["kernelms"])
        absorbed.append(st["energyabs"] / st["energytot"])
        ndet.append(out["detp"].shape[1] if "detp" in out else 0)
        print("step: job %.2f ms (transport %.2f ms), wall %.1f ms" % (job_ms[-1], kernel_ms[-1], wall[-1]),
              file=sys.stderr, flush=True)

    ms_per_step = float(np.mean(job_ms))
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
            "parallelism": "single GPU; fast math: one resident wave of workgroups drawing photons at run time",
        },
        "transport_ms_per_step": round(float(np.mean(kernel_ms)), 3),
        "wall_ms_per_step": round(float(np.mean(wall)), 3),
        "absorbed_fraction": round(float(np.mean(absorbed)), 6),
        "detected_per_step": int(np.mean(ndet)),
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle

        ncpu = host_threads()
        # CPU baseline: the -O3 float32 build of the C restatement, OpenMP on the
        # host threads, timed on a bounded sample of the same workload
        cpu_n = int(args.cpu_photons)
        ccfg = workloads.sphere_cube_workload(nphoton=cpu_n, autopilot=0, nthread=65536, nblocksize=64)
        base = oracle.run(ccfg, nthread=65536, nomp=ncpu, fast=True)
        cpu_rate = cpu_n / (base["seconds"] * 1e3)
        # event counts per photon for the operation model: the checker build counts them
        ncount = int(args.count_photons)
        ref = oracle.run(workloads.sphere_cube_workload(nphoton=ncount, autopilot=0, nthread=65536, nblocksize=64),
                         nthread=65536, nomp=ncpu)
        opp, depp = ops_per_photon(ref["counters"])
        transport_ms = float(np.mean(kernel_ms))
        achieved = nphoton * opp / (transport_ms * 1e-3) / 1e12
        traffic = read_traffic()
        res["roofline"] = {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS, 3),
            "unit": "Tops/s",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
            "kernel": "transport_kernel<savedet,reflect,simple,full>",
            "kernel_ms": round(transport_ms, 3),
            "ops_per_photon": round(opp, 1),
            "frac_of_fp32_vector_peak": round(achieved / FP32_PEAK_TFLOPS, 4),
            "fluence_deposits_per_photon": round(depp, 2),
        }
        # second ceiling: float64 atomics at scattered addresses, measured live
        ceiling = ctypes.c_double(0.0)
        if mcxcl_amd._native.lib.mcx_device_atomic_ceiling(0, 60 * 60 * 60 * 10, ctypes.byref(ceiling)) == 0 and \
                traffic and traffic.get("global_atomics_per_photon"):
            rate = nphoton * traffic["global_atomics_per_photon"] / (transport_ms * 1e-3)
            res["roofline"]["atomic"] = {
                "achieved": round(rate, 0),
                "peak": round(ceiling.value, 0),
                "unit": "atomics/s",
                "frac": round(rate / ceiling.value, 4),
                "atomics_per_photon": traffic["global_atomics_per_photon"],
            }
        res["cpu_baseline"] = {
            "value": round(cpu_rate, 2),
            "unit": "photons/ms",
            "cores": ncpu,
            "kind": "port",
            "sample": "%d photons of the same workload (65536-thread map): the C restatement of mcx_core.cl "
                      "semantics built -O3 -march=x86-64-v3 with libm float32 math and float32 tallies, OpenMP "
                      "on %d threads, %.1f s" % (cpu_n, ncpu, base["seconds"]),
            "absorbed_fraction": round(base["stat"]["energyabs"] / base["stat"]["energytot"], 6),
            "nproc": os.cpu_count(),
            "cpu_model": cpu_model(),
        }
    if args.adjoint_photons > 0:
        res["adjoint"] = adjoint_leg(args.adjoint_photons, args.adjoint_fd_voxels, hrel=args.adjoint_fd_step)
    if args.inverse_photons > 0:
        res["inverse"] = inverse_leg(args.inverse_photons, args.inverse_fd_voxels)

    if args.phiphi_photons > 0:
        res["phiphi"] = phiphi_leg(args.phiphi_photons)
    if args.batch_instances > 0:
        res["batch"] = batch_leg(args.batch_instances, args.batch_steps, 0 if args.no_cpu else 200)
    if args.feature_photons > 0:
        res["features"] = features_leg(args.feature_photons)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
